"""Summarise rocprofv3 CSV output (kernel stats and PMC counter collection).

    python -m pardiso2_amd.rocprof_summary stats <dir> [--top N]
    python -m pardiso2_amd.rocprof_summary trace <dir> --kernel SUBSTR [--last N]
    python -m pardiso2_amd.rocprof_summary pmc <dir> [--kernel SUBSTR] [--last N]
    python -m pardiso2_amd.rocprof_summary window <dir> --since SUBSTR [--nth K] [--top N] [--by-grid]

`stats` reads the *_kernel_stats.csv that `rocprofv3 --kernel-trace --stats
--output-format csv` writes; `trace` reads the same run's *_kernel_trace.csv
and prints the duration statistics of the matching dispatches (the last N of
them with --last); `pmc` reads *_counter_collection.csv of a `rocprofv3 --pmc
... --output-format csv` run and prints, per kernel name and counter, the
number of dispatches and the mean value per dispatch (of the last N);
`window` takes the kernel trace from the K-th last dispatch whose name holds
SUBSTR to the end (e.g. one timed solve step) and prints per-kernel totals,
the window's span and the share of it the device spent in kernels (with
--by-grid, one line per kernel and launch shape, so the levels of a hierarchy
show apart).
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sys
from collections import defaultdict


def _find(d, suffix):
    hits = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not hits:
        raise SystemExit(f"no *{suffix} under {d}")
    return hits


def _col(row, *names):
    for n in names:
        if n in row and row[n] != "":
            return row[n]
    raise KeyError(f"none of {names} in {list(row)}")


def stats(d, top):
    rows = []
    for f in _find(d, "kernel_stats.csv"):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    tot = sum(float(_col(r, "TotalDurationNs")) for r in rows)
    rows.sort(key=lambda r: -float(_col(r, "TotalDurationNs")))
    print(f"{'calls':>8} {'total ms':>10} {'avg us':>10} {'min us':>9} {'max us':>9} {'%':>6}  kernel")
    for r in rows[:top]:
        t = float(_col(r, "TotalDurationNs"))
        print(f"{int(_col(r, 'Calls')):8d} {t / 1e6:10.3f} {float(_col(r, 'AverageNs')) / 1e3:10.2f} "
              f"{float(_col(r, 'MinNs')) / 1e3:9.2f} {float(_col(r, 'MaxNs')) / 1e3:9.2f} {100 * t / tot:6.2f}  "
              f"{_col(r, 'Name', 'KernelName')[:110]}")
    print(f"total kernel time {tot / 1e6:.3f} ms over {sum(int(_col(r, 'Calls')) for r in rows)} dispatches")


def trace(d, kernel, last):
    rows = []
    for f in _find(d, "kernel_trace.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                name = _col(r, "Kernel_Name", "KernelName")
                if kernel in name:
                    t0, t1 = float(_col(r, "Start_Timestamp")), float(_col(r, "End_Timestamp"))
                    rows.append((int(_col(r, "Dispatch_Id", "Correlation_Id")), name, t1 - t0))
    rows.sort()
    if last:
        rows = rows[-last:]
    if not rows:
        raise SystemExit(f"no dispatch of {kernel}")
    ds = sorted(r[2] for r in rows)
    print(f"kernel {rows[-1][1][:110]}")
    print(f"dispatches {len(ds)}  avg us {sum(ds) / len(ds) / 1e3:.3f}  min us {ds[0] / 1e3:.3f}  "
          f"median us {ds[len(ds) // 2] / 1e3:.3f}  max us {ds[-1] / 1e3:.3f}")


def window(d, since, nth, top, by_grid=False):
    rows = []
    for f in _find(d, "kernel_trace.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                name = _col(r, "Kernel_Name", "KernelName")
                if by_grid:  # one line per launch shape: grid x workgroup size
                    name = f"[grid {_col(r, 'Grid_Size_X', 'Grid_Size')}/{_col(r, 'Workgroup_Size_X', 'Workgroup_Size')}] {name}"
                rows.append((float(_col(r, "Start_Timestamp")), float(_col(r, "End_Timestamp")), name))
    rows.sort()
    marks = [i for i, r in enumerate(rows) if since in r[2]]
    if len(marks) < nth:
        raise SystemExit(f"fewer than {nth} dispatches of {since}")
    rows = rows[marks[-nth]:]
    span = rows[-1][1] - rows[0][0]
    per = defaultdict(lambda: [0, 0.0])
    for t0, t1, name in rows:
        per[name][0] += 1
        per[name][1] += t1 - t0
    busy = sum(v[1] for v in per.values())
    print(f"window: {len(rows)} dispatches, span {span / 1e6:.3f} ms, in kernels {busy / 1e6:.3f} ms "
          f"({100 * busy / span:.1f} %)")
    print(f"{'calls':>8} {'total ms':>10} {'avg us':>10} {'% span':>7}  kernel")
    for name, (cnt, t) in sorted(per.items(), key=lambda kv: -kv[1][1])[:top]:
        print(f"{cnt:8d} {t / 1e6:10.3f} {t / cnt / 1e3:10.2f} {100 * t / span:7.2f}  {name[:110]}")


def pmc(d, kernel, last=0):
    per = defaultdict(float)  # (kernel, counter, dispatch) -> value summed over the row's dimensions
    for f in _find(d, "counter_collection.csv"):
        with open(f) as fh:
            for i, r in enumerate(csv.DictReader(fh)):
                name = _col(r, "Kernel_Name", "KernelName", "Kernel-Name")
                if kernel and kernel not in name:
                    continue
                try:
                    disp = _col(r, "Dispatch_Id", "DispatchId", "Correlation_Id")
                except KeyError:
                    disp = i
                cn = _col(r, "Counter_Name", "CounterName")
                per[(name, cn, disp)] += float(_col(r, "Counter_Value", "CounterValue"))
    acc = defaultdict(list)
    for (name, cn, disp), v in sorted(per.items(), key=lambda kv: (kv[0][0], kv[0][1], int(kv[0][2]))):
        acc[(name, cn)].append(v)
    if last:
        acc = {key: v[-last:] for key, v in acc.items()}
    for (name, cn), v in sorted(acc.items()):
        print(f"{cn:>12} dispatches {len(v):5d} mean {sum(v) / len(v):14.1f}  {name[:100]}")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "trace", "pmc", "window"])
    ap.add_argument("dir")
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--kernel", default=None)
    ap.add_argument("--last", type=int, default=0)
    ap.add_argument("--since", default=None)
    ap.add_argument("--nth", type=int, default=1)
    ap.add_argument("--by-grid", action="store_true", help="window: one line per kernel and launch shape")
    a = ap.parse_args(argv)
    if a.mode == "stats":
        stats(a.dir, a.top)
    elif a.mode == "trace":
        trace(a.dir, a.kernel or "", a.last)
    elif a.mode == "window":
        window(a.dir, a.since or "", a.nth, a.top, a.by_grid)
    else:
        pmc(a.dir, a.kernel, a.last)


if __name__ == "__main__":
    sys.exit(main())
