"""Summarise rocprofv3 CSV output (kernel stats and PMC counter collection).

    python -m pardiso2_amd.rocprof_summary stats <dir> [--top N]
    python -m pardiso2_amd.rocprof_summary trace <dir> --kernel SUBSTR [--last N]
    python -m pardiso2_amd.rocprof_summary pmc <dir> [--kernel SUBSTR] [--last N]
    python -m pardiso2_amd.rocprof_summary window <dir> --since SUBSTR [--nth K] [--top N] [--by-grid]
    python -m pardiso2_amd.rocprof_summary traffic <dir> --nodes N --entry NAME:SUBSTR:BYTES_PER_NODE ... [--out F]
        [--copy-per-node K]

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
show apart).  `traffic` reads <dir>/<NAME>_FETCH_SIZE and <dir>/<NAME>_WRITE_SIZE
(one `--pmc` run each, NAME from the entries and `copy`), scales both counters
by what they read for k_stream_copy (K*N doubles in, K*N out per launch, K = 4
by default) and
prints, per entry, the HBM bytes per launch of the kernel whose name holds
SUBSTR (mean of its last --last dispatches, default 10) against its
algorithmic bytes (BYTES_PER_NODE * N); --out writes the same as JSON.
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


def window(d, since, nth, top, by_grid=False, through=None):
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
    if through:  # up to and including the first later dispatch of this kernel
        ends = [i for i, r in enumerate(rows) if i > 0 and through in r[2]]
        if ends:
            rows = rows[:ends[0] + 1]
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


def _pmc_means(d, kernel, last=0):
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
    return acc


def pmc(d, kernel, last=0):
    for (name, cn), v in sorted(_pmc_means(d, kernel, last).items()):
        print(f"{cn:>12} dispatches {len(v):5d} mean {sum(v) / len(v):14.1f}  {name[:100]}")


def _counter(d, substr, counter, last):
    acc = {k: v for k, v in _pmc_means(d, substr, last).items() if k[1] == counter}
    if len(acc) != 1:
        raise SystemExit(f"{d}: {len(acc)} kernels named *{substr}* with {counter}: {sorted(k[0] for k in acc)}")
    (name, _), v = next(iter(acc.items()))
    return name, sum(v) / len(v)


def traffic(d, nodes, entries, last, out, copy_per_node=4):
    copy_bytes = copy_per_node * nodes * 8.0
    _, cf = _counter(os.path.join(d, "copy_FETCH_SIZE"), "k_stream_copy", "FETCH_SIZE", last)
    _, cw = _counter(os.path.join(d, "copy_WRITE_SIZE"), "k_stream_copy", "WRITE_SIZE", last)
    ff, wf = copy_bytes / (cf * 1024), copy_bytes / (cw * 1024)
    res = {"nodes": nodes, "fetch_factor": ff, "write_factor": wf, "kernels": {}}
    print(f"calibration: k_stream_copy moves {copy_bytes / 1024:.1f} KiB each way; FETCH_SIZE reads {cf:.1f} KiB "
          f"(factor {ff:.5f}), WRITE_SIZE reads {cw:.1f} KiB (factor {wf:.5f})")
    print(f"{'entry':10} {'FETCH KiB':>12} {'WRITE KiB':>12} {'HBM bytes':>14} {'algorithmic':>14} {'ratio':>6}  instance")
    for e in entries:
        key, sub, bpn = e.rsplit(":", 2)
        name, fk = _counter(os.path.join(d, f"{key}_FETCH_SIZE"), sub, "FETCH_SIZE", last)
        _, wk = _counter(os.path.join(d, f"{key}_WRITE_SIZE"), sub, "WRITE_SIZE", last)
        hbm = fk * 1024 * ff + wk * 1024 * wf
        alg = float(bpn) * nodes
        res["kernels"][key] = {"instance": name, "fetch_kib": fk, "write_kib": wk, "hbm_bytes_per_launch": hbm,
                               "algorithmic_bytes": alg}
        print(f"{key:10} {fk:12.1f} {wk:12.1f} {hbm:14.0f} {alg:14.0f} {hbm / alg:6.3f}  {name[:90]}")
    if out:
        import json
        with open(out, "w") as f:
            json.dump(res, f, indent=1)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "trace", "pmc", "window", "traffic"])
    ap.add_argument("dir")
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--kernel", default=None)
    ap.add_argument("--last", type=int, default=0)
    ap.add_argument("--since", default=None)
    ap.add_argument("--through", default=None, help="window: end with the first later dispatch of this kernel")
    ap.add_argument("--nth", type=int, default=1)
    ap.add_argument("--by-grid", action="store_true", help="window: one line per kernel and launch shape")
    ap.add_argument("--nodes", type=int, default=0, help="traffic: finest-level nodes of the run")
    ap.add_argument("--entry", action="append", default=[], help="traffic: NAME:KERNEL_SUBSTR:BYTES_PER_NODE")
    ap.add_argument("--out", default=None, help="traffic: JSON output file")
    ap.add_argument("--copy-per-node", type=int, default=4,
                    help="traffic: doubles per node the calibrating k_stream_copy moves each way")
    a = ap.parse_args(argv)
    if a.mode == "stats":
        stats(a.dir, a.top)
    elif a.mode == "trace":
        trace(a.dir, a.kernel or "", a.last)
    elif a.mode == "window":
        window(a.dir, a.since or "", a.nth, a.top, a.by_grid, a.through)
    elif a.mode == "traffic":
        traffic(a.dir, a.nodes, a.entry, a.last or 10, a.out, a.copy_per_node)
    else:
        pmc(a.dir, a.kernel, a.last)


if __name__ == "__main__":
    sys.exit(main())
