#!/usr/bin/env python
"""Stokes solves/s and p99 latency to a 1e-8 FP64 relative residual on one MI355X.

Workload (default): single 257^3-node solve, 7 multigrid levels
(257/131/68/36/20/12/8), eight seeded spheres of radius 0.03-0.1 box,
viscosity 1e-3..1e3 times the 1e21 Pa s matrix (log-uniform), density 3200
in 3300 kg/m^3, 2x2x2 jittered markers per cell, free-slip box, gravity -z.  Node fields are built on the device from 2x2x2 markers per cell
once; a step is one solve from those resident node fields: prepare
(viscosity restriction to all levels, stencil viscosities, right-hand side,
coarsest inverse) + hydrostatic guess + FP64 flexible GMRES with the FP32
V-cycle / BFBt block preconditioner until ||W r||/||W b|| <= 1e-8.
--workload randfield --n 129 and --workload sinker --n 33 select the other
single-solve problems; --workload batched and --workload served the
64-instance batch and the Poisson request stream.

The line also carries one implicit temperature step (titerate, parallel/heat3mg.c) of the
257^3 heat box (random conductivity, fixed temperatures on the z faces, 6 levels, V(3,3),
to heatmin 1e-4): its time, the roofline of its colour pass, and the oracle's step at 65^3 on
the host cores beside the device's on that box.

The line also carries the roofline of the finest-level FP32 smoother
instance the timed solve runs (64 algorithmic bytes per node and sweep,
HIP events on the library stream, HBM bytes from the committed rocprofv3
counters of that instance), the oracle's residual of the device solution,
and the CPU baseline: full FP64 solves of the 33^3 instance of the same
workload by the oracle on all host cores (median over repeats, CPU model and
sockets recorded) with the device's rate on that same problem beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
SMOOTHER_BYTES_PER_NODE = 64.0   # FP32 sweep: x(16) + b(16) + 4 viscosities(16) read, x(16) written
RESIDUAL_BYTES_PER_NODE = 112.0  # FP64 residual: x(32) + b(32) + 4 viscosities(16) read, r(32) written


KERNELS = {"smooth32": 0, "smooth64": 1, "apply64": 2, "res32": 3, "res64": 4, "copy": 5}
HEAT_KERNELS = {"heatgs": 0, "heatres": 2, "heatcopy": 3}
# one colour pass of the temperature sweep (k_heat_gs over the finest level): x (8) and kt (4) of every
# node read, b (8), rho cp (4) and the boundary flag (1) of the colour's half read, x (8) of that half written
HEAT_GS_BYTES_PER_NODE = 12.0 + 21.0 / 2


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_problem(workload: str, n: int, seed: int):
    from pardiso2_amd import setups
    L = setups.coarsest_levels(n)
    if workload == "randfield":
        pb = setups.random_field(n, L, seed=seed)
    elif workload == "sinker":
        pb = setups.sinker(n, L, seed=seed)
    elif workload == "spheres":
        pb = setups.random_spheres(n, L, seed=seed)
    else:
        raise SystemExit(f"unknown workload {workload}")
    pb.bc = setups.free_slip_lines(n, n, n, gauge=None)
    return pb


def sweeps_for(L: int):
    pre = [3, 6, 8, 8, 12, 12, 12, 12, 12, 12, 12][:L] + [0]
    return pre, pre


def host_cpu():
    """CPU model, socket and physical core count from /proc/cpuinfo, and the
    CPUs this process may run on."""
    model, sockets, cores, phys = None, set(), set(), None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                    sockets.add(v)
                elif k == "core id":
                    cores.add((phys, v))
    except OSError:
        pass
    return {"cpu_model": model, "sockets": len(sockets) or None, "physical_cores": len(cores) or None,
            "cpus_allowed": len(os.sched_getaffinity(0)), "nproc": os.cpu_count()}


CPU_SAMPLE_N = 33
HOST = host_cpu()  # before any OpenMP runtime binds this process's threads


def load_oracle():
    """The CPU oracle, OpenMP threads bound to the host cores (set before its
    OpenMP runtime starts)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    O.build()
    return O


def cpu_baseline(pb, budget_s: float, repeats: int = 5, device=True):
    """Measured full FP64 solves on the host cores.

    The oracle's FP64 GCR (velocity V-cycle of the reference's smoother and
    transfers, mass-matrix Schur scaling; all OpenMP threads) solves ``pb`` to
    the same weighted relative residual 1e-8 from the same hydrostatic guess,
    ``repeats`` times (at least 3; fewer than ``repeats`` only when the budget
    runs out); median and p99 wall clock.  With ``device`` the library solves
    the same problem the same number of times for a like-for-like rate, and
    its mixed-precision solution is compared with the oracle's converged one.
    """
    host = HOST
    O = load_oracle()
    pre, post = sweeps_for(pb.levels)
    t0 = time.perf_counter()
    o = O.from_problem(pb)
    setup_s = time.perf_counter() - t0
    times, its, x_or = [], [], None
    while len(times) < repeats and (len(times) < 3 or sum(times) < budget_s):
        ts = time.perf_counter()
        it, x, hist, _ = o.gcr(pre, post, mode=1, tol=1e-8, maxit=2000, restart=32)
        times.append(time.perf_counter() - ts)
        its.append(int(it))
        x_or = x
        if it < 0 or hist[-1] > 1e-8:
            raise RuntimeError(f"the oracle's GCR (mass-matrix Schur scaling) stalls on this {pb.n}^3 problem: "
                               f"{hist[-1]:.2e} after {it} iterations, no full CPU solve to 1e-8")
    med = float(np.median(times))
    out = {"value": 1.0 / med, "unit": "solves/s", "cores": O.threads(), "kind": "port",
           "sample": (f"full FP64 solves of the {pb.n}^3 instance of this workload (same generator and seed) by "
                      f"the oracle's GCR + velocity V-cycle (reference smoother and transfers, mass-matrix Schur "
                      f"scaling) to weighted relative residual 1e-8, {len(times)} repeats, median wall clock; "
                      f"oracle setup (assembly, coarse LU) {setup_s:.2f} s not counted; the 129^3 and 257^3 "
                      f"instances do not fit a bounded CPU run"),
           "median_s": med, "p99_s": float(np.percentile(times, 99)), "repeats": len(times), "iterations": its,
           "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES"),
           **host}
    if device:
        from pardiso2_amd import _lib as S
        ctx = S.from_problem(pb)
        ctx.autotune(0)
        opts = S.default_opts(pb.levels, precision=0, pre=pre, post=post, tol=1e-8, maxit=800, restart=128)
        dts, dits = [], []
        for _ in range(len(times) + 1):  # the first solve warms the context
            ts = time.perf_counter()
            ctx.prepare()
            ctx.hydrostatic_guess()
            st = ctx.solve(opts)
            dts.append(time.perf_counter() - ts)
            dits.append(st.iterations)
        out["device_same_sample"] = {"value": 1.0 / float(np.median(dts[1:])), "unit": "solves/s",
                                     "median_ms": float(np.median(dts[1:])) * 1e3, "iterations": dits[-1],
                                     "precision": "mixed"}
        out["device_vs_cpu"] = out["device_same_sample"]["value"] / out["value"]
        out["device_error_vs_oracle"] = solution_diff(ctx.solution(), x_or, ctx.info[0][0], ctx.info[0][1])
        ctx.close()
    return out


def banded_baseline(pb, repeats: int = 3):
    """The reference's secondary CPU path: the banded direct solve gausmat3
    (parallel/gausjomp.c:70-254, rows pipelined over all host threads as
    there), restated by the oracle, on ``pb``; median of ``repeats`` full
    solves, the oracle's relative residual of the result."""
    O = load_oracle()
    o = O.from_problem(pb)
    times, band, rc, x = [], 0, 0, None
    for _ in range(repeats):
        ts = time.perf_counter()
        rc, x, band = o.band_solve(threads=0)
        times.append(time.perf_counter() - ts)
        if rc:
            raise RuntimeError(f"banded solve failed ({rc})")
    rel = float(o.relres_rows(x)[0])
    med = float(np.median(times))
    return {"value": 1.0 / med, "unit": "solves/s", "cores": O.threads(), "kind": "port",
            "sample": (f"full banded direct FP64 solves (gausmat3, band {band}) of the {pb.n}^3 instance of this "
                       f"workload, {len(times)} repeats, median wall clock, assembly not counted"),
            "median_s": med, "n": pb.n, "band": int(band), "unknowns": 4 * pb.n ** 3, "relres": rel}


def solution_diff(x, ref, X, Y):
    """Max and RMS difference of solution x from ref, velocities and
    pressures separately, relative to ref's max and RMS; pressures compared
    at a common gauge (cell 2,2,2)."""
    def gauged(v):
        v = np.array(v, dtype=np.float64)
        p = v[0::4]
        v[0::4] = p - p[2 * X * Y + 2 * Y + 2]
        return v
    a, b = gauged(x), gauged(ref)
    out = {}
    for name, sl in (("velocity", np.arange(a.size) % 4 != 0), ("pressure", np.arange(a.size) % 4 == 0)):
        d, r = a[sl] - b[sl], b[sl]
        out[name] = {"max_rel": float(np.abs(d).max() / max(np.abs(r).max(), 1e-300)),
                     "rms_rel": float(np.sqrt(np.mean(d * d)) / max(np.sqrt(np.mean(r * r)), 1e-300))}
    return out


def solution_error(ctx, x):
    """Difference of solution x from the context's current (FP64) solution."""
    ctx.pressure_gauge(2, 2, 2, 0.0)
    return solution_diff(x, ctx.solution(), ctx.info[0][0], ctx.info[0][1])


def smoother_instance(ctx, tile):
    """Template instance name of the FP32 tile smoother the V-cycle runs on
    the finest level with this tile (the name rocprofv3 reports)."""
    ty, tx, tz, sv = tile
    ch = "true" if ctx.tile_chains(0, tile) else "false"
    pairs = (ty + 1) // 2 * tx * tz
    if not sv and pairs <= 512:  # one node pair per thread: the prefetching kernel
        shape = {(32, 4, 4): "32, 4, 4", (32, 4, 8): "32, 4, 8", (32, 8, 4): "32, 8, 4"}.get((ty, tx, tz), "0, 0, 0")
        # uniform spacing as the library tests it: one FP32 inverse spacing per axis
        inv = [(1.0 / np.diff(a)).astype(np.float32) for a in ctx._g]
        uni = "true" if all(np.all(v == v[0]) for v in inv) else "false"
        return f"k_vsmooth32<{shape}, {ch}, {uni}, {256 if pairs <= 256 else 512}>"
    if sv:
        return f"k_smooth<float, true, 0, 0, 0, {ch}>"
    if (ty, tx, tz) == (32, 4, 4):
        return f"k_smooth<float, false, 32, 4, 4, {ch}>"
    return f"k_smooth<float, false, 0, 0, 0, {ch}>"


def smoother_roofline(ctx, nodes):
    ms = ctx.time_kernel(0, 0, reps=50)
    gbs = SMOOTHER_BYTES_PER_NODE * nodes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": f"k_smooth<float> finest level of a {round(nodes ** (1 / 3))}^3 instance",
            "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS, "traffic": None,
            "ms_per_launch": ms, "bytes_per_launch": SMOOTHER_BYTES_PER_NODE * nodes}


def cpu_sample(pb, budget_s, device=True):
    try:
        return cpu_baseline(pb, budget_s, device=device)
    except Exception as e:  # the baseline must not hide the device result
        return {"value": None, "unit": "solves/s", "cores": None, "kind": "port", "sample": f"failed: {e}", **HOST}


def heat_levels(n: int) -> int:
    """Coarse levels of the temperature V-cycle: the hierarchy down to 12 lines per axis (the
    reference's cycle diverges on these boxes with its 8-line level below the 12-line one)."""
    from pardiso2_amd import setups
    L = setups.coarsest_levels(n)
    return L - 1 if L >= 2 else L


def heat_setup(n: int):
    from pardiso2_amd import _lib as S, setups
    Lt = heat_levels(n)
    hp = setups.heat_box(n, setups.coarsest_levels(n))
    return hp, Lt, S.from_heat_problem(hp)


def heat_opts(Lt: int):
    return dict(levels=Lt, pre=[3] * (Lt + 1), post=[0] + [3] * Lt, heatmin=1e-4, maxit=400)


def bench_heat(args):
    """One implicit temperature step (titerate) of the heat box on the device, the roofline of its
    colour pass, and the oracle's step on the 65^3 box on the host cores beside the device's."""
    hp, Lt, ctx = heat_setup(args.heat_n)
    n0 = ctx.n0
    ko = heat_opts(Lt)

    def step(c, h):
        c.heat_set_fields(h.ro, h.cp, h.kt, h.tk, h.tk1)
        return c.heat_titerate(h.timestep, **ko)

    st = step(ctx, hp)
    times, its = [], []
    for _ in range(args.heat_steps):
        st = step(ctx, hp)
        times.append(st.seconds)
        its.append(st.iterations)
    ms_gs = ctx.heat_time(0, 0, reps=50)
    gbs = HEAT_GS_BYTES_PER_NODE * n0 / (ms_gs * 1e-3) / 1e9
    ms_sweep = ctx.heat_time(1, 0, reps=20)
    traffic = None
    tf = os.path.join(HERE, "profiles", f"heat{args.heat_n}_colour_pass_pmc.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tj = json.load(f)
        if tj.get("nodes") == n0:
            traffic = tj.get("kernels", {}).get("heatgs", {}).get("hbm_bytes_per_launch")
    log(f"temperature step {args.heat_n}^3: {np.median(times) * 1e3:.1f} ms, {its} iterations, heatsum "
        f"{st.heatsum:.2e}; colour pass {ms_gs:.4f} ms ({gbs:.0f} GB/s)")
    ctx.close()
    out = {"workload": f"heat box {args.heat_n}^3, {Lt + 1} levels, V(3,3), heatmin 1e-4, timestep 1e4 s",
           "ms_per_step": float(np.median(times) * 1e3), "steps_per_s": float(1.0 / np.median(times)),
           "iterations": its, "heatsum": st.heatsum, "bondsum": st.bondsum, "substeps": st.substeps,
           "roofline": {"bound": "hbm", "kernel": "k_heat_gs (one colour, finest level)", "achieved": gbs,
                        "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS, "traffic": traffic,
                        "ms_per_launch": ms_gs, "bytes_per_launch": HEAT_GS_BYTES_PER_NODE * n0},
           "sweep_ms": ms_sweep, "cpu_baseline": None}
    if not args.no_cpu:
        try:
            n = 65
            hp65, Lt65, c65 = heat_setup(n)
            k65 = heat_opts(Lt65)
            c65.heat_set_fields(hp65.ro, hp65.cp, hp65.kt, hp65.tk, hp65.tk1)
            d65 = c65.heat_titerate(hp65.timestep, **k65)
            c65.close()
            O = load_oracle()
            o = O.Oracle(hp65.gx, hp65.gy, hp65.gz, hp65.levels)
            h = O.Heat(o, Lt65)
            h.bc_lines(hp65.bc)
            ts = []
            for _ in range(3):
                h.set_fields(hp65.ro, hp65.cp, hp65.kt, hp65.tk, hp65.tk1)
                h.set(hp65.timestep)
                t0 = time.perf_counter()
                r = h.titerate(hp65.timestep, 25.0, k65["pre"], k65["post"], heatmin=1e-4, maxit=400)
                ts.append(time.perf_counter() - t0)
            med = float(np.median(ts))
            out["cpu_baseline"] = {"value": 1.0 / med, "unit": "temperature steps/s", "cores": O.threads(),
                                   "kind": "port", "sample": f"oracle titerate, heat box 65^3, {Lt65 + 1} levels, "
                                   f"same cycle and heatmin, {r['iterations']} iterations, median of 3",
                                   "device_same_problem_steps_per_s": 1.0 / d65.seconds,
                                   "device_iterations": d65.iterations}
            log(f"temperature step cpu baseline 65^3: {med * 1e3:.1f} ms ({r['iterations']} iterations), "
                f"device {d65.seconds * 1e3:.1f} ms ({d65.iterations})")
        except Exception as e:
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    return out


def bench_batched(args):
    """64 independent instances, sizes drawn from {33,49,65,97}^3, 1-4 seeded
    inclusions each (contrast 1e0..1e6); a step prepares and solves all of
    them concurrently (stk_solve_batch, one stream per instance)."""
    from pardiso2_amd import _lib as S
    from pardiso2_amd import setups
    rng = np.random.default_rng(args.seed)
    sizes = [int(v) for v in rng.choice([33, 49, 65, 97], size=args.instances)]
    problems = [setups.inclusions(n, seed=args.seed * 1000 + i) for i, n in enumerate(sizes)]
    ctxs = [S.from_problem(pb) for pb in problems]
    log(f"{len(ctxs)} instances ready: " + ", ".join(f"{n}^3 x{sizes.count(n)}" for n in sorted(set(sizes))))
    gkw = {} if args.graph is None else {"graph": args.graph}
    # Krylov vectors borrowed per solve (history = 1): memory follows the
    # solves in flight, not the 64 resident instances, which leaves room for
    # restart 256 (hard inclusion instances stall under restart 128)
    mk = lambda pb, prec: S.default_opts(pb.levels, precision=prec, tol=1e-8, maxit=args.maxit,
                                         restart=args.batch_restart, history=1, **gkw)
    opts = [mk(pb, args.precision) for pb in problems]
    for i in range(args.warmup):
        rc, st = S.solve_batch(ctxs, opts, threads=args.threads, prepare=True)
        log(f"warmup {i}: rc {rc}, iterations {min(s.iterations for s in st)}..{max(s.iterations for s in st)}")
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        rc, st = S.solve_batch(ctxs, opts, threads=args.threads, prepare=True)
        lat.append(time.perf_counter() - ts)
        log(f"batch: rc {rc}, {lat[-1]:.2f} s, worst relres {max(s.relres for s in st):.2e}")
    total = time.perf_counter() - t0
    conv = sum(1 for s in st if s.status == 0 and s.relres <= 1e-8)
    # the plain FP64 single-instance path over the same instances, one at a time
    ref = None
    if not args.no_fp64_ref:
        o64 = [mk(pb, 1) for pb in problems]
        ts = time.perf_counter()
        its64 = []
        for c, o in zip(ctxs, o64):
            c.prepare()
            c.hydrostatic_guess()
            its64.append(c.solve(o).iterations)
        t64 = time.perf_counter() - ts
        ref = {"fp64_serial_solves_per_s": len(ctxs) / t64, "fp64_serial_s": t64,
               "fp64_iterations_mean": float(np.mean(its64))}
        log(f"plain FP64, one instance at a time: {t64:.2f} s")
    big = max(range(len(ctxs)), key=lambda i: sizes[i])
    out = {
        "metric": "Stokes solves/s to 1e-8 FP64 relative residual (batched instances)",
        "value": args.instances * args.steps / total, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": total / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 smoother / f64 outer" if args.precision == 0 else "f64", "data": "synthetic",
        "config": {"workload": f"batched: {args.instances} instances, sizes from {{33,49,65,97}}^3, 1-4 inclusions, "
                               f"contrast 1e0..1e6, {args.threads} concurrent streams",
                   "graph": int(opts[0].graph), "restart": args.batch_restart, "history": "borrowed per solve",
                   "sizes": {str(n): sizes.count(n) for n in sorted(set(sizes))}, "seed": args.seed, "tol": 1e-8},
        "converged": conv, "p50_ms": float(np.percentile(np.array(lat) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(lat) * 1e3, 99)),
        "iterations_mean": float(np.mean([s.iterations for s in st])),
        "fp64_single_instance": ref,
        "speedup_vs_fp64_single_instance": (args.instances * args.steps / total) / ref["fp64_serial_solves_per_s"]
        if ref else None,
        "roofline": smoother_roofline(ctxs[big], ctxs[big].n0),
        "cpu_baseline": None if args.no_cpu else cpu_sample(
            problems[min(range(len(sizes)), key=lambda i: sizes[i])], args.cpu_budget),
    }
    print(json.dumps(out), flush=True)
    for c in ctxs:
        c.close()


def bench_served(args):
    """Poisson arrivals of Stokes requests, sizes {33,49,65,97,129}^3 with
    weights .3/.25/.2/.15/.1, random-field viscosity (contrast 1e2..1e6),
    two-phase density; continuous batching over resident contexts."""
    from pardiso2_amd import serve
    from pardiso2_amd import setups
    sizes, weights = [33, 49, 65, 97, 129], [0.3, 0.25, 0.2, 0.15, 0.1]
    rng = np.random.default_rng(args.seed)
    payloads = {}
    for n in sizes:
        payloads[n] = [serve.payload_from_problem(setups.served_request(n, args.seed * 100 + k,
                                                                        float(rng.uniform(2.0, 6.0))))
                       for k in range(3)]
    log("payloads ready (node fields of 3 seeded requests per size)")
    trace = serve.poisson_trace(args.rate, args.duration, sizes, weights, args.seed)
    srv = serve.Server(workers=args.threads, opts_kw=dict(precision=args.precision, tol=1e-8, maxit=args.maxit))
    srv.start()
    # warm the per-size context pools (one request of each size, not timed)
    warm = [(0.0, n) for n in sizes]
    serve.run_stream(srv, warm, payloads, seed=args.seed)
    srv.done.clear()
    log(f"replaying {len(trace)} arrivals over {args.duration:.0f} s at {args.rate:.0f}/s")
    done, t0 = serve.run_stream(srv, trace, payloads, seed=args.seed + 1, progress=log)
    srv.stop()
    lat = np.array([(r.done - r.arrival) * 1e3 for r in done])
    makespan = max(r.done for r in done) - t0
    ok = [r for r in done if not r.error and r.stats is not None and r.stats.status == 0]
    out = {
        "metric": "Stokes solves/s and p50/p99 time-to-converged at 1e-8 (served stream)",
        "value": len(ok) / makespan, "unit": "solves/s", "n_gpus": 1, "steps": len(trace), "warmup": len(warm),
        "ms_per_step": makespan / max(len(trace), 1) * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 smoother / f64 outer" if args.precision == 0 else "f64",
        "data": "synthetic",
        "config": {"workload": f"served: Poisson {args.rate:.0f} req/s for {args.duration:.0f} s, sizes "
                               "{33,49,65,97,129}^3 (.3/.25/.2/.15/.1), random-field contrast 1e2..1e6",
                   "workers": args.threads, "seed": args.seed, "tol": 1e-8},
        "requests": len(trace), "converged": len(ok), "errors": sum(1 for r in done if r.error),
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "service_p50_ms": float(np.percentile([(r.done - r.start) * 1e3 for r in done], 50)),
        "makespan_s": makespan, "contexts_created": srv.pool.created,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="spheres")
    ap.add_argument("--n", type=int, default=257)
    ap.add_argument("--seed", type=int, default=None, help="problem seed (default: 1 for the sinker, else 2)")
    ap.add_argument("--precision", type=int, default=0, help="0 mixed FP32/FP64, 1 FP64")
    ap.add_argument("--restart", type=int, default=128)
    ap.add_argument("--maxit", type=int, default=800)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--banded-n", type=int, default=17,
                    help="grid of the banded direct CPU solve (gausmat3) beside the CPU baseline; 0: skip")
    ap.add_argument("--no-fp64-ref", action="store_true")
    ap.add_argument("--heat-n", type=int, default=257, help="temperature step box (0 skips it)")
    ap.add_argument("--heat-steps", type=int, default=3)
    ap.add_argument("--kernel", choices=sorted(KERNELS) + sorted(HEAT_KERNELS), default=None,
                    help="only launch one finest-level kernel --reps times (for counter runs) and exit")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--instances", type=int, default=64, help="batched workload: number of instances")
    ap.add_argument("--threads", type=int, default=16, help="batched/served: host workers (one stream each)")
    ap.add_argument("--batch-restart", type=int, default=256, help="batched workload: FGMRES restart")
    ap.add_argument("--rate", type=float, default=50.0, help="served workload: arrivals per second")
    ap.add_argument("--duration", type=float, default=60.0, help="served workload: arrival window in seconds")
    ap.add_argument("--graph", type=int, default=None, help="1: replay the preconditioner from HIP graphs")
    args = ap.parse_args()
    if args.seed is None:
        args.seed = {"sinker": 1}.get(args.workload, 2)
    if args.gpus != 1:
        print("bench.py measures one GPU: multi-GPU runs are not part of this build "
              "(instances shard by request; see the design notes).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "served":
        return bench_served(args)

    from pardiso2_amd import _lib as S
    if args.kernel in HEAT_KERNELS:  # counter and trace runs of the temperature kernels (heat box, --heat-n)
        hp, Lt, hctx = heat_setup(args.heat_n)
        hctx.heat_titerate(hp.timestep, **heat_opts(Lt))
        ms = hctx.heat_time(HEAT_KERNELS[args.kernel], 0, reps=args.reps)
        log(f"kernel {args.kernel}: {args.reps} launches, {ms:.4f} ms each, {hctx.n0} nodes")
        return
    pb = make_problem(args.workload, args.n, args.seed)
    L = pb.levels
    pre, post = sweeps_for(L)
    ctx = S.from_problem(pb)
    tiles = ctx.autotune(args.precision)
    if args.kernel:  # counter and trace runs: reps+1 launches of this kernel after the tile autotuning
        ms = ctx.time_kernel(KERNELS[args.kernel], 0, reps=args.reps)
        log(f"kernel {args.kernel}: {args.reps} launches, {ms:.4f} ms each, finest tile {tiles[0]}, "
            f"{smoother_instance(ctx, tiles[0])}")
        return
    log(f"{args.workload} {args.n}^3 ready on the device ({ctx.n0} nodes, {L + 1} levels), smoother tiles {tiles}")
    opts = S.default_opts(L, precision=args.precision, pre=pre, post=post, tol=1e-8, maxit=args.maxit,
                          restart=args.restart)

    def step():
        ctx.prepare()
        ctx.hydrostatic_guess()
        return ctx.solve(opts)

    for i in range(args.warmup):
        st = step()
        log(f"warmup {i}: {st.iterations} iterations, relres {st.relres:.3e}, status {st.status}")
    lat, its, rels, solve_s = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        st = step()
        lat.append(time.perf_counter() - ts)
        its.append(st.iterations)
        rels.append(st.relres)
        solve_s.append(st.seconds)
        log(f"step: {st.iterations} iterations, relres {st.relres:.3e}, {lat[-1] * 1e3:.1f} ms")
        if st.status != 0:
            print(f"warning: solve status {st.status} relres {st.relres:.3e}", file=sys.stderr)
    total = time.perf_counter() - t0
    lat_ms = np.array(lat) * 1e3

    # roofline of the finest-level FP32 smoother, the instance the timed
    # solve runs there (tuned tile, boundary rows in the tiles), and of the
    # FP64 residual (HIP events on the library stream)
    n0 = ctx.n0
    inst = smoother_instance(ctx, tiles[0])
    ms_sm = ctx.time_kernel(0, 0, reps=50)
    ms_res = ctx.time_kernel(4, 0, reps=50)
    gbs_sm = SMOOTHER_BYTES_PER_NODE * n0 / (ms_sm * 1e-3) / 1e9
    gbs_res = RESIDUAL_BYTES_PER_NODE * n0 / (ms_res * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(HERE, "profiles", f"{args.workload}{args.n}_finest_kernels_pmc.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tj = json.load(f)
        k = tj.get("kernels", {}).get("smooth32", {})
        if tj.get("nodes") == n0 and inst in k.get("instance", ""):  # counters of this very instance only
            traffic = k.get("hbm_bytes_per_launch")

    # one marker-advection step (RK4) of the problem's own markers through the
    # timed solution: host coordinates in and out, so PCIe is included
    adv = None
    if pb.mx is not None and len(pb.mx) > 0:
        mx, my, mz = (np.array(a, dtype=np.float32) for a in (pb.mx, pb.my, pb.mz))
        vmax = float(np.abs(ctx.solution()[1::4]).max())
        dt_adv = 0.01 * float(pb.gx[-1]) / max(vmax, 1e-300)
        ts = time.perf_counter()
        ctx.move_markers(mx, my, mz, timestep=dt_adv, want_eii=False)
        t_adv = time.perf_counter() - ts
        adv = {"markers": int(len(mx)), "ms": t_adv * 1e3, "markers_per_s": len(mx) / t_adv,
               "note": "one RK4 step incl. host<->device copies of the coordinates (the kernel alone: profiles/)"}
        log(f"marker advection: {len(mx)} markers in {t_adv * 1e3:.1f} ms")
        del mx, my, mz

    x_dev = ctx.solution()
    fp64_its, err = None, None
    if not args.no_fp64_ref:
        o64 = S.default_opts(L, precision=1, pre=pre, post=post, tol=1e-8, maxit=args.maxit, restart=args.restart)
        ctx.prepare()
        ctx.hydrostatic_guess()
        s64 = ctx.solve(o64)
        fp64_its = s64.iterations
        log(f"fp64 reference solve: {s64.iterations} iterations, relres {s64.relres:.3e}, {s64.seconds:.2f} s")
        err = solution_error(ctx, x_dev)

    heat = None
    if args.heat_n > 0:
        try:
            heat = bench_heat(args)
        except Exception as e:
            log(f"temperature step failed: {e}")
            heat = {"error": str(e)}

    cpu, orel, banded = None, None, None
    log(f"smoother {ms_sm:.4f} ms ({gbs_sm:.0f} GB/s), fp64 residual {ms_res:.4f} ms ({gbs_res:.0f} GB/s)")
    if not args.no_cpu:
        # the oracle's own residual of the timed solve's solution, on the device's node fields
        # (rows on the fly, no assembly)
        try:
            O = load_oracle()
            o = O.Oracle(pb.gx, pb.gy, pb.gz, L)
            k = pb.knobs
            o.knobs(k["p0koef"], k["p1koef"], k["p2koef"], k["v0koef"], k["v1koef"], k["nubeg"], k["nuend"],
                    k["viscmod"], pb.gravity)
            o.bc_lines(pb.bc)
            f = ctx.fields()
            o.set_fields(f["nu"], f["ro"], f["nxx"], f["nxy"], f["nxz"], f["nyz"])
            del f
            orel = float(o.relres_rows(x_dev)[0])
            log(f"oracle residual of the device solution: {orel:.3e}")
            del o
        except Exception as e:
            log(f"oracle residual failed: {e}")
        ctx.close()
        spb = pb if pb.n <= CPU_SAMPLE_N else make_problem(args.workload, CPU_SAMPLE_N, args.seed)
        cpu = cpu_sample(spb, args.cpu_budget)
        log(f"cpu baseline: {cpu.get('value')} solves/s ({cpu.get('sample')})")
        if args.banded_n > 0:
            try:
                banded = banded_baseline(make_problem(args.workload, args.banded_n, args.seed))
                log(f"cpu banded direct: {banded['median_s']:.2f} s per solve at {args.banded_n}^3 "
                    f"(band {banded['band']}, relres {banded['relres']:.1e})")
            except Exception as e:
                banded = {"value": None, "error": str(e)}

    out = {
        "metric": "Stokes solves/s to 1e-8 FP64 relative residual",
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 smoother / f64 outer" if args.precision == 0 else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} {args.n}^3 single solve, {L + 1} levels, free slip, gravity -z",
                   "nodes": n0, "levels": L + 1, "seed": args.seed, "tol": 1e-8,
                   "precision": "mixed" if args.precision == 0 else "fp64"},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "iterations": its,
        "relres": max(rels),
        "reference_norms": {k: v for k, v in st.norms.as_dict().items()
                            if k in ("stoksum", "stoksum1", "stoksum2", "contsum", "contsumb", "bondsum", "maxvxyz",
                                     "converged")},
        "solve_ms_device": float(np.mean(solve_s) * 1e3),
        "fp64_iterations": fp64_its,
        "error_vs_fp64_solve": err,
        "oracle_relres": orel,
        "smoother_tiles": tiles,
        "roofline": {"bound": "hbm", "kernel": f"{inst}, finest level, tile {tiles[0][:3]}", "achieved": gbs_sm,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs_sm / PEAK_HBM_GBS, "traffic": traffic,
                     "ms_per_launch": ms_sm, "bytes_per_launch": SMOOTHER_BYTES_PER_NODE * n0},
        "residual_fp64": {"achieved": gbs_res, "unit": "GB/s", "frac": gbs_res / PEAK_HBM_GBS,
                          "ms_per_launch": ms_res},
        "cpu_baseline": cpu,
        "cpu_banded_direct": banded,
        "marker_advection": adv,
        "temperature_step": heat,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
