"""Request loop of independent Stokes solves on one GPU (continuous batching).

Requests (a grid size plus node fields) arrive over time.  A request is
admitted as soon as a worker slot is free and the device has memory for its
context.  Each worker takes a resident context of the request's size from a
per-size pool (hierarchy, boundary tables, work vectors and Krylov history
stay allocated in HBM between requests), loads the node fields, prepares
(viscosity restriction, stencil viscosities, coarsest inverse, Schur
hierarchy) and solves on its own stream, so the work of requests that are in
flight together overlaps on the device.  Results do not depend on what else is
running: every context is deterministic and private.
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as S
from . import setups


def context_bytes(n: int, levels: int, restart: int) -> int:
    """Device bytes of one context of an n^3 box: per level about 415 B/node
    (geometry planes, FP32 and FP64 vectors, Schur planes in natural and
    colour-compact order, Schur vectors), plus 2 (restart + 1) FP64 Krylov
    vectors of 4 unknowns per finest node."""
    total, m = 0, n
    for _ in range(levels + 1):
        total += 415 * m ** 3
        m = (m - 5) // 2 + 5
    return int(total + 2 * (restart + 1) * 32 * n ** 3 + (64 << 20))


@dataclass(order=True)
class Request:
    arrival: float
    rid: int = field(compare=False)
    n: int = field(compare=False)
    payload: dict = field(compare=False)
    done: float = field(default=0.0, compare=False)
    start: float = field(default=0.0, compare=False)
    stats: object = field(default=None, compare=False)
    solution: object = field(default=None, compare=False)
    error: str = field(default="", compare=False)


def payload_from_problem(pb) -> dict:
    """Node fields of a problem (averaged from its markers on the device)."""
    ctx = S.Stokes(pb.gx, pb.gy, pb.gz, pb.levels)
    try:
        k = pb.knobs
        ctx.knobs(k["p0koef"], k["p1koef"], k["p2koef"], k["v0koef"], k["v1koef"], k["nubeg"], k["nuend"],
                  k["viscmod"], pb.gravity)
        ctx.markers_to_nodes(pb.mx, pb.my, pb.mz, pb.mnu, pb.mro)
        f = ctx.fields()
        n0 = ctx.n0
        return dict(nu=f["nu"][:n0].copy(), ro=f["ro"][:n0].copy(), nxx=f["nxx"], nxy=f["nxy"], nxz=f["nxz"],
                    nyz=f["nyz"], knobs=dict(k), gravity=pb.gravity, levels=pb.levels, grid=(pb.gx, pb.gy, pb.gz))
    finally:
        ctx.close()


class ContextPool:
    """Resident contexts per grid size, each handed to one request at a time.

    ``factory(n, payload)`` builds a new context (a library context with the
    free-slip boundary lines of an n^3 box unless given)."""

    def __init__(self, restart: int, factory=None):
        self.restart = restart
        self.free: dict[int, list] = {}
        self.lock = threading.Lock()
        self.created = 0
        self.evicted = 0
        self.factory = factory or self._make

    @staticmethod
    def _make(n: int, payload: dict):
        gx, gy, gz = payload["grid"]
        ctx = S.Stokes(gx, gy, gz, payload["levels"])
        ctx.bc_lines(setups.free_slip_lines(n, n, n, gauge=None))
        return ctx

    def take_idle(self, n: int):
        """An idle resident context of size n, removed from the pool, or None."""
        with self.lock:
            lst = self.free.get(n)
            return lst.pop() if lst else None

    def create(self, n: int, payload: dict):
        ctx = self.factory(n, payload)
        with self.lock:
            self.created += 1
        return ctx

    def release(self, n: int, ctx):
        with self.lock:
            self.free.setdefault(n, []).append(ctx)

    def evict_other(self, n: int) -> bool:
        """Free the idle contexts of every size but n; True if any was freed."""
        with self.lock:
            victims = [c for m, lst in self.free.items() if m != n for c in lst]
            for m in list(self.free):
                if m != n:
                    self.free[m] = []
            self.evicted += len(victims)
        for c in victims:
            c.close()
        return bool(victims)

    def close(self):
        with self.lock:
            for lst in self.free.values():
                for c in lst:
                    c.close()
            self.free.clear()


class Server:
    """Worker threads over a shared admission queue.

    Admission by device memory: a request reuses an idle resident context of
    its size when one exists; otherwise a new context is admitted only when
    the free device memory, less the bytes admitted for contexts that are not
    fully allocated yet (a context reaches its full size during its first
    solve, when the Krylov vectors are allocated), covers the context's bytes
    plus a reserve.  A request that does not fit first frees the idle
    contexts of other sizes, then waits for running solves to end.  When no
    solve is running and nothing is pending, it is admitted anyway, so that a
    request larger than the device fails with an error instead of waiting
    forever."""

    def __init__(self, workers: int = 16, opts_kw: dict | None = None, reserve_bytes: int = 8 << 30,
                 keep_solutions: bool = False, mem_info=None, factory=None):
        self.workers = workers
        self.opts_kw = dict(opts_kw or {})
        self.opts_kw.setdefault("graph", 0)  # 1: each context replays its captured preconditioner
        self.restart = self.opts_kw.get("restart", 128)
        self.pool = ContextPool(self.restart, factory)
        self.reserve = reserve_bytes
        self.keep = keep_solutions
        self.mem_info = mem_info or S.mem_info
        self.q: queue.Queue = queue.Queue()
        self.done: list = []
        self.lock = threading.Lock()
        self.mem_cv = threading.Condition()
        self.pending = 0   # bytes admitted for contexts not yet fully allocated
        self.busy = 0      # requests holding a context
        self.threads: list = []

    def _context_for(self, n: int, payload: dict):
        """(context, bytes still pending for it): an idle resident one, or a
        new one once admitted."""
        need = context_bytes(n, payload["levels"], self.restart)
        with self.mem_cv:
            while True:
                ctx = self.pool.take_idle(n)
                if ctx is not None:
                    self.busy += 1
                    return ctx, 0
                free, _ = self.mem_info()
                fits = free - self.pending - need >= self.reserve
                if not fits and self.pool.evict_other(n):
                    continue
                if fits or (self.busy == 0 and self.pending == 0):
                    self.pending += need
                    self.busy += 1
                    break
                self.mem_cv.wait(0.01)
        try:
            return self.pool.create(n, payload), need
        except Exception:
            self._finish(need)
            raise

    def _finish(self, pending: int):
        with self.mem_cv:
            self.pending -= pending
            self.busy -= 1
            self.mem_cv.notify_all()

    def _work(self):
        while True:
            req = self.q.get()
            if req is None:
                return
            p = req.payload
            try:
                ctx, pend = self._context_for(req.n, p)
            except Exception as e:  # a failed request must still complete
                req.error, req.done = repr(e), time.perf_counter()
                with self.lock:
                    self.done.append(req)
                continue
            req.start = time.perf_counter()
            try:
                k = p["knobs"]
                ctx.knobs(k["p0koef"], k["p1koef"], k["p2koef"], k["v0koef"], k["v1koef"], k["nubeg"], k["nuend"],
                          k["viscmod"], p["gravity"])
                ctx.set_fields(p["nu"], p["ro"], p["nxx"], p["nxy"], p["nxz"], p["nyz"])
                ctx.prepare()
                ctx.hydrostatic_guess()
                req.stats = ctx.solve(S.default_opts(p["levels"], **self.opts_kw))
                if self.keep:
                    req.solution = ctx.solution()
            except Exception as e:
                req.error = repr(e)
            finally:
                self.pool.release(req.n, ctx)
                self._finish(pend)
            req.done = time.perf_counter()
            with self.lock:
                self.done.append(req)

    def start(self):
        for _ in range(self.workers):
            t = threading.Thread(target=self._work, daemon=True)
            t.start()
            self.threads.append(t)

    def submit(self, req: Request):
        self.q.put(req)

    def stop(self):
        for _ in self.threads:
            self.q.put(None)
        for t in self.threads:
            t.join()
        self.pool.close()


def poisson_trace(rate: float, duration: float, sizes, weights, seed: int):
    """Arrival times (s) and sizes of a Poisson stream."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append((t, int(rng.choice(sizes, p=np.asarray(weights) / np.sum(weights)))))


def run_stream(server: Server, trace, payloads: dict, seed: int = 0, progress=None, every: float = 10.0):
    """Replay an arrival trace against a started server; returns the finished
    requests.  payloads[n] is a list of payload dicts for size n.  progress,
    if given, is called with a status line every `every` seconds."""
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    last = t0

    def report():
        nonlocal last
        now = time.perf_counter()
        if progress is not None and now - last >= every:
            last = now
            with server.lock:
                nd = len(server.done)
            progress(f"{now - t0:.0f} s: {nd} of {len(trace)} requests done")

    for rid, (ta, n) in enumerate(trace):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        p = payloads[n][int(rng.integers(len(payloads[n])))]
        server.submit(Request(arrival=t0 + ta, rid=rid, n=n, payload=p))
        report()
    while True:
        with server.lock:
            if len(server.done) >= len(trace):
                break
        report()
        time.sleep(0.01)
    return sorted(server.done, key=lambda r: r.rid), t0
