"""Admission of the request loop by device memory, on the host with a
simulated device: contexts are fake objects whose bytes count against a
fixed capacity (half at creation, the rest at their first solve, as the
library allocates Krylov vectors on the first solve)."""
import threading
import time

import numpy as np
import pytest

from pardiso2_amd import serve

GB = 1 << 30


class FakeDevice:
    def __init__(self, total):
        self.total = total
        self.used = 0
        self.peak = 0
        self.lock = threading.Lock()
        self.live = 0

    def alloc(self, b):
        with self.lock:
            self.used += b
            self.peak = max(self.peak, self.used)
            if self.used > self.total:
                raise MemoryError("device out of memory")

    def free(self, b):
        with self.lock:
            self.used -= b

    def mem_info(self):
        with self.lock:
            return self.total - self.used, self.total


class FakeCtx:
    def __init__(self, dev, n, payload, restart, solve_s=0.02):
        self.dev, self.n = dev, n
        self.bytes = serve.context_bytes(n, payload["levels"], restart)
        self.half = self.bytes // 2
        dev.alloc(self.half)
        self.first = True
        self.closed = False
        self.solve_s = solve_s
        with dev.lock:
            dev.live += 1

    def knobs(self, *a):
        pass

    def set_fields(self, *a):
        pass

    def prepare(self):
        pass

    def hydrostatic_guess(self):
        pass

    def solve(self, opts):
        if self.first:
            self.dev.alloc(self.bytes - self.half)
            self.first = False
        time.sleep(self.solve_s)
        return type("St", (), {"status": 0, "relres": 1e-9, "iterations": 1})()

    def solution(self):
        return np.zeros(1)

    def close(self):
        assert not self.closed
        self.closed = True
        self.dev.free(self.half if self.first else self.bytes)
        with self.dev.lock:
            self.dev.live -= 1


def _payload(levels):
    return {"levels": levels, "knobs": dict(p0koef=1, p1koef=1, p2koef=0, v0koef=1, v1koef=1, nubeg=1, nuend=1,
                                            viscmod=0), "gravity": (0, 0, 0), "nu": None, "ro": None, "nxx": None,
            "nxy": None, "nxz": None, "nyz": None, "grid": None}


def _server(dev, workers, reserve, restart=128):
    def factory(n, payload):
        return FakeCtx(dev, n, payload, restart)
    return serve.Server(workers=workers, opts_kw={"restart": restart}, reserve_bytes=reserve,
                        mem_info=dev.mem_info, factory=factory)


def test_concurrent_admissions_never_overcommit(monkeypatch):
    monkeypatch.setattr(serve.S, "default_opts", lambda levels, **kw: None)
    one = serve.context_bytes(97, 4, 128)
    dev = FakeDevice(total=int(5.5 * one) + GB)
    srv = _server(dev, workers=16, reserve=GB)
    srv.start()
    p = {97: [_payload(4)]}
    trace = [(0.0, 97)] * 40
    done, _ = serve.run_stream(srv, trace, p)
    srv.stop()
    assert len(done) == 40 and not any(r.error for r in done)
    # 16 workers, room for 5 contexts: never more than fits beside the reserve
    assert dev.peak <= dev.total - GB
    assert srv.pool.created <= 5
    assert srv.pending == 0 and srv.busy == 0
    assert dev.live == 0 and dev.used == 0  # stop() freed every pooled context


def test_idle_contexts_of_other_sizes_are_evicted(monkeypatch):
    monkeypatch.setattr(serve.S, "default_opts", lambda levels, **kw: None)
    a = serve.context_bytes(65, 3, 128)
    b = serve.context_bytes(129, 5, 128)
    dev = FakeDevice(total=b + a + GB + (1 << 20))
    srv = _server(dev, workers=4, reserve=GB)
    srv.start()
    p = {65: [_payload(3)], 129: [_payload(5)]}
    # four 65^3 requests leave idle contexts behind; a 129^3 request then
    # needs their memory
    done, _ = serve.run_stream(srv, [(0.0, 65)] * 4 + [(0.3, 129)], p)
    assert not any(r.error for r in done)
    assert srv.pool.evicted >= 1
    assert dev.peak <= dev.total
    srv.stop()
    assert dev.used == 0


def test_request_larger_than_device_fails_instead_of_waiting(monkeypatch):
    monkeypatch.setattr(serve.S, "default_opts", lambda levels, **kw: None)
    dev = FakeDevice(total=GB)
    srv = _server(dev, workers=2, reserve=GB // 2)
    srv.start()
    t = time.perf_counter()
    done, _ = serve.run_stream(srv, [(0.0, 129)], {129: [_payload(5)]})
    srv.stop()
    assert time.perf_counter() - t < 10.0
    assert len(done) == 1 and done[0].error
    assert srv.pending == 0 and srv.busy == 0
