"""Concurrent instances: a batch (one stream per instance) and the served
request loop give every instance exactly the result it has alone."""
import numpy as np
import pytest

from pardiso2_amd import _lib as S
from pardiso2_amd import serve, setups

pytestmark = pytest.mark.gpu


def _problems():
    return [setups.sinker(17, 2, eta_in=1e22), setups.inclusions(21, seed=4), setups.inclusions(25, seed=5),
            setups.inclusions(21, seed=6)]


def _alone(pb):
    g = S.from_problem(pb)
    try:
        g.hydrostatic_guess()
        st = g.solve(S.default_opts(pb.levels))
        return g.solution(), st
    finally:
        g.close()


def test_batch_equals_alone():
    probs = _problems()
    alone = [_alone(pb) for pb in probs]
    ctxs = [S.from_problem(pb) for pb in probs]
    try:
        rc, st = S.solve_batch(ctxs, [S.default_opts(pb.levels) for pb in probs], threads=len(ctxs), prepare=True)
        assert rc == 0
        for (xa, sa), c, s in zip(alone, ctxs, st):
            assert s.status == 0 and s.relres <= 1e-8
            assert s.iterations == sa.iterations
            np.testing.assert_array_equal(c.solution(), xa)
        # a second batch on the same resident contexts reproduces the first
        rc2, st2 = S.solve_batch(ctxs, [S.default_opts(pb.levels) for pb in probs], threads=2, prepare=True)
        for (xa, _), c in zip(alone, ctxs):
            np.testing.assert_array_equal(c.solution(), xa)
    finally:
        for c in ctxs:
            c.close()


def test_borrowed_krylov_vectors_give_the_same_solves():
    """history = 1 (Krylov vectors from the shared pool, restart 256) gives
    bit-identical solves to the context's own vectors, alone and in a batch
    whose instances outnumber the workers (sets are reused across sizes)."""
    probs = _problems()
    ref = []
    for pb in probs:
        g = S.from_problem(pb)
        try:
            g.hydrostatic_guess()
            st = g.solve(S.default_opts(pb.levels, restart=256))
            ref.append((g.solution(), st.iterations))
        finally:
            g.close()
    ctxs = [S.from_problem(pb) for pb in probs]
    try:
        for (xr, itr), c, pb in zip(ref, ctxs, probs):
            c.hydrostatic_guess()
            st = c.solve(S.default_opts(pb.levels, restart=256, history=1))
            assert st.status == 0 and st.iterations == itr
            np.testing.assert_array_equal(c.solution(), xr)
        rc, st = S.solve_batch(ctxs, [S.default_opts(pb.levels, restart=256, history=1) for pb in probs], threads=2,
                               prepare=True)
        assert rc == 0
        for (xr, itr), c, s in zip(ref, ctxs, st):
            assert s.iterations == itr
            np.testing.assert_array_equal(c.solution(), xr)
    finally:
        for c in ctxs:
            c.close()


def test_served_stream_equals_alone():
    pbs = {17: setups.served_request(17, 1, 4.0), 21: setups.served_request(21, 2, 5.0)}
    payloads = {n: [serve.payload_from_problem(pb)] for n, pb in pbs.items()}
    alone = {}
    for n, pb in pbs.items():
        p = payloads[n][0]
        g = S.Stokes(*p["grid"], p["levels"])
        try:
            k = p["knobs"]
            g.knobs(k["p0koef"], k["p1koef"], k["p2koef"], k["v0koef"], k["v1koef"], k["nubeg"], k["nuend"],
                    k["viscmod"], p["gravity"])
            g.bc_lines(setups.free_slip_lines(n, n, n, gauge=None))
            g.set_fields(p["nu"], p["ro"], p["nxx"], p["nxy"], p["nxz"], p["nyz"])
            g.prepare()
            g.hydrostatic_guess()
            g.solve(S.default_opts(p["levels"]))
            alone[n] = g.solution()
        finally:
            g.close()
    srv = serve.Server(workers=3, keep_solutions=True, reserve_bytes=1 << 30)
    srv.start()
    trace = [(0.0, 17), (0.0, 21), (0.001, 17), (0.002, 21), (0.003, 17), (0.004, 21)]
    done, _ = serve.run_stream(srv, trace, payloads)
    srv.stop()
    assert len(done) == len(trace)
    for r in done:
        assert not r.error, r.error
        assert r.stats.status == 0
        np.testing.assert_array_equal(r.solution, alone[r.n])
