"""The preconditioner replayed from a captured HIP graph gives the eager
launches' results exactly: every replay is checked implicitly by the solve
(each Krylov step replays the graph between eager orthogonalisation work and
host synchronisations), and the final iterates must match bit for bit, also
after a re-prepare with new fields that refills the captured buffers."""
import numpy as np
import pytest

from pardiso2_amd import _lib as S
from pardiso2_amd import setups

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_matches_eager(precision):
    pb = setups.random_field(33, setups.coarsest_levels(33), lo=19.0, hi=23.0, field_n=33)
    g = S.from_problem(pb)
    try:
        res = {}
        for graph in (0, 1, 1):
            g.prepare()
            g.hydrostatic_guess()
            st = g.solve(S.default_opts(pb.levels, precision=precision, graph=graph))
            assert st.status == 0 and st.relres <= 1e-8
            res.setdefault(graph, []).append((st.iterations, g.solution()))
        (ie, xe), = res[0]
        for it, x in res[1]:
            assert it == ie
            np.testing.assert_array_equal(x, xe)
        # new node fields, same sizes: the capture is kept and refilled
        f = g.fields()
        scale = np.float32(3.0)
        g.set_fields(f["nu"][:g.n0] * scale, f["ro"][:g.n0], f["nxx"] * scale, f["nxy"] * scale, f["nxz"] * scale,
                     f["nyz"] * scale)
        out = []
        for graph in (0, 1):
            g.prepare()
            g.hydrostatic_guess()
            st = g.solve(S.default_opts(pb.levels, precision=precision, graph=graph))
            out.append((st.iterations, g.solution()))
        assert out[0][0] == out[1][0]
        np.testing.assert_array_equal(out[0][1], out[1][1])
    finally:
        g.close()


def test_concurrent_captures_from_two_threads():
    """Two host threads each capture and replay their own context's
    preconditioner at the same time, while a third allocates and releases
    contexts: every solve equals the same context's eager solve bit for bit."""
    import threading
    probs = [setups.random_field(n, setups.coarsest_levels(n), seed=s, lo=19.0, hi=23.0, field_n=33)
             for n, s in ((33, 1), (25, 2))]
    ctxs = [S.from_problem(pb) for pb in probs]
    try:
        eager = []
        for g, pb in zip(ctxs, probs):
            g.hydrostatic_guess()
            st = g.solve(S.default_opts(pb.levels, graph=0))
            eager.append((st.iterations, g.solution()))
        out = [[] for _ in ctxs]
        errs = []
        stop = threading.Event()

        def solver(i):
            try:
                g, pb = ctxs[i], probs[i]
                for _ in range(3):
                    g.prepare()
                    g.hydrostatic_guess()
                    st = g.solve(S.default_opts(pb.levels, graph=1))
                    out[i].append((st.status, st.iterations, g.solution()))
            except Exception as e:  # reported below
                errs.append(e)

        def churn():
            try:
                pb = setups.sinker(17, setups.coarsest_levels(17))
                while not stop.is_set():
                    S.from_problem(pb).close()
            except Exception as e:
                errs.append(e)

        th = [threading.Thread(target=solver, args=(i,)) for i in range(len(ctxs))]
        ch = threading.Thread(target=churn)
        ch.start()
        for t in th:
            t.start()
        for t in th:
            t.join()
        stop.set()
        ch.join()
        assert not errs, errs
        for (ie, xe), res in zip(eager, out):
            assert len(res) == 3
            for status, it, x in res:
                assert status == 0 and it == ie
                np.testing.assert_array_equal(x, xe)
    finally:
        for g in ctxs:
            g.close()
