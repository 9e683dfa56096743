"""Device solves: convergence to the build's FP64 relative residual, checked
with the oracle's own assembled operator; mixed and FP64 modes agree; error
paths return status codes."""
import numpy as np
import pytest

from pardiso2_amd import _lib as S
from pardiso2_amd import setups

pytestmark = pytest.mark.gpu


def _problem(kind, n, **kw):
    L = setups.coarsest_levels(n)
    pb = setups.sinker(n, L, **kw) if kind == "sinker" else setups.random_field(n, L, **kw)
    pb.bc = setups.free_slip_lines(n, n, n, gauge=None)
    return pb


@pytest.fixture(scope="module")
def sinker33(oracle_mod):
    pb = _problem("sinker", 33)
    return pb, oracle_mod.from_problem(pb)


@pytest.mark.parametrize("precision", [0, 1])
def test_sinker33_converges_and_oracle_agrees(sinker33, precision):
    pb, o = sinker33
    g = S.from_problem(pb)
    try:
        g.hydrostatic_guess()
        st = g.solve(precision=precision, tol=1e-8)
        assert st.status == 0, (st.status, st.relres, st.iterations)
        assert st.relres <= 1e-8
        x = g.solution()
        o.arr("sol0")[: x.size] = x
        rel, _ = o.relres()
        assert rel <= 1.0000001e-8
    finally:
        g.close()


def test_mixed_and_fp64_solutions_agree(sinker33):
    pb, o = sinker33
    g = S.from_problem(pb)
    try:
        sols = []
        for prec in (0, 1):
            g.hydrostatic_guess()
            st = g.solve(precision=prec, tol=1e-10)
            assert st.relres <= 1e-10
            g.pressure_gauge(2, 2, 2, 0.0)
            sols.append(g.solution())
        vel = np.arange(sols[0].size) % 4 != 0
        vs = np.abs(sols[1][vel]).max()
        np.testing.assert_allclose(sols[0][vel], sols[1][vel], atol=1e-6 * vs)
        pr = ~vel
        ps = np.abs(sols[1][pr]).max()
        np.testing.assert_allclose(sols[0][pr], sols[1][pr], atol=1e-6 * ps)
    finally:
        g.close()


def test_device_solution_matches_oracle_gcr(oracle_mod):
    pb = _problem("sinker", 17, eta_in=1e22)
    o = oracle_mod.from_problem(pb)
    it, xo, hist, _ = o.gcr([2, 4, 0], [2, 4, 0], mode=1, tol=1e-10, maxit=400, restart=128)
    assert hist[-1] <= 1e-10
    g = S.from_problem(pb)
    try:
        g.hydrostatic_guess()
        st = g.solve(precision=1, tol=1e-10, pre=[2, 4, 0], post=[2, 4, 0])
        assert st.relres <= 1e-10
        xg = g.solution()
        vel = np.arange(xg.size) % 4 != 0
        np.testing.assert_allclose(xg[vel], xo[vel], atol=1e-6 * np.abs(xo[vel]).max())
    finally:
        g.close()


def test_random_field_small_converges(oracle_mod):
    pb = _problem("randfield", 33, lo=19.0, hi=23.0)
    g = S.from_problem(pb)
    try:
        g.hydrostatic_guess()
        st = g.solve(tol=1e-8)
        assert st.status == 0 and st.relres <= 1e-8, (st.relres, st.iterations)
    finally:
        g.close()


def test_vpiterate_entry_point_solves_bound_context(sinker33):
    pb, o = sinker33
    g = S.from_problem(pb)
    try:
        g.bind()
        g.hydrostatic_guess()
        S.lib().vpiterate(0, 0)
        rel, blocks = g.relres()
        assert rel <= 1e-8
    finally:
        g.close()


def test_error_paths_return_codes():
    pb = _problem("sinker", 17)
    g = S.Stokes(pb.gx, pb.gy, pb.gz, pb.levels)
    try:
        o = S.default_opts(pb.levels)
        assert S.lib().stk_solve(g.h, S.C.byref(o), None) == -7  # solve before prepare
        assert S.lib().stk_prepare(g.h) == -7                    # no fields yet
        bad = np.full(g.n0, -1.0, np.float32)
        ok = np.ones(g.n0, np.float32)
        fp = S._fp
        assert S.lib().stk_set_fields(g.h, fp(bad), fp(ok), fp(ok), fp(ok), fp(ok), fp(ok)) == -1
        g.set_fields(ok * 1e21, ok * 3300, ok * 1e21, ok * 1e21, ok * 1e21, ok * 1e21)
        # no boundary equations: the outer layer's rows would leave the grid (STK_EBCFORM)
        assert S.lib().stk_prepare(g.h) == -9
    finally:
        g.close()
    L = S.lib()
    assert L.gausmat4(1, 0, 0, 0) in (0, -8)
    # grid lines that do not increase strictly are refused
    for k in (4, len(pb.gx) - 1):
        gx = pb.gx.copy()
        gx[k] = gx[k - 1]
        with pytest.raises(RuntimeError):
            S.Stokes(gx, pb.gy, pb.gz, pb.levels)


def test_solve_reports_nonconvergence_instead_of_exiting():
    pb = _problem("sinker", 17, eta_in=1e23)
    g = S.from_problem(pb)
    try:
        g.hydrostatic_guess()
        st = g.solve(tol=1e-14, maxit=3)
        assert st.status == -5 and st.iterations == 3
    finally:
        g.close()
