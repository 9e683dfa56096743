"""The reference-named entry points acting on a caller's own arrays
(stk_bind_arrays): the reference's multigrid loop written against
val1/lin0/num0/pos0/sol0/fre0/fre1 runs unchanged with the library in place
of parallel/gaus3mg_MG.c and the solver routines of parallel/move3mg.c.

- gausmat2 stores every assembled row exactly where and as the oracle's row
  store holds it (parallel/gaus3mg_MG.c:311-365);
- gausmat4 over that row store on the device is bit-identical to the
  oracle's gausmat4 in the same multicolour order, sol0 and fre0, on every
  level (parallel/move3mg.c:3855-3960);
- vprestrict, vpprolong and ronurestrict on the arrays are bit-identical to
  the oracle's (move3mg.c:3966, 3535, 3387);
- one V-cycle through the reference names (pardisomat for the coarsest
  level) matches the oracle's V-cycle to the coarse solve's rounding;
- vpiterate solves from the arrays' fields and boundary tables and leaves
  sol0, pr, vx, vy, vz and the stress planes in them.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from pardiso2_amd import _lib as S
from pardiso2_amd import setups

pytestmark = pytest.mark.gpu

N = 17


def _say(msg):
    print(f"[compat] {msg}", flush=True)
    sys.stdout.flush()


def _caller_arrays(o):
    """The caller's globals: copies of the oracle's, in the reference's types."""
    a = {k: np.array(o.arr(k), copy=True) for k in ("sol0", "fre0", "fre1", "val1", "lin0", "num0", "pos0", "nu",
                                                   "ro", "nd", "nxx", "nxy", "nxz", "nyz", "bondm")}
    a["bondv1"] = np.array(o.arr("bondv1"), copy=True)
    a["bondn1"] = np.array(o.arr("bondn1"), copy=True)
    a["bondnum"] = np.array([o.nbond], dtype=np.int64)
    return a


@pytest.fixture(scope="module")
def sinker(oracle_mod):
    L = setups.coarsest_levels(N)
    pb = setups.sinker(N, L, eta_in=1e22, eta_out=1e20)
    pb.bc = setups.free_slip_lines(N, N, N, gauge=(2, 2, 2))
    o = oracle_mod.from_problem(pb)
    g = S.from_problem(pb)
    g.bind()
    yield pb, o, g
    S.bind_arrays(None)
    g.close()


def _rows(o, l):
    b, n = 4 * o.info[l][5], 4 * o.info[l][4]
    return b, b + n


def test_gausmat2_fills_the_callers_row_store(sinker):
    pb, o, g = sinker
    ref = _caller_arrays(o)
    a = {k: np.zeros_like(v) for k, v in ref.items()}
    keep = S.bind_arrays(a)
    un1 = (C.c_long * 64)()
    ui1 = (C.c_double * 64)()
    num0, pos0 = o.arr("num0"), o.arr("pos0")
    stored = 0
    for r in range(4 * o.ntot):
        if not num0[r]:
            continue
        idx, val, rhs = o.row(r)
        un1[0] = len(idx)
        for k, (j, v) in enumerate(zip(idx, val)):
            un1[k + 1] = int(j)
            ui1[k + 1] = float(v)
        ui1[0] = rhs
        assert S.lib().gausmat2(1, r, int(pos0[r]), un1, ui1) == 0
        stored += 1
    np.testing.assert_array_equal(a["num0"], ref["num0"])
    live = ref["num0"] > 0
    np.testing.assert_array_equal(a["pos0"][live], ref["pos0"][live])
    np.testing.assert_array_equal(a["fre1"][live], ref["fre1"][live])
    used = np.zeros(a["val1"].size, bool)
    for r in np.nonzero(live)[0]:
        used[ref["pos0"][r]:ref["pos0"][r] + ref["num0"][r]] = True
    np.testing.assert_array_equal(a["val1"][used], ref["val1"][used])
    np.testing.assert_array_equal(a["lin0"][used], ref["lin0"][used])
    _say(f"gausmat2: {stored} rows stored as the oracle's row store holds them")
    del keep


def test_gausmat4_on_the_callers_rows_is_bit_identical(sinker):
    pb, o, g = sinker
    rng = np.random.default_rng(4)
    for l in range(pb.levels + 1):
        b, e = _rows(o, l)
        x0 = rng.standard_normal(e - b) * 1e-3
        o.arr("sol0")[b:e] = x0
        a = _caller_arrays(o)
        S.bind_arrays(a)
        assert S.lib().gausmat4(3, e - 1, b, l) == 0
        o.gausmat4_colored(3, l)
        np.testing.assert_array_equal(a["sol0"][b:e], o.arr("sol0")[b:e], err_msg=f"level {l}")
        np.testing.assert_array_equal(a["fre0"][b:e], o.arr("fre0")[b:e], err_msg=f"level {l}")
        # only the level's rows change
        np.testing.assert_array_equal(a["sol0"][:b], o.arr("sol0")[:b])
        assert np.abs(a["sol0"][b:e] - x0).max() > 0
    _say("gausmat4 on the caller's rows: bit-identical on every level")


def test_transfers_on_the_callers_arrays_are_bit_identical(sinker):
    pb, o, g = sinker
    rng = np.random.default_rng(5)
    for l in range(1, pb.levels + 1):
        bf, ef = _rows(o, l - 1)
        bc, ec = _rows(o, l)
        o.arr("fre0")[bf:ef] = rng.standard_normal(ef - bf)
        a = _caller_arrays(o)
        S.bind_arrays(a)
        S.lib().vprestrict(l)
        o.vprestrict(l)
        np.testing.assert_array_equal(a["fre1"][bc:ec], o.arr("fre1")[bc:ec], err_msg=f"vprestrict {l}")
        np.testing.assert_array_equal(a["sol0"][bc:ec], o.arr("sol0")[bc:ec], err_msg=f"vprestrict {l}")
        o.arr("sol0")[bc:ec] = rng.standard_normal(ec - bc)
        o.arr("sol0")[bf:ef] = rng.standard_normal(ef - bf)
        a = _caller_arrays(o)
        S.bind_arrays(a)
        S.lib().vpprolong(l, 0)
        o.vpprolong(l, 0)
        np.testing.assert_array_equal(a["sol0"][bf:ef], o.arr("sol0")[bf:ef], err_msg=f"vpprolong {l}")
        a = _caller_arrays(o)
        a["nu"][o.info[l][5]:o.info[l][5] + o.info[l][4]] = 0.0
        a["ro"][o.info[l][5]:o.info[l][5] + o.info[l][4]] = 0.0
        S.bind_arrays(a)
        S.lib().ronurestrict(l)
        np.testing.assert_array_equal(a["nu"], o.arr("nu"), err_msg=f"ronurestrict {l}")
        np.testing.assert_array_equal(a["ro"], o.arr("ro"), err_msg=f"ronurestrict {l}")
    _say("vprestrict / vpprolong / ronurestrict on the caller's arrays: bit-identical")


def test_vcycle_through_the_reference_names(sinker):
    pb, o, g = sinker
    L = pb.levels
    pre = [2] * (L + 1)
    post = [2] * (L + 1)
    b0, e0 = _rows(o, 0)
    rng = np.random.default_rng(6)
    o.arr("sol0")[:] = 0.0
    o.arr("sol0")[b0:e0] = rng.standard_normal(e0 - b0) * 1e-3
    a = _caller_arrays(o)
    S.bind_arrays(a)
    un1 = (C.c_long * 64)()
    ui1 = (C.c_double * 64)()
    for l in range(L):
        b, e = _rows(o, l)
        assert S.lib().gausmat4(pre[l], e - 1, b, l) == 0
        S.lib().vprestrict(l + 1)
    # coarsest level: its rows with the restricted right-hand side, then the solve
    bL, eL = _rows(o, L)
    for r in range(bL, eL):
        i = r - bL
        if a["num0"][r] == 0:
            un1[0], un1[1], ui1[0], ui1[1] = 1, i, 0.0, 1.0
        else:
            p, k = a["pos0"][r], a["num0"][r]
            un1[0] = int(k)
            for q in range(k):
                un1[q + 1] = int(a["lin0"][p + q] - bL)
                ui1[q + 1] = float(a["val1"][p + q])
            ui1[0] = float(a["fre1"][r])
        assert S.lib().pardisomat(1, i, r, 0, 0, un1, ui1) == 0
    assert S.lib().pardisomat(0, eL - bL, 0, 0, 0, un1, ui1) == 0
    S.lib().vpprolong(L, 0)
    for l in range(L - 1, 0, -1):
        b, e = _rows(o, l)
        assert S.lib().gausmat4(post[l], e - 1, b, l) == 0
        S.lib().vpprolong(l, 0)
    assert S.lib().gausmat4(post[0], e0 - 1, b0, 0) == 0
    o.vcycle(pre, post, mode=2)
    ref = o.arr("sol0")[b0:e0]
    np.testing.assert_allclose(a["sol0"][b0:e0], ref, rtol=1e-8, atol=1e-8 * np.abs(ref).max())
    _say(f"V-cycle through the reference names: max diff {np.abs(a['sol0'][b0:e0] - ref).max():.2e} "
         f"of {np.abs(ref).max():.2e}")


def test_vpiterate_solves_from_the_callers_arrays(oracle_mod):
    L = setups.coarsest_levels(N)
    pb = setups.sinker(N, L, eta_in=1e22, eta_out=1e20)
    pb.bc = setups.free_slip_lines(N, N, N, gauge=None)
    o = oracle_mod.from_problem(pb)
    g = S.from_problem(pb)
    try:
        g.bind()
        a = _caller_arrays(o)
        a["sol0"][:] = 0.0
        n0 = o.info[0][4]
        for k in ("pr",):
            a[k] = np.zeros(n0)
        for k in ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz", "exx", "eyy", "ezz", "exy", "exz",
                  "eyz"):
            a[k] = np.zeros(n0, np.float32)
        S.bind_arrays(a)
        S.lib().vpiterate(0, 0)
        x = a["sol0"][:4 * n0]
        rel = o.relres_rows(x)[0]
        _say(f"vpiterate on the caller's arrays: oracle relres {rel:.2e}")
        assert rel <= 1.0000001e-8
        np.testing.assert_array_equal(a["vx"][np.nonzero(a["vx"])], x[1::4][np.nonzero(a["vx"])].astype(np.float32))
        sig, eps = g.stress_strain()
        for k, name in enumerate(("sxx", "syy", "szz", "sxy", "sxz", "syz")):
            np.testing.assert_array_equal(a[name], sig[k].astype(np.float32), err_msg=name)
        for k, name in enumerate(("exx", "eyy", "ezz", "exy", "exz", "eyz")):
            np.testing.assert_array_equal(a[name], eps[k].astype(np.float32), err_msg=name)
        assert np.abs(a["sxy"]).max() > 0
    finally:
        S.bind_arrays(None)
        g.close()


def test_vpiterate_on_the_periodic_shear_box(oracle_mod):
    """The reference's own setup, periodic in x and y (parallel/init_inclusion.t3c),
    solved by vpiterate from the caller's arrays; pr on line 0 of each
    periodic axis is the last line's (move3mg.c:3321-3379), and the stress
    planes are the reference's periodic ones."""
    pb = setups.periodic_shear(N, count=1, r=0.25)
    o = oracle_mod.from_problem(pb)
    g = S.from_problem(pb)
    try:
        g.bind()
        a = _caller_arrays(o)
        a["sol0"][:] = 0.0
        n0 = o.info[0][4]
        a["pr"] = np.zeros(n0)
        for k in ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz", "exx", "eyy", "ezz", "exy", "exz",
                  "eyz"):
            a[k] = np.zeros(n0, np.float32)
        S.bind_arrays(a)
        S.lib().vpiterate(0, 0)
        x = a["sol0"][:4 * n0]
        rel = o.relres_rows(x)[0]
        _say(f"vpiterate on the periodic shear box: oracle relres {rel:.2e}")
        assert rel <= 1.0000001e-8
        pr = a["pr"].reshape(N, N, N)  # [m3, m1, m2]
        np.testing.assert_array_equal(pr[1:, 0, 1:], pr[1:, N - 1, 1:])
        np.testing.assert_array_equal(pr[1:, :, 0], pr[1:, :, N - 1])
        sig, eps = g.stress_strain()
        for k, name in enumerate(("sxx", "syy", "szz", "sxy", "sxz", "syz")):
            np.testing.assert_array_equal(a[name], sig[k].astype(np.float32), err_msg=name)
        sxx = a["sxx"].reshape(N, N, N)
        np.testing.assert_array_equal(sxx[1:, 0, 1:-1], sxx[1:, N - 1, 1:-1])
    finally:
        S.bind_arrays(None)
        g.close()


def test_walls_on_the_outer_lines_are_refused_with_their_own_code():
    """The reference's layout with the walls on grid lines 0 and n-1 leaves
    outer-layer velocities without a boundary equation, whose rows would
    read nodes outside the grid: stk_prepare refuses it with STK_EBCFORM."""
    n = 13
    X = Y = Z = n - 1
    lines = [(1, 0, 0, 0, Y - 1, 0, Z - 1, 0.0, 0.0, 0, 0, 0), (1, X, X, 0, Y - 1, 0, Z - 1, 0.0, 0.0, 0, 0, 0),
             (2, 0, X - 1, 0, 0, 0, Z - 1, 0.0, 0.0, 0, 0, 0), (2, 0, X - 1, Y, Y, 0, Z - 1, 0.0, 0.0, 0, 0, 0),
             (3, 0, X - 1, 0, Y - 1, 0, 0, 0.0, 0.0, 0, 0, 0), (3, 0, X - 1, 0, Y - 1, Z, Z, 0.0, 0.0, 0, 0, 0),
             (0, 2, 2, 2, 2, 2, 2, 0.0, 0.0, 0, 0, 0)]
    pb = setups.sinker(n, setups.coarsest_levels(n))
    pb.bc = lines
    g = S.Stokes(pb.gx, pb.gy, pb.gz, pb.levels)
    try:
        k = pb.knobs
        g.knobs(k["p0koef"], k["p1koef"], k["p2koef"], k["v0koef"], k["v1koef"], k["nubeg"], k["nuend"],
                k["viscmod"], pb.gravity)
        g.bc_lines(pb.bc)
        g.markers_to_nodes(pb.mx, pb.my, pb.mz, pb.mnu, pb.mro)
        assert S.lib().stk_prepare(g.h) == -9
    finally:
        g.close()


def test_solve_with_the_reference_gauge_row(oracle_mod):
    """The reference's pressure gauge line (P 2 2 2 2 2 2 0 0,
    parallel/init_inclusion.t3c:53) kept: the solve still reaches 1e-8 by
    the oracle and the gauge cell holds its constant."""
    pb = setups.sinker(N, setups.coarsest_levels(N), eta_in=1e22, eta_out=1e20)
    pb.bc = setups.free_slip_lines(N, N, N, gauge=(2, 2, 2))
    g = S.from_problem(pb)
    try:
        g.hydrostatic_guess()
        st = g.solve(precision=0, tol=1e-8)
        assert st.status == 0
        x = g.solution()
        o = oracle_mod.from_problem(pb)
        o.arr("sol0")[: x.size] = x
        rel, _ = o.relres()
        _say(f"solve with the gauge row: {st.iterations} iterations, oracle relres {rel:.2e}")
        assert rel <= 1.0000001e-8
        assert abs(x[4 * (2 * N * N + 2 * N + 2)]) <= 1e-9 * np.abs(x[0::4]).max()
    finally:
        g.close()
