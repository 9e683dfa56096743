"""Device kernels against the CPU oracle on the same inputs, through the C ABI.

Bit-exact where the reference's arithmetic order is sequential and the
device repeats it (marker averaging, viscosity restriction, FP64 residual,
coloured smoother sweep, restriction, prolongation); within 1e-12 relative
otherwise.
"""
import numpy as np
import pytest

from pardiso2_amd import _lib as S
from pardiso2_amd import setups

pytestmark = pytest.mark.gpu


def _pair(oracle_mod, n=17, L=2, p2=0.1, eta_in=1e21, viscmod=0, seed=1):
    pb = setups.sinker(n, L, eta_in=eta_in, eta_out=1e20, seed=seed)
    pb.bc = setups.free_slip_lines(n, n, n, gauge=None)
    pb.knobs.update(dict(p2koef=p2, viscmod=viscmod))
    o = oracle_mod.from_problem(pb)
    g = S.from_problem(pb)
    return pb, o, g


@pytest.fixture(scope="module")
def pair(oracle_mod):
    pb, o, g = _pair(oracle_mod)
    yield pb, o, g
    g.close()


def test_marker_averaging_bit_exact(pair):
    pb, o, g = pair
    f = g.fields()
    for k in ("nu", "ro", "nxx", "nxy", "nxz", "nyz"):
        a = f[k]
        np.testing.assert_array_equal(a, o.arr(k)[: a.size], err_msg=k)


def test_viscosity_restriction_bit_exact_and_nd(pair):
    pb, o, g = pair
    f = g.fields()
    n0 = o.info[0][4]
    np.testing.assert_array_equal(f["nu"][n0:], o.arr("nu")[n0:])
    np.testing.assert_array_equal(f["ro"][n0:], o.arr("ro")[n0:])
    # nd (cell viscosity of the continuity rows) on every continuity row
    num0, bondm = o.arr("num0"), o.arr("bondm")
    cont = np.where((num0[0::4] > 0) & (bondm[0::4] == 0))[0]
    np.testing.assert_array_equal(f["nd"][cont], o.arr("nd")[cont])


def test_fp64_residual_matches_oracle(pair):
    pb, o, g = pair
    n4 = 4 * o.info[0][4]
    rng = np.random.default_rng(0)
    x = rng.standard_normal(n4) * 1e-15
    o.arr("sol0")[:n4] = x
    o.residual(0)
    r_or = o.arr("fre0")[:n4].copy()
    g.vec_set(S.SOL0, 0, x)
    g.vec_set(S.FRE1, 0, o.arr("fre1")[:n4].copy())
    g.smooth(0, 0)
    r_g = g.vec_get(S.FRE0, 0)
    rows = o.arr("num0")[:n4] > 0
    np.testing.assert_allclose(r_g[rows], r_or[rows], rtol=1e-12, atol=1e-12 * np.abs(r_or).max())
    assert np.mean(r_g[rows] == r_or[rows]) > 0.999  # same summation order: bit-identical


def test_operator_apply_matches_assembled_rows(pair):
    pb, o, g = pair
    n4 = 4 * o.info[0][4]
    rng = np.random.default_rng(2)
    x = rng.standard_normal(n4)
    y = g.apply(0, x)
    num0, pos0, lin0, val1 = o.arr("num0"), o.arr("pos0"), o.arr("lin0"), o.arr("val1")
    yo = np.zeros(n4)
    for r in range(n4):
        k = num0[r]
        if k:
            p = pos0[r]
            yo[r] = val1[p:p + k] @ x[lin0[p:p + k]]
    np.testing.assert_allclose(y, yo, rtol=1e-12, atol=1e-12 * np.abs(yo).max())


@pytest.mark.parametrize("tile", [(32, 4, 4), (8, 2, 3), (8, 2, -3), (16, 4, -2)])
def test_coloured_sweep_matches_oracle(oracle_mod, tile):
    pb, o, g = _pair(oracle_mod, seed=3)
    try:
        n4 = 4 * o.info[0][4]
        rng = np.random.default_rng(4)
        x = rng.standard_normal(n4) * 1e-16
        o.arr("sol0")[:n4] = x
        g.vec_set(S.SOL0, 0, x)
        g.vec_set(S.FRE1, 0, o.arr("fre1")[:n4].copy())
        o.sweep_tiles(3, 0, tile[0], tile[1], abs(tile[2]))
        g.smooth(0, 3, tile=tile)
        if tile == (32, 4, 4):  # this case runs the boundary chains inside the tiles
            assert g.tile_chains(0, tile)
        xo = o.arr("sol0")[:n4]
        xg = g.vec_get(S.SOL0, 0)
        np.testing.assert_allclose(xg, xo, rtol=1e-12, atol=1e-12 * np.abs(xo).max())
        fo = o.arr("fre0")[:n4]
        fg = g.vec_get(S.FRE0, 0)
        rows = o.arr("num0")[:n4] > 0
        np.testing.assert_allclose(fg[rows], fo[rows], rtol=1e-10, atol=1e-10 * np.abs(fo).max())
    finally:
        g.close()


def test_sweep_boundary_rows_both_ways(oracle_mod):
    """FP64 tile sweeps equal the oracle's both when the tiles set the boundary
    rows themselves (chains composed once per call, rows written in LDS or
    stored directly) and when a separate chain pass follows every sweep (some
    chain's row and partner fall in different tiles).  27^3, 3 levels: both
    kinds occur among these tiles and levels."""
    pb, o, g = _pair(oracle_mod, n=27, L=3, seed=6)
    try:
        cands = [(32, 4, 4), (8, 2, 3), (16, 4, 2), (4, 4, 4), (3, 3, 3), (5, 2, 7), (2, 2, 2), (7, 3, 5)]
        found = {}
        for l in range(3):
            for t in cands:
                found.setdefault(g.tile_chains(l, t), (l, t))
        assert set(found) == {True, False}, found
        rng = np.random.default_rng(9)
        for inplace, (l, t) in sorted(found.items()):
            b4, n4 = 4 * o.info[l][5], 4 * o.info[l][4]
            x = rng.standard_normal(n4) * 1e-16
            f = rng.standard_normal(n4) * (o.arr("num0")[b4:b4 + n4] > 0)
            o.arr("sol0")[b4:b4 + n4] = x
            o.arr("fre1")[b4:b4 + n4] = f
            g.vec_set(S.SOL0, l, x)
            g.vec_set(S.FRE1, l, f)
            o.sweep_tiles(3, l, t[0], t[1], t[2])
            g.smooth(l, 3, tile=t, precision=1)
            xo = o.arr("sol0")[b4:b4 + n4]
            np.testing.assert_allclose(g.vec_get(S.SOL0, l), xo, rtol=1e-12, atol=1e-12 * np.abs(xo).max(),
                                       err_msg=f"level {l} tile {t} in-tile chains {inplace}")
    finally:
        g.close()


@pytest.mark.parametrize("tile,vonly", [((32, 4, 4), False), ((8, 2, 3), False), ((16, 4, -2), False),
                                        ((0, 0, 0), False), ((0, 0, 0), True), ((32, 4, 4), True)])
def test_fp32_sweep_matches_fp64_oracle_to_fp32_accuracy(oracle_mod, tile, vonly):
    """The FP32 smoother of the mixed-precision V-cycle (fused multiply-adds,
    4/3 and 2/3 as constants, hardware reciprocal of the diagonal) makes the
    oracle's FP64 row updates to FP32 accuracy: max error <= 2e-5 of max |x|
    after 3 sweeps.  A wrong coefficient or slot is an O(1) error.  Tile
    (0, 0, 0) is the LDS-resident whole-level kernel (level 1, 11^3 nodes);
    velocity-only on it is the V-cycle's kernel with precomputed rows and
    composed boundary chains."""
    pb, o, g = _pair(oracle_mod, seed=3)
    o.set_vonly(vonly)
    try:
        l = 1 if tile == (0, 0, 0) else 0
        X, Y, Z = o.info[l][:3]
        b4, n4 = 4 * o.info[l][5], 4 * o.info[l][4]
        rng = np.random.default_rng(8)
        x = rng.standard_normal(n4) * 1e-16
        f = rng.standard_normal(n4) * (o.arr("num0")[b4:b4 + n4] > 0)
        o.arr("sol0")[b4:b4 + n4] = x
        o.arr("fre1")[b4:b4 + n4] = f
        g.vec_set(S.SOL0, l, x)
        g.vec_set(S.FRE1, l, f)
        if tile == (0, 0, 0):
            o.sweep_tiles(3, l, Y, X, Z)
        else:
            o.sweep_tiles(3, l, tile[0], tile[1], abs(tile[2]))
        g.smooth(l, 3, tile=tile, precision=0, velocity_only=vonly)
        xo = o.arr("sol0")[b4:b4 + n4]
        xg = g.vec_get(S.SOL0, l)
        err = np.abs(xg - xo).max() / np.abs(xo).max()
        assert err <= 2e-5, err
        assert err > 0  # the FP32 path really ran in FP32
    finally:
        g.close()


def test_whole_level_sweep_matches_oracle(oracle_mod):
    """The LDS-resident whole-level smoother equals the tile sweep with one
    tile covering the level (level 1 of a 17^3 hierarchy: 11^3 nodes)."""
    pb, o, g = _pair(oracle_mod, seed=5)
    try:
        l = 1
        X, Y, Z = o.info[l][:3]
        b4, n4 = 4 * o.info[l][5], 4 * o.info[l][4]
        rng = np.random.default_rng(6)
        x = rng.standard_normal(n4) * 1e-16
        f = rng.standard_normal(n4) * (o.arr("num0")[b4:b4 + n4] > 0)
        o.arr("sol0")[b4:b4 + n4] = x
        o.arr("fre1")[b4:b4 + n4] = f
        g.vec_set(S.SOL0, l, x)
        g.vec_set(S.FRE1, l, f)
        o.sweep_tiles(3, l, Y, X, Z)
        g.smooth(l, 3, tile=(0, 0, 0))
        xo = o.arr("sol0")[b4:b4 + n4]
        np.testing.assert_allclose(g.vec_get(S.SOL0, l), xo, rtol=1e-12, atol=1e-12 * np.abs(xo).max())
    finally:
        g.close()


def test_restriction_and_prolongation_bit_exact(pair):
    pb, o, g = pair
    g.bind()
    L = S.lib()
    n4 = 4 * o.info[0][4]
    rng = np.random.default_rng(5)
    r = rng.standard_normal(n4)
    o.arr("fre0")[:n4] = r
    g.vec_set(S.FRE0, 0, r)
    o.vprestrict(1)
    L.vprestrict(1)
    b1, n41 = 4 * o.info[1][5], 4 * o.info[1][4]
    np.testing.assert_array_equal(g.vec_get(S.FRE1, 1), o.arr("fre1")[b1:b1 + n41])
    xc = rng.standard_normal(n41)
    xf = rng.standard_normal(n4)
    o.arr("sol0")[b1:b1 + n41] = xc
    o.arr("sol0")[:n4] = xf
    g.vec_set(S.SOL0, 1, xc)
    g.vec_set(S.SOL0, 0, xf)
    o.vpprolong(1, 0)
    L.vpprolong(1, 0)
    np.testing.assert_array_equal(g.vec_get(S.SOL0, 0), o.arr("sol0")[:n4])


def test_restriction_bit_exact_with_five_entry_gathers(oracle_mod):
    """27^3 -> 16^3 -> 10^3 -> 7^3: on the last restriction a coarse index
    gathers 5 fine indices on some axes (the 6-wide padded kernel); the
    others use the 4-wide one."""
    pb, o, g = _pair(oracle_mod, n=27, L=3, seed=6)
    try:
        g.bind()
        rng = np.random.default_rng(9)
        for l in (1, 2, 3):
            bf, nf4 = 4 * o.info[l - 1][5], 4 * o.info[l - 1][4]
            r = rng.standard_normal(nf4)
            o.arr("fre0")[bf:bf + nf4] = r
            g.vec_set(S.FRE0, l - 1, r)
            o.vprestrict(l)
            S.lib().vprestrict(l)
            bc, nc4 = 4 * o.info[l][5], 4 * o.info[l][4]
            np.testing.assert_array_equal(g.vec_get(S.FRE1, l), o.arr("fre1")[bc:bc + nc4], err_msg=f"level {l}")
    finally:
        g.close()


@pytest.mark.parametrize("n,L", [(17, 2), (27, 3)])
def test_fp32_velocity_restriction_matches_oracle(oracle_mod, n, L):
    """The V-cycle's FP32 restriction (axis by axis) equals the oracle's
    vprestrict on every velocity row to FP32 accuracy (max error <= 1e-5 of
    max |b|), pressure rows 0; 4- and 5-entry gathers."""
    pb, o, g = _pair(oracle_mod, n=n, L=L, seed=7)
    try:
        g.bind()
        rng = np.random.default_rng(11)
        for l in range(1, L + 1):
            bf, nf4 = 4 * o.info[l - 1][5], 4 * o.info[l - 1][4]
            r = rng.standard_normal(nf4)
            o.arr("fre0")[bf:bf + nf4] = r
            g.vec_set(S.FRE0, l - 1, r)
            o.vprestrict(l)
            g.vrestrict32(l)
            bc, nc4 = 4 * o.info[l][5], 4 * o.info[l][4]
            bo = o.arr("fre1")[bc:bc + nc4].copy()
            bg = g.vec_get(S.FRE1, l)
            v = np.arange(nc4) % 4 != 0
            assert not np.any(bg[~v])
            err = np.abs(bg[v] - bo[v]).max() / np.abs(bo[v]).max()
            assert err <= 1e-5, (l, err)
    finally:
        g.close()


def test_ronurestrict_entry_point(pair):
    pb, o, g = pair
    g.bind()
    before = g.fields()["nu"].copy()
    S.lib().ronurestrict(1)
    after = g.fields()["nu"]
    np.testing.assert_array_equal(after, before)  # recomputing level 1 from level 0 is idempotent


def test_coarse_solve_matches_dense_lu(oracle_mod):
    n = 13
    pb = setups.sinker(n, 1, eta_in=1e21, eta_out=1e20)
    pb.bc = setups.free_slip_lines(n, n, n, gauge=(2, 2, 2))
    o = oracle_mod.from_problem(pb)
    g = S.from_problem(pb)
    try:
        L = pb.levels
        bL, nL = 4 * o.info[L][5], 4 * o.info[L][4]
        rng = np.random.default_rng(7)
        rhs = rng.standard_normal(nL) * (o.arr("num0")[bL:bL + nL] > 0)
        o.arr("fre1")[bL:bL + nL] = rhs
        o.coarse_solve()
        g.vec_set(S.FRE1, L, rhs)
        g.coarse_solve()
        xo = o.arr("sol0")[bL:bL + nL]
        xg = g.vec_get(S.SOL0, L)
        np.testing.assert_allclose(xg, xo, rtol=1e-9, atol=1e-9 * np.abs(xo).max())
    finally:
        g.close()


def test_pardisomat_entry_point(oracle_mod):
    """Rows handed over in the reference's un1/ui1 format (pardisomat(1,..)),
    then pardisomat(0,..) solves the coarsest system on the device and stores
    the solution at sol0[4*(mgp[L]+node)+c]."""
    import ctypes as C
    n = 13
    pb = setups.sinker(n, 1, eta_in=1e21, eta_out=1e20)
    pb.bc = setups.free_slip_lines(n, n, n, gauge=(2, 2, 2))
    o = oracle_mod.from_problem(pb)
    g = S.from_problem(pb)
    try:
        g.bind()
        L = pb.levels
        bL, nL = 4 * o.info[L][5], 4 * o.info[L][4]
        num0 = o.arr("num0")
        un1 = (C.c_long * 64)()
        ui1 = (C.c_double * 64)()
        rng = np.random.default_rng(8)
        rhs = rng.standard_normal(nL)
        for i in range(nL):
            r = bL + i
            if num0[r] == 0:
                un1[0], un1[1], ui1[0], ui1[1] = 1, i, 0.0, 1.0
                rhs[i] = 0.0
            else:
                idx, val, _ = o.row(r)
                un1[0] = len(idx)
                for k, (j, v) in enumerate(zip(idx - bL, val)):
                    un1[k + 1] = int(j)
                    ui1[k + 1] = float(v)
                ui1[0] = rhs[i]
            assert S.lib().pardisomat(1, i, r, 0, 0, un1, ui1) == 0
        assert S.lib().pardisomat(0, nL, 0, 0, 0, un1, ui1) == 0
        o.arr("fre1")[bL:bL + nL] = rhs
        o.coarse_solve()
        np.testing.assert_allclose(g.vec_get(S.SOL0, L), o.arr("sol0")[bL:bL + nL], rtol=1e-9,
                                   atol=1e-9 * np.abs(o.arr("sol0")[bL:bL + nL]).max())
    finally:
        g.close()


def test_pardisomat_identity_and_singular_system():
    """The device Gauss-Jordan behind pardisomat(0,..): an identity system
    comes back exactly (pivots 1, no factors), and the same system with one
    empty row is reported as singular (STK_ESINGULAR, -4) instead of solved.
    pardisomat(1,..) sets a row as the reference does (a second assembly
    replaces the first), so a later valid system solves to its own values."""
    import ctypes as C
    n = 13
    pb = setups.sinker(n, 1, eta_in=1e21, eta_out=1e20)
    pb.bc = setups.free_slip_lines(n, n, n, gauge=(2, 2, 2))
    g = S.from_problem(pb)
    try:
        g.bind()
        L = pb.levels
        nL = 4 * g.info[L][4]
        un1 = (C.c_long * 4)()
        ui1 = (C.c_double * 4)()
        rhs = np.random.default_rng(3).standard_normal(nL)

        def hand_over(empty):
            for i in range(nL):
                un1[0] = 0 if i == empty else 1
                un1[1], ui1[0], ui1[1] = i, rhs[i], 1.0
                assert S.lib().pardisomat(1, i, 0, 0, 0, un1, ui1) == 0

        hand_over(empty=-1)
        assert S.lib().pardisomat(0, nL, 0, 0, 0, un1, ui1) == 0
        np.testing.assert_array_equal(g.vec_get(S.SOL0, L), rhs)
        hand_over(empty=nL // 3)
        assert S.lib().pardisomat(0, nL, 0, 0, 0, un1, ui1) == -4
        rhs *= 2.0
        hand_over(empty=-1)
        assert S.lib().pardisomat(0, nL, 0, 0, 0, un1, ui1) == 0
        np.testing.assert_array_equal(g.vec_get(S.SOL0, L), rhs)
    finally:
        g.close()


def test_small_dense_elimination_entry_points():
    import ctypes as C
    L = S.lib()
    a = np.array([[4.0, 1.0, 0.0, 1.0], [2.0, 5.0, 1.0, 2.0], [0.0, 1.0, 3.0, 3.0]])  # 3 x (3+1)
    m = np.ascontiguousarray(a.copy())
    x = np.zeros(3)
    L.gausmat(3, m.ctypes.data_as(C.POINTER(C.c_double)), x.ctypes.data_as(C.POINTER(C.c_double)))
    np.testing.assert_allclose(a[:, :3] @ x, a[:, 3], rtol=1e-13)
    L.matclear(3, m.ctypes.data_as(C.POINTER(C.c_double)))
    assert not m.any()
