"""Oracle checked against what the mathematics fixes: closed forms, symmetry,
duality, invariants of the transfers, exact-solution fixed points, and
agreement of the direct and iterative solves."""
import numpy as np
import pytest
import scipy.linalg as sla

from pardiso2_amd import setups


def _aniso_grid(n, sizes=(1.0, 0.7, 1.3)):
    return [setups.uniform_grid(n, s) for s in sizes]


def _assembled_dense(o, level=0):
    info = o.info[level]
    base = 4 * info[5]
    n4 = 4 * info[4]
    A = np.zeros((n4, n4))
    b = np.zeros(n4)
    num0 = o.arr("num0")
    for i in range(n4):
        r = base + i
        if num0[r] == 0:
            A[i, i] = 1.0
            continue
        idx, val, rhs = o.row(r)
        for j, v in zip(idx - base, val):
            A[i, j] += v
        b[i] = rhs
    return A, b


def test_level_sizes_and_coarse_lines(oracle_mod):
    assert setups.levels_for(33, 3) == [33, 19, 12, 8]
    assert setups.levels_for(129, 5) == [129, 67, 36, 20, 12, 8]
    assert setups.levels_for(257, 6) == [257, 131, 68, 36, 20, 12, 8]
    g = setups.uniform_grid(33)
    o = oracle_mod.Oracle(g, g, g, 3)
    for l, step in ((1, 2), (2, 4), (3, 8)):
        c = o.grid(l, 0)
        m = len(c)
        np.testing.assert_array_equal(c[:3], g[:3])
        np.testing.assert_array_equal(c[-3:], g[-3:])
        for i in range(3, m - 3):
            assert c[i] == g[(i - 2) * step + 2]


def test_vx_diagonal_closed_form(oracle_mod):
    n = 11
    gx, gy, gz = _aniso_grid(n)
    o = oracle_mod.Oracle(gx, gy, gz, 0)
    eta = 3.0
    o.knobs(p0koef=1, p1koef=1, p2koef=0.0, nubeg=eta, nuend=eta)
    o.assemble()
    dx, dy, dz = gx[1] - gx[0], gy[1] - gy[0], gz[1] - gz[0]
    for (m1, m2, m3), c, lead in (((5, 4, 3), 1, dx), ((4, 5, 3), 2, dy), ((4, 3, 5), 3, dz)):
        r = 4 * (m3 * n * n + m1 * n + m2) + c
        idx, val, _ = o.row(r)
        assert idx[0] == r
        others = {1: (dy, dz), 2: (dx, dz), 3: (dx, dy)}[c]
        expect = -eta * (8.0 / 3.0 / lead ** 2 + 2.0 / others[0] ** 2 + 2.0 / others[1] ** 2)
        assert val[0] == pytest.approx(expect, rel=1e-13)
        # the row annihilates constants of its own component (no zero-order term)
        own = [v for j, v in zip(idx, val) if j % 4 == c]
        assert abs(sum(own)) < 1e-10 * abs(val[0])


def test_velocity_block_symmetric_and_duality(oracle_mod):
    """On a uniform grid with variable viscosity and p2koef=0 the velocity
    block is symmetric and the momentum P columns are the transposed
    continuity rows (A_vp = A_pv^T)."""
    n = 9
    g = setups.box_grid(n)
    o = oracle_mod.Oracle(g, g, g, 0)
    o.knobs(p0koef=1, p1koef=1, p2koef=0.0, nubeg=1e-3, nuend=1e3)
    o.bc_lines(setups.free_slip_lines(n, n, n, gauge=None))
    rng = np.random.default_rng(3)
    for name in ("nu", "nxx", "nxy", "nxz", "nyz"):
        a = o.arr(name)
        a[:] = np.exp(rng.uniform(-3, 3, a.size)).astype(np.float32)
    o.assemble()
    A, _ = _assembled_dense(o)
    num0 = o.arr("num0")[: A.shape[0]]
    bondm = o.arr("bondm")[: A.shape[0]]
    eq = (num0 > 0) & (bondm == 0)
    vel = eq & (np.arange(A.shape[0]) % 4 != 0)
    pres = eq & (np.arange(A.shape[0]) % 4 == 0)
    Avv = A[np.ix_(vel, vel)]
    np.testing.assert_allclose(Avv, Avv.T, rtol=1e-12, atol=1e-12 * np.abs(Avv).max())
    Avp = A[np.ix_(vel, pres)]
    Apv = A[np.ix_(pres, vel)]
    np.testing.assert_allclose(Avp, Apv.T, rtol=1e-13, atol=1e-13)


def test_dense_lu_matches_scipy(oracle_mod):
    rng = np.random.default_rng(1)
    n = 300
    a = rng.standard_normal((n, n))
    a[:, 7] *= 1e-3  # force pivoting
    b = rng.standard_normal(n)
    rc, lu, piv = oracle_mod.dense_lu(a)
    assert rc == 0
    x = oracle_mod.dense_lu_solve(lu, piv, b)
    lu2, piv2 = sla.lu_factor(a)
    np.testing.assert_array_equal(piv, piv2)
    np.testing.assert_allclose(x, sla.lu_solve((lu2, piv2), b), rtol=1e-11, atol=1e-12)
    assert np.linalg.norm(a @ x - b) / np.linalg.norm(b) < 1e-13


def _small_gauged(oracle_mod, n=9, eta_in=1e21):
    pb = setups.sinker(n, 0, eta_in=eta_in, eta_out=1e20)
    pb.bc = setups.free_slip_lines(n, n, n, gauge=(2, 2, 2))
    return pb, oracle_mod.from_problem(pb, L=0)


def test_band_solver_matches_dense_direct(oracle_mod):
    pb, o = _small_gauged(oracle_mod)
    A, b = _assembled_dense(o)
    xd = np.linalg.solve(A, b)
    rc, xb, bw = o.band_solve()
    assert rc == 0 and bw > 0
    scale = np.abs(xd).max()
    np.testing.assert_allclose(xb, xd, rtol=1e-7, atol=1e-9 * scale)


@pytest.mark.parametrize("threads", [1, 3, 4])
def test_pipelined_band_solver_is_bit_identical(oracle_mod, threads):
    # the reference's thread structure (rows round-robin, each waiting for the
    # rows it eliminates with) repeats the serial elimination exactly
    pb = setups.sinker(11, 1, eta_in=1e22, eta_out=1e20)
    pb.bc = setups.free_slip_lines(11, 11, 11, gauge=(2, 2, 2))
    o = oracle_mod.from_problem(pb)
    rc, xs, bw = o.band_solve()
    rc2, xm, bw2 = o.band_solve(threads=threads)
    assert rc == rc2 == 0 and bw == bw2
    np.testing.assert_array_equal(xm, xs)
    assert o.relres_rows(xm)[0] < 1e-12


def test_exact_solution_is_fixed_point_of_smoother(oracle_mod):
    pb, o = _small_gauged(oracle_mod)
    A, b = _assembled_dense(o)
    x = np.linalg.solve(A, b)
    n4 = x.size
    o.arr("sol0")[:n4] = x
    o.gausmat4(2, 0)
    y = o.arr("sol0")[:n4]
    np.testing.assert_allclose(y, x, rtol=1e-8, atol=1e-10 * np.abs(x).max())
    o.arr("sol0")[:n4] = x
    o.sweep_tiles(2, 0, 4, 2, 2)
    np.testing.assert_allclose(o.arr("sol0")[:n4], x, rtol=1e-8, atol=1e-10 * np.abs(x).max())


def test_restriction_and_prolongation_of_constants(oracle_mod):
    n, L = 21, 1
    g = setups.box_grid(n)
    o = oracle_mod.Oracle(g, g, g, L)
    o.knobs(p0koef=1, p1koef=1, p2koef=0, nubeg=1, nuend=1)
    o.bc_lines(setups.free_slip_lines(n, n, n, gauge=None))
    o.assemble()
    f0, fc = o.info[0], o.info[1]
    n4f, n4c = 4 * f0[4], 4 * fc[4]
    num0 = o.arr("num0")
    bondm = o.arr("bondm")
    fre0 = o.arr("fre0")
    fre0[:n4f] = 0.0
    for c in (1, 2, 3):
        fre0[c:n4f:4] = 7.5
    o.vprestrict(1)
    base = 4 * fc[5]
    fre1c = o.arr("fre1")[base:base + n4c]
    eqc = (num0[base:base + n4c] > 0) & (bondm[base:base + n4c] == 0)
    velc = eqc & (np.arange(n4c) % 4 != 0)
    np.testing.assert_allclose(fre1c[velc], 7.5, rtol=1e-12)
    # prolongation of a constant correction reproduces it wherever the fine
    # point is bracketed by its coarse stencil (weights in [0,1])
    sol0 = o.arr("sol0")
    sol0[:] = 0
    sol0[base:base + n4c] = 2.0
    o.vpprolong(1, 0)
    xf = sol0[:n4f]
    X = f0[0]
    inner = np.zeros(n4f, bool)
    for m3 in range(3, X - 3):
        for m1 in range(3, X - 3):
            for m2 in range(3, X - 3):
                nn = m3 * X * X + m1 * X + m2
                inner[4 * nn + 1:4 * nn + 4] = True
    np.testing.assert_allclose(xf[inner], 2.0, rtol=1e-12)


def test_ronurestrict_and_marker_averaging_of_constants(oracle_mod):
    n, L = 17, 2
    g = setups.box_grid(n)
    o = oracle_mod.Oracle(g, g, g, L)
    o.knobs(nubeg=1e18, nuend=1e25)
    mx, my, mz = setups.markers_regular(g, g, g, jitter=0.3, seed=5)
    o.ronurecalc(mx, my, mz, np.full(mx.size, 4e21, np.float32), np.full(mx.size, 3300.0, np.float32))
    m3, m1, m2 = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    m3, m1, m2 = m3.ravel(), m1.ravel(), m2.ravel()
    valid = {"nu": np.ones(n ** 3, bool), "nxx": (m1 > 0) & (m2 > 0) & (m3 > 0),
             "nxy": m3 < n - 1, "nxz": m2 < n - 1, "nyz": m1 < n - 1}
    for name, ok in valid.items():
        a = o.arr(name)[: n ** 3]
        assert np.allclose(a[ok], np.float32(4e21), rtol=1e-6), name
        # positions no marker reaches keep the clamp of their initial zero
        assert np.all(a[~ok] == np.float32(1e18)), name
    np.testing.assert_allclose(o.arr("ro")[: n ** 3], 3300.0, rtol=1e-6)
    o.ronurestrict_all()
    np.testing.assert_allclose(o.arr("nu")[n ** 3:], np.float32(4e21), rtol=1e-6)
    np.testing.assert_allclose(o.arr("ro")[n ** 3:], 3300.0, rtol=1e-6)


def test_gcr_solution_matches_direct_solve(oracle_mod):
    """Without the gauge row the pressure level is free; the gauge row only
    replaces a continuity equation the others imply, so the velocities agree
    and the pressures agree up to a constant."""
    n = 9
    pb, o = _small_gauged(oracle_mod, n=n, eta_in=1e22)
    A, b = _assembled_dense(o)
    xd = np.linalg.solve(A, b)
    pb2 = setups.sinker(n, 1, eta_in=1e22, eta_out=1e20)
    pb2.bc = setups.free_slip_lines(n, n, n, gauge=None)
    o2 = oracle_mod.from_problem(pb2)
    it, x, hist, _ = o2.gcr([4, 0], [4, 0], mode=1, tol=1e-11, maxit=300, restart=128)
    assert hist[-1] <= 1e-11
    vel = np.arange(x.size) % 4 != 0
    vs = np.abs(xd[vel]).max()
    np.testing.assert_allclose(x[vel], xd[vel], atol=1e-6 * vs)
    num0 = o.arr("num0")[: x.size]
    pr = (np.arange(x.size) % 4 == 0) & (num0 > 0)
    dp = x[pr] - xd[pr]
    assert np.abs(dp - dp.mean()).max() <= 1e-6 * np.abs(xd[pr]).max()


def test_equation_box_of_ghost_layer_boxes(oracle_mod):
    """With walls on lines 1 and n-2 every level's equation rows sit on nodes
    1..n-2 in each direction; a sweep whose single tile spans that box equals
    one whose single tile spans the whole level (no equation row is left out)."""
    for n in (17, 33):
        pb = setups.sinker(n, setups.coarsest_levels(n))
        o = oracle_mod.from_problem(pb)
        for l in range(pb.levels + 1):
            X, Y, Z = o.info[l][:3]
            assert o.eq_box(l) == (1, 1, 1, Y - 1, X - 1, Z - 1)
    pb = setups.sinker(17, 2, seed=5)
    o = oracle_mod.from_problem(pb)
    n4 = 4 * o.info[0][4]
    x = np.random.default_rng(3).standard_normal(n4) * 1e-15
    o.arr("sol0")[:n4] = x
    o.sweep_tiles(2, 0, 15, 15, 15)  # one tile over the box (1..15)
    y_box = o.arr("sol0")[:n4].copy()
    o.arr("sol0")[:n4] = x
    o.sweep_tiles(2, 0, 64, 64, 64)  # one tile over the whole level
    np.testing.assert_array_equal(o.arr("sol0")[:n4], y_box)


def test_tile_index_division_by_reciprocal():
    """The smoother decodes tile indices as umulhi(i, ceil(2^32/d)), exact for
    numerators below 2^16 and divisors 2 <= d < 2^16."""
    i = np.arange(1 << 16, dtype=np.uint64)
    for d in list(range(2, 140)) + [255, 256, 1000, 4097, 65535]:
        m = np.uint64(((1 << 32) + d - 1) // d)
        np.testing.assert_array_equal((i * m) >> np.uint64(32), i // np.uint64(d))
