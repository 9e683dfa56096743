"""The oracle's post-solve pass pinned by closed forms: the stress and
strain-rate recompute (parallel/move3mg.c:2908-3046) and the reference's
convergence norms stoksum/stoksum1/stoksum2/contsum/contsumb/bondsum
(parallel/move3mg.c:3090-3238, stop test 3282).

- Shear pairs (velocity component c varying as a square along axis a != c,
  edge viscosity of the (c, a) plane layered along a): the discrete strain
  rate on edge k is exactly the grid coordinate g[k], the stress 2 eta_k g[k],
  and the momentum error of a c-row between edges k and k+1 is
  (2 eta_{k+1} g[k+1] - 2 eta_k g[k]) / (g[k+1] - g[k]).
- Normal pairs (c varying as a square along its own axis, cell viscosity
  layered along it): strain rate 4/3 of the cell-centre coordinate, stress
  8/3 eta of it, and div v = twice the cell centre.
- Each check fails when the layer is read one edge or cell off.
- The hydrostatic state has zero momentum and continuity errors; a single
  violated boundary equation gives bondsum = 1/sqrt(rows); the 12-node mean
  of the error scale weighs the node lines 1:10:1 across the row.
"""
import numpy as np
import pytest

from pardiso2_amd import setups

N = 13
SIZES = (1.0, 0.8, 1.3)
EDGE = {frozenset((0, 1)): "nxy", frozenset((0, 2)): "nxz", frozenset((1, 2)): "nyz"}
SIG = {frozenset((0, 1)): 3, frozenset((0, 2)): 4, frozenset((1, 2)): 5}


def _layers(m, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(1, 200, m).astype(np.float64)


def _box(oracle_mod, gravity=(0.0, 0.0, 0.0), viscmod=0):
    g = [setups.box_grid(N, s) for s in SIZES]
    o = oracle_mod.Oracle(g[0], g[1], g[2], 0)
    o.knobs(p0koef=1.0, p1koef=1.0, p2koef=0.0, nubeg=1e-3, nuend=1e3, viscmod=viscmod, gravity=gravity)
    o.bc_lines(setups.free_slip_lines(N, N, N, gauge=None))
    n0 = o.info[0][4]
    for k in ("nu", "nxx", "nxy", "nxz", "nyz"):
        o.arr(k)[:n0] = 1.0
    o.arr("ro")[:n0] = 0.0
    return o, g


def _idx():
    m3, m1, m2 = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return m1.ravel(), m2.ravel(), m3.ravel()


def _centres(g):
    """Staggered positions along one axis: the mid-point of line k and k+1
    (the last one past the grid, never read by a row in range)."""
    return np.append(0.5 * (g[:-1] + g[1:]), g[-1] + 0.5 * (g[-1] - g[-2]))


def _eq_rows(o, c):
    """Momentum rows of component c the reference's error loop visits."""
    n0 = o.info[0][4]
    m = _idx()
    inr = [None, (m[1] < N - 1) & (m[2] < N - 1), (m[0] < N - 1) & (m[2] < N - 1), (m[0] < N - 1) & (m[1] < N - 1)][c]
    return inr & (o.arr("bondm")[c:4 * n0:4] == 0)


def _shear_state(o, g, c, a, eta):
    n0 = o.info[0][4]
    m = _idx()
    o.arr(EDGE[frozenset((c - 1, a))])[:n0] = eta[m[a]]
    x = np.zeros(4 * n0)
    x[c::4] = _centres(g[a])[m[a]] ** 2
    return x


def _skf1(g):
    h = [gg[1] - gg[0] for gg in g]
    return (h[0] ** 2 + h[1] ** 2 + h[2] ** 2) / 3.0


@pytest.mark.parametrize("c,a", [(1, 1), (1, 2), (2, 0), (2, 2), (3, 0), (3, 1)])
def test_shear_stress_and_momentum_error_closed_form(oracle_mod, c, a):
    o, g = _box(oracle_mod)
    eta = _layers(N, 10 * c + a)
    x = _shear_state(o, g, c, a, eta)
    sig, eps, out = o.post(x)
    m = _idx()
    ka = m[a]
    ga = g[a]
    k = SIG[frozenset((c - 1, a))]
    # the strain rate on each computed edge is the grid coordinate of the edge
    # computed edges: lines 1 .. N-2 of both in-plane axes, below the last line of the third
    rng_ok = np.ones(N ** 3, bool)
    other = ({0, 1, 2} - {c - 1, a}).pop()
    for ax in (c - 1, a):
        rng_ok &= (m[ax] >= 1) & (m[ax] <= N - 2)
    rng_ok &= m[other] <= N - 2
    np.testing.assert_allclose(eps[k][rng_ok], ga[ka[rng_ok]], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(sig[k][rng_ok], 2.0 * eta[ka[rng_ok]] * ga[ka[rng_ok]], rtol=1e-13, atol=1e-12)
    assert not np.any(sig[k][~rng_ok])
    for q in range(6):
        if q != k:
            assert not np.any(sig[q]), q
    # momentum errors: only the c-rows, from the two edges around each row
    rows = _eq_rows(o, c)
    kk = ka[rows]
    err = (2 * eta[kk + 1] * ga[kk + 1] - 2 * eta[kk] * ga[kk]) / (ga[kk + 1] - ga[kk])
    nrows = sum(int(_eq_rows(o, cc).sum()) for cc in (1, 2, 3))
    assert out[3] == nrows
    np.testing.assert_allclose(out[0], np.sum(err ** 2), rtol=1e-12)
    np.testing.assert_allclose(out[1], np.sum((err * _skf1(g)) ** 2), rtol=1e-12)
    np.testing.assert_allclose(out[2], np.sum((err * _skf1(g)) ** 2 * err ** 2), rtol=1e-12)
    # the layer read one edge off does not reproduce the error
    e1 = (2 * eta[kk] * ga[kk + 1] - 2 * eta[kk - 1] * ga[kk]) / (ga[kk + 1] - ga[kk])
    assert abs(np.sum(e1 ** 2) - out[0]) > 1e-3 * out[0]
    # no divergence, no continuity error
    assert out[4] == 0.0 and out[6] == 0.0
    norms = oracle_mod.post_norms(out)
    v = x[c::4][[None, (m[1] < N - 1) & (m[2] < N - 1), (m[0] < N - 1) & (m[2] < N - 1),
                 (m[0] < N - 1) & (m[1] < N - 1)][c]]
    mv = (v.max() - v.min()) / np.sqrt(3.0)
    np.testing.assert_allclose(norms["maxvxyz"], mv, rtol=1e-14)
    np.testing.assert_allclose(norms["stoksum"], np.sqrt(np.sum(err ** 2) / nrows), rtol=1e-12)
    np.testing.assert_allclose(norms["stoksum1"], np.sqrt(np.sum((err * _skf1(g)) ** 2) / nrows) / mv, rtol=1e-12)


@pytest.mark.parametrize("c", [1, 2, 3])
def test_normal_stress_momentum_and_continuity_closed_form(oracle_mod, c):
    a = c - 1
    o, g = _box(oracle_mod)
    n0 = o.info[0][4]
    eta = _layers(N, 40 + c)
    m = _idx()
    ka = m[a]
    ga = g[a]
    o.arr("nxx")[:n0] = eta[ka]
    x = np.zeros(4 * n0)
    x[c::4] = ga[ka] ** 2
    sig, eps, out = o.post(x)
    cells = (m[0] >= 1) & (m[1] >= 1) & (m[2] >= 1)
    gc = np.zeros(N)
    gc[1:] = 0.5 * (ga[:-1] + ga[1:])   # cell centre of cell k (lines k-1, k)
    np.testing.assert_allclose(eps[a][cells], 4.0 / 3.0 * gc[ka[cells]], rtol=1e-13)
    np.testing.assert_allclose(sig[a][cells], 8.0 / 3.0 * eta[ka[cells]] * gc[ka[cells]], rtol=1e-13)
    for b in {0, 1, 2} - {a}:
        np.testing.assert_allclose(eps[b][cells], -2.0 / 3.0 * gc[ka[cells]], rtol=1e-13)
    assert not np.any(sig[3:])
    # nu extrema: the clamped cell viscosities of the normal stresses
    assert out[18] == eta[1:].min() and out[19] == eta[1:].max()
    rows = _eq_rows(o, c)
    kk = ka[rows]
    h = ga[1] - ga[0]
    err = 8.0 / 3.0 * (eta[kk + 1] * gc[kk + 1] - eta[kk] * gc[kk]) / h
    np.testing.assert_allclose(out[0], np.sum(err ** 2), rtol=1e-11)
    e1 = 8.0 / 3.0 * (eta[kk] * gc[kk + 1] - eta[kk - 1] * gc[kk]) / h
    assert abs(np.sum(e1 ** 2) - out[0]) > 1e-3 * out[0]
    # continuity: div v = 2 * cell centre, scaled by the mean spacing
    hs = [gg[1] - gg[0] for gg in g]
    skf = (hs[0] + hs[1] + hs[2]) / 3.0
    pbc = o.arr("bondm")[0:4 * n0:4] != 0
    cv = 2.0 * gc[ka] * skf
    np.testing.assert_allclose(out[4], np.sum(cv[cells & ~pbc] ** 2), rtol=1e-12)
    np.testing.assert_allclose(out[6], np.sum(cv[cells & pbc] ** 2), rtol=1e-12)
    assert out[5] == np.sum(cells & ~pbc) and out[7] == np.sum(cells & pbc)
    norms = oracle_mod.post_norms(out)
    # the velocity range runs over the rows in range: vx on every x line, vy/vz up to line N-2
    v = x[c::4][[None, (m[1] < N - 1) & (m[2] < N - 1), (m[0] < N - 1) & (m[2] < N - 1),
                 (m[0] < N - 1) & (m[1] < N - 1)][c]]
    mv = (v.max() - v.min()) / np.sqrt(3.0)
    np.testing.assert_allclose(norms["contsum"], np.sqrt(np.sum(cv[cells & ~pbc] ** 2) / np.sum(cells & ~pbc)) / mv,
                               rtol=1e-12)


def test_hydrostatic_state_has_zero_errors(oracle_mod):
    gz = -9.81
    rho = 3300.0
    o, g = _box(oracle_mod, gravity=(0.0, 0.0, gz))
    n0 = o.info[0][4]
    o.arr("ro")[:n0] = rho
    m = _idx()
    zc = np.zeros(N)
    zc[1:] = 0.5 * (g[2][:-1] + g[2][1:])
    x = np.zeros(4 * n0)
    x[0::4] = gz * rho * zc[m[2]]
    _, _, out = o.post(x)
    assert out[0] <= (1e-12 * rho * abs(gz)) ** 2 * out[3]
    assert out[4] == 0.0 and out[6] == 0.0 and out[8] == 0.0
    norms = oracle_mod.post_norms(out)
    assert norms["maxvxyz"] == 1e-8          # no velocity: the reference's floor
    assert oracle_mod.post_norms(out, divvmax=0.5)["maxvxyz"] == 0.5
    # the gravity term alone (pressure left at zero) is the whole error of every z-row
    x[0::4] = 0.0
    _, _, out0 = o.post(x)
    nz = int(_eq_rows(o, 3).sum())
    np.testing.assert_allclose(out0[0], nz * (gz * rho) ** 2, rtol=1e-13)


def test_bondsum_of_one_violated_boundary_equation(oracle_mod):
    o, g = _box(oracle_mod)
    n0 = o.info[0][4]
    x = np.zeros(4 * n0)
    # vx ghost row (m1=4, m2=0, m3=5): a copy of its interior neighbour
    r = 4 * (5 * N * N + 4 * N + 0) + 1
    assert o.arr("bondm")[r] != 0
    x[r] = 1.0
    _, _, out = o.post(x)
    m = _idx()
    nb = 0
    for c, inr in ((1, (m[1] < N - 1) & (m[2] < N - 1)), (2, (m[0] < N - 1) & (m[2] < N - 1)),
                   (3, (m[0] < N - 1) & (m[1] < N - 1))):
        nb += int(np.sum(inr & (o.arr("bondm")[c:4 * n0:4] != 0)))
    assert out[9] == nb
    assert out[8] == 1.0
    np.testing.assert_allclose(oracle_mod.post_norms(out)["bondsum"], np.sqrt(1.0 / nb), rtol=1e-15)


@pytest.mark.parametrize("viscmod", [0, 1, 2])
def test_error_scale_weighs_node_lines_one_ten_one(oracle_mod, viscmod):
    """x-rows of the (vx, y) shear state with the node viscosity layered in
    x: the 12-node mean of xstokserr (move3mg.c:867-869) is (c + 10a + b)/12
    of the lines m1-1, m1, m1+1 (geometric / harmonic for viscmod 1 / 2)."""
    o, g = _box(oracle_mod, viscmod=viscmod)
    n0 = o.info[0][4]
    m = _idx()
    nu = _layers(N, 77)
    o.arr("nu")[:n0] = nu[m[0]]
    eta = np.ones(N)
    x = _shear_state(o, g, 1, 1, eta)
    _, _, out = o.post(x)
    rows = _eq_rows(o, 1)
    k1 = m[0][rows]
    a, lo, hi = nu[k1], nu[k1 - 1], nu[k1 + 1]
    if viscmod == 0:
        ne = (lo + 10 * a + hi) / 12.0
    elif viscmod == 1:
        ne = np.exp((np.log(lo) + 10 * np.log(a) + np.log(hi)) / 12.0)
    else:
        ne = 12.0 / (1 / lo + 10 / a + 1 / hi)
    kk = m[1][rows]
    gy = g[1]
    err = (2 * gy[kk + 1] - 2 * gy[kk]) / (gy[kk + 1] - gy[kk])
    h = [gg[1] - gg[0] for gg in g]
    s1 = (h[0] ** 2 + h[1] ** 2 + h[2] ** 2) / 3.0 / ne
    np.testing.assert_allclose(out[1], np.sum((err * s1) ** 2), rtol=1e-12)
    # an unweighted 4-node mean of line m1 alone would not match
    assert abs(np.sum((err * (h[0] ** 2 + h[1] ** 2 + h[2] ** 2) / 3.0 / a) ** 2) - out[1]) > 1e-3 * out[1]
