"""The oracle's periodic parts pinned by closed forms (CPU).

- Marker folds (parallel/mark3mg.c:893-955, 1032-1111): on an unjittered
  2x2x2 lattice of a box periodic in x, the property p = x averaged to a node
  on line 0 (or its copy, line n-1) takes the markers on both sides of the
  wrap, at x and at x - period, with hat weights 3/4 and 1/4 each: the value
  is exactly half the period.  Without the fold line 0 would see only the
  markers right of it (3h/8).  Properties that sit on x lines (nu, ro, nxy,
  nxz) fold; nxx (cells) and nyz (x edges, centred between x lines) do not.
- Rows of a periodic box (parallel/move3mg.c:27-34, 400-429, 834-847):
  exactly constant viscosity in the reference's simple-shear box
  (parallel/init_inclusion.t3c) has the exact discrete solution vx linear in
  z between the held cell centres, vy = vz = P = 0, and a v_y(x) that is
  periodic in x but not constant fails it.
"""
import numpy as np
import pytest

from pardiso2_amd import setups


def _fold_oracle(oracle_mod, per):
    n = 9
    g = setups.uniform_grid(n, 1.0)
    mx, my, mz = setups.markers_regular(g, g, g, jitter=0.0)
    o = oracle_mod.Oracle(g, g.copy(), g.copy(), 1, periodic=per)
    o.knobs(nubeg=-1e30, nuend=1e30, viscmod=0)
    o.ronurecalc(mx, my, mz, mx.astype(np.float32), mx.astype(np.float32))
    return n, g, o


def _field(o, name, n):
    return o.arr(name)[: n ** 3].reshape(n, n, n)  # [m3, m1, m2]


def test_fold_gives_half_period_on_the_wrap_line(oracle_mod):
    n, g, o = _fold_oracle(oracle_mod, (1, 0, 0))
    L = g[-1] - g[0]
    for name in ("nu", "ro", "nxy", "nxz"):
        f = _field(o, name, n)
        # interior of the other two axes (their node weights are complete there)
        for m1 in (0, n - 1):
            v = f[2:-2, m1, 2:-2]
            np.testing.assert_allclose(v, L / 2, rtol=2e-6, err_msg=f"{name} line {m1}")
        # lines away from the wrap keep p = x at the node
        np.testing.assert_allclose(f[2:-2, 3, 2:-2], g[3], rtol=2e-6, err_msg=name)


def test_fold_leaves_cell_and_x_edge_properties(oracle_mod):
    n, g, o = _fold_oracle(oracle_mod, (1, 0, 0))
    h = g[1] - g[0]
    nyz = _field(o, "nyz", n)
    # nyz on line 0 averages cell 0 only: its centre h/2
    np.testing.assert_allclose(nyz[2:-2, 0, 2:-2], h / 2, rtol=2e-6)
    nxx = _field(o, "nxx", n)
    np.testing.assert_allclose(nxx[2:-2, n - 1, 2:-2], g[n - 1] - h / 2, rtol=2e-6)


def test_without_fold_the_wrap_line_differs(oracle_mod):
    # the same markers in a walled box: line 0 sees only the markers right of it
    n, g, o = _fold_oracle(oracle_mod, (0, 0, 0))
    h = g[1] - g[0]
    nu = _field(o, "nu", n)
    np.testing.assert_allclose(nu[2:-2, 0, 2:-2], 3 * h / 8, rtol=2e-6)
    assert abs(3 * h / 8 - (g[-1] - g[0]) / 2) > 0.1


def test_fold_corner_takes_all_four_sides(oracle_mod):
    # x and y periodic: nu at (0, 0) averages the four quadrants around the wrap
    n = 9
    g = setups.uniform_grid(n, 1.0)
    mx, my, mz = setups.markers_regular(g, g, g, jitter=0.0)
    o = oracle_mod.Oracle(g, g.copy(), g.copy(), 1, periodic=(1, 1, 0))
    o.knobs(nubeg=-1e30, nuend=1e30, viscmod=0)
    p = (mx + 2.0 * my).astype(np.float32)
    o.ronurecalc(mx, my, mz, p, p)
    nu = _field(o, "nu", n)
    L = 1.0
    for m1 in (0, n - 1):
        for m2 in (0, n - 1):
            np.testing.assert_allclose(nu[2:-2, m1, m2], L / 2 + 2.0 * L / 2, rtol=2e-6)


def _shear_oracle(oracle_mod, n=17):
    pb = setups.periodic_shear(n, count=0)
    N = n ** 3
    o = oracle_mod.Oracle(pb.gx, pb.gy, pb.gz, pb.levels, periodic=pb.periodic)
    k = pb.knobs
    o.knobs(k["p0koef"], k["p1koef"], 0.1, k["v0koef"], k["v1koef"], k["nubeg"], k["nuend"], k["viscmod"],
            pb.gravity)
    o.bc_lines(pb.bc)
    one = np.ones(N, np.float32)
    o.set_fields(one, one * 3000, one, one, one, one)
    o.ronurestrict_all()
    o.assemble()
    return pb, o


def _couette(pb, n):
    zc = (pb.gz[:-1] + pb.gz[1:]) / 2
    x = np.zeros(4 * n ** 3)
    for m3 in range(n - 1):
        v = 1 - 2 * (zc[m3] - zc[0]) / (zc[n - 2] - zc[0])
        for m1 in range(n):
            x[4 * ((m3 * n + m1) * n + np.arange(n - 1)) + 1] = v
    return x


def test_periodic_shear_couette_is_exact(oracle_mod):
    n = 17
    pb, o = _shear_oracle(oracle_mod, n)
    rel, _ = o.relres_rows(_couette(pb, n))
    assert rel < 1e-12


def test_periodic_shear_rejects_nonconstant_periodic_vy(oracle_mod):
    # vy = cos(2 pi x) is periodic in x but its shear stress varies: not a solution
    n = 17
    pb, o = _shear_oracle(oracle_mod, n)
    x = _couette(pb, n)
    xc = (pb.gx[:-1] + pb.gx[1:]) / 2
    for m3 in range(1, n - 2):
        for m1 in range(n - 1):
            x[4 * ((m3 * n + m1) * n + np.arange(n)) + 2] = 1e-3 * np.cos(2 * np.pi * xc[m1])
    rel, _ = o.relres_rows(x)
    assert rel > 1e-6


# ---- marker interpolation and advection across a periodic wrap (mark3mg.c:2963-3760, 337-485) ----
NP = 13


def _pgrid(n, size, seed):
    rng = np.random.default_rng(seed)
    h = rng.uniform(0.7, 1.3, n - 1)
    return np.concatenate([[0.0], np.cumsum(h / h.sum() * size)])


@pytest.fixture(scope="module")
def pbox(oracle_mod):
    g = [_pgrid(NP, s, k + 7) for k, s in enumerate((1.0, 0.8, 1.3))]
    o = oracle_mod.Oracle(g[0], g[1], g[2], 0, periodic=(1, 1, 0))
    return o, g


def _cos(p, size):
    return np.cos(2 * np.pi * p / size)


def _wrap_lin(f_last, f_first, p, c_last, c_first, d, size):
    # linear interpolation from the last centre to the first one across the wrap
    P = p + size if p < c_first else p
    t = (P - c_last) / d
    return (1 - t) * f_last + t * f_first


def test_velocity_across_the_wrap(pbox):
    o, g = pbox
    gx = g[0]
    L = gx[-1]
    n0 = NP ** 3
    cx = 0.5 * (gx[:-1] + gx[1:])  # cell centres 0..n-2
    x = np.zeros(4 * n0)
    m1 = np.tile(np.repeat(np.arange(NP), NP), NP)  # node order (m3, m1, m2)
    vy = np.where(m1 < NP - 1, _cos(cx[np.minimum(m1, NP - 2)], L), 0.0)
    x[2::4] = vy  # vy sits at x cell centres: constant in y and z here
    eps = np.zeros((6, n0))
    d = (gx[1] - gx[0] + gx[-1] - gx[-2]) / 2
    f = lambda i: float(np.float32(_cos(cx[i], L)))
    for px in (0.25 * cx[0], 0.9 * cx[0], cx[-1] + 0.3 * (L - cx[-1]), L):
        v, _ = o.interp(x, eps, (px, 0.4, 0.65))
        want = _wrap_lin(f(NP - 2), f(0), px, cx[-1], cx[0], d, L)
        assert abs(v[1] - want) < 1e-12, (px, v[1], want)
    # inside the centres: the plain bracket
    v, _ = o.interp(x, eps, (0.5 * (cx[3] + cx[4]), 0.4, 0.65))
    t = (0.5 * (cx[3] + cx[4]) - cx[3]) / (cx[4] - cx[3])
    assert abs(v[1] - ((1 - t) * f(3) + t * f(4))) < 1e-12


def test_velocity_across_the_wrap_differs_from_the_walled_clamp(oracle_mod, pbox):
    _, g = pbox
    ow = oracle_mod.Oracle(g[0], g[1], g[2], 0)
    n0 = NP ** 3
    gx = g[0]
    cx = 0.5 * (gx[:-1] + gx[1:])
    m1 = np.tile(np.repeat(np.arange(NP), NP), NP)
    x = np.zeros(4 * n0)
    x[2::4] = np.where(m1 < NP - 1, _cos(cx[np.minimum(m1, NP - 2)], gx[-1]), 0.0)
    v, _ = ow.interp(x, np.zeros((6, n0)), (0.25 * cx[0], 0.4, 0.65))
    assert abs(v[1] - float(np.float32(_cos(cx[0], gx[-1])))) < 1e-12  # walled: held at centre 0


def test_strain_rates_across_the_wrap(pbox):
    o, g = pbox
    gx = g[0]
    L = gx[-1]
    n0 = NP ** 3
    m1 = np.tile(np.repeat(np.arange(NP), NP), NP)
    x = np.zeros(4 * n0)
    cc = np.concatenate([[0.0], 0.5 * (gx[:-1] + gx[1:])])  # cell m (1..n-1) centre, upper-corner index
    cx = 0.5 * (gx[:-1] + gx[1:])
    d = (gx[1] - gx[0] + gx[-1] - gx[-2]) / 2
    eps = np.zeros((6, n0))
    eps[0] = np.where(m1 >= 1, _cos(cc[m1], L), 0.0)                     # exx at cells 1..n-1
    eps[5] = np.where(m1 < NP - 1, _cos(cx[np.minimum(m1, NP - 2)], L), 0.0)  # eyz at x centres 0..n-2
    eps[3] = _cos(gx[m1], L)                                             # exy at x lines 0..n-1
    f0 = lambda i: float(np.float32(_cos(cc[i], L)))
    f5 = lambda i: float(np.float32(_cos(cx[i], L)))
    f3 = lambda i: float(np.float32(_cos(gx[i], L)))
    # normal rates: cells n-1 and 1 across the wrap, on both sides
    for px in (0.2 * cc[1], cc[-1] + 0.5 * (L - cc[-1])):
        _, e = o.interp(x, eps, (px, 0.4, 0.65))
        assert abs(e[0] - _wrap_lin(f0(NP - 1), f0(1), px, cc[-1], cc[1], d, L)) < 1e-12
    # shear across a periodic axis: the wrap below centre 0 only; between centres 0 and 1 the reference
    # holds centre 0, above centre n-2 it holds centre n-2
    _, e = o.interp(x, eps, (0.3 * cx[0], 0.4, 0.65))
    assert abs(e[5] - _wrap_lin(f5(NP - 2), f5(0), 0.3 * cx[0], cx[-1], cx[0], d, L)) < 1e-12
    _, e = o.interp(x, eps, (0.5 * (cx[0] + cx[1]), 0.4, 0.65))
    assert abs(e[5] - f5(0)) < 1e-12
    _, e = o.interp(x, eps, (cx[-1] + 0.5 * (L - cx[-1]), 0.4, 0.65))
    assert abs(e[5] - f5(NP - 2)) < 1e-12
    # shear along a periodic line axis: lines 0 and 1 bracket the point (walled: held at line 1)
    px = 0.4 * gx[1]
    _, e = o.interp(x, eps, (px, 0.4, 0.65))
    t = px / gx[1]
    assert abs(e[3] - ((1 - t) * f3(0) + t * f3(1))) < 1e-12


def test_translation_across_the_wrap(pbox):
    o, g = pbox
    L = g[0][-1]
    n0 = NP ** 3
    U = 0.7
    x = np.zeros(4 * n0)
    x[1::4] = U
    eps = np.zeros((6, n0))
    dt = 0.1
    mx = np.array([L - 0.02, L - 0.2, 0.01], np.float32)
    my = np.array([0.3, 0.5, 0.7], np.float32)
    mz = np.array([0.4, 0.6, 0.9], np.float32)
    markk = np.ones(3, np.float32)
    want = (mx.astype(np.float64) + U * dt) % L
    o.move_markers(x, eps, mx, my, mz, markk=markk, timestep=dt)
    assert o.wraps_out == 0
    np.testing.assert_allclose(mx, want, atol=2e-7)
    assert np.all(markk == 1.0) and np.all(my == np.float32([0.3, 0.5, 0.7]))
    # a step longer than the period leaves a point outside after one wrap (the reference exits)
    mx2 = np.array([0.5], np.float32)
    o.move_markers(x, eps, mx2, my[:1].copy(), mz[:1].copy(), timestep=3.0 * L / U)
    assert o.wraps_out > 0
