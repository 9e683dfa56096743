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
