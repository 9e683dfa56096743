"""The oracle's marker advection (movemark with vxyzcalc1 and epscalc1,
parallel/mark3mg.c:174-629, 2963, 3244) pinned by closed forms on a
non-uniform grid:

- a linear velocity field sampled at the staggered positions is reproduced
  at interior points (trilinear interpolation is exact for it), and beyond
  the first or last centre each component is held at that centre's value;
- its recomputed strain rates are constant, so their interpolation is the
  constant everywhere and the invariant is sqrt((exx^2+eyy^2+ezz^2)/2 +
  exy^2+exz^2+eyz^2);
- one RK4 step of dp/dt = A p + b is p + dt (I + dt A/2 + dt^2 A^2/6 +
  dt^3 A^3/24)(A p + b), one Euler step p + dt (A p + b), and the
  accumulated strain grows by the invariant times dt;
- markers leaving the box lose markk; outgrid 1 leaves outside markers
  alone, outgrid 2 keeps only their motion along the axes they are outside of.
"""
import numpy as np
import pytest

from pardiso2_amd import setups

N = 13
A = np.array([[0.3, -0.2, 0.1], [0.25, -0.1, 0.05], [-0.15, 0.4, -0.2]])
B = np.array([0.5, -0.3, 0.2])


def _grid(n, size, seed):
    """Non-uniform lines from 0: spacings in [0.7h, 1.3h]."""
    rng = np.random.default_rng(seed)
    h = rng.uniform(0.7, 1.3, n - 1)
    return np.concatenate([[0.0], np.cumsum(h / h.sum() * size)])


@pytest.fixture(scope="module")
def box(oracle_mod):
    g = [_grid(N, s, k) for k, s in enumerate((1.0, 0.8, 1.3))]
    o = oracle_mod.Oracle(g[0], g[1], g[2], 0)
    o.knobs(p0koef=1.0, p1koef=1.0, p2koef=0.0, nubeg=1e-3, nuend=1e3, viscmod=0)
    o.bc_lines(setups.free_slip_lines(N, N, N, gauge=None))
    n0 = o.info[0][4]
    for k in ("nu", "nxx", "nxy", "nxz", "nyz"):
        o.arr(k)[:n0] = 1.0
    c = [np.append(0.5 * (gg[:-1] + gg[1:]), gg[-1]) for gg in g]  # centres (last entry never read)
    m3, m1, m2 = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    m1, m2, m3 = m1.ravel(), m2.ravel(), m3.ravel()
    pos = {1: (g[0][m1], c[1][m2], c[2][m3]), 2: (c[0][m1], g[1][m2], c[2][m3]), 3: (c[0][m1], c[1][m2], g[2][m3])}
    x = np.zeros(4 * n0)
    for comp in (1, 2, 3):
        p = np.stack(pos[comp])
        x[comp::4] = A[comp - 1] @ p + B[comp - 1]
    _, eps, _ = o.post(x)
    return o, g, c, x, eps


def _v(p):
    return A @ p + B


def test_linear_velocity_is_reproduced_inside(box):
    o, g, c, x, eps = box
    rng = np.random.default_rng(1)
    lo = np.array([c[0][0], c[1][0], c[2][0]])
    hi = np.array([c[0][N - 2], c[1][N - 2], c[2][N - 2]])
    for _ in range(200):
        p = lo + rng.random(3) * (hi - lo)
        v, _ = o.interp(x, eps, p)
        np.testing.assert_allclose(v, _v(p), rtol=0, atol=2e-7)


def test_velocity_is_held_beyond_the_outer_centres(box):
    o, g, c, x, eps = box
    rng = np.random.default_rng(2)
    size = np.array([g[0][-1], g[1][-1], g[2][-1]])
    for _ in range(200):
        p = rng.random(3) * size
        v, _ = o.interp(x, eps, p)
        for comp in range(3):
            q = p.copy()
            for a in range(3):  # a component on lines follows p along its own axis; across, the centres bound it
                if a != comp:
                    q[a] = min(max(q[a], c[a][0]), c[a][N - 2])
            assert abs(v[comp] - _v(q)[comp]) <= 2e-7, (p, comp)


def test_strain_rates_of_a_linear_field_and_their_invariant(box):
    o, g, c, x, eps = box
    div = np.trace(A)
    e_ref = np.array([A[0, 0] - div / 3, A[1, 1] - div / 3, A[2, 2] - div / 3, (A[0, 1] + A[1, 0]) / 2,
                      (A[0, 2] + A[2, 0]) / 2, (A[1, 2] + A[2, 1]) / 2])
    rng = np.random.default_rng(3)
    size = np.array([g[0][-1], g[1][-1], g[2][-1]])
    for _ in range(100):
        p = rng.random(3) * size
        _, e = o.interp(x, eps, p)
        np.testing.assert_allclose(e, e_ref, rtol=0, atol=1e-6)
    # a mistaken half on the shear terms would fail this
    eii = np.sqrt(0.5 * (e_ref[:3] ** 2).sum() + (e_ref[3:] ** 2).sum())
    mx, my, mz = (np.full(4, v, np.float32) for v in (0.5, 0.4, 0.6))
    got = o.move_markers(x, eps, mx, my, mz, timestep=0.0)
    np.testing.assert_allclose(got, eii, rtol=2e-6)


def _rk4(p, dt):
    k = _v(p)
    return p + dt * (k + dt / 2 * (A @ k) + dt ** 2 / 6 * (A @ A @ k) + dt ** 3 / 24 * (A @ A @ A @ k))


@pytest.mark.parametrize("markmod", [0, 1])
def test_one_step_of_a_linear_flow(box, markmod):
    o, g, c, x, eps = box
    rng = np.random.default_rng(4)
    n = 64
    lo = np.array([c[0][1], c[1][1], c[2][1]])
    hi = np.array([c[0][N - 3], c[1][N - 3], c[2][N - 3]])
    p0 = (lo + rng.random((n, 3)) * (hi - lo)).astype(np.float32)
    mx, my, mz = (np.ascontiguousarray(p0[:, a]) for a in range(3))
    marke = np.full(n, 0.5, np.float32)
    dt = 0.08
    eii = o.move_markers(x, eps, mx, my, mz, marke=marke, timestep=dt, markmod=markmod)
    for i in range(n):
        p = p0[i].astype(np.float64)
        want = _rk4(p, dt) if markmod != 1 else p + dt * _v(p)
        got = np.array([mx[i], my[i], mz[i]], np.float64)
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-6)
        # Euler and RK4 differ far beyond the tolerance here
        other = p + dt * _v(p) if markmod != 1 else _rk4(p, dt)
        assert np.abs(got - other).max() > 1e-4
    np.testing.assert_allclose(marke, 0.5 + eii.astype(np.float64) * dt, rtol=1e-6)


def test_markers_leaving_or_outside_the_box(box):
    o, g, c, x, eps = box
    size = np.array([g[0][-1], g[1][-1], g[2][-1]])
    # leaving through the far x face (vx > 0 there for this field)
    p = np.array([size[0] - 0.001, 0.4, 0.5])
    assert _v(p)[0] > 0
    markk = np.array([300.0, 300.0], np.float32)
    mx, my, mz = (np.array([v, 0.5], np.float32) for v in p)
    o.move_markers(x, eps, mx, my, mz, markk=markk, timestep=0.5)
    assert mx[0] > size[0] and markk[0] == 0.0
    assert markk[1] == 300.0  # the marker that stays keeps it
    # already outside: outgrid 1 leaves it alone, outgrid 2 keeps the motion along x only
    q = np.array([-0.05, 0.4, 0.5])
    for outgrid in (1, 2, 0):
        mx, my, mz = (np.array([v], np.float32) for v in q)
        o.move_markers(x, eps, mx, my, mz, timestep=0.1, outgrid=outgrid, markmod=1)
        got = np.array([mx[0], my[0], mz[0]], np.float64)
        v = _v(np.array([0.0, q[1], q[2]]))  # interpolated at the point clamped onto the box
        if outgrid == 1:
            np.testing.assert_array_equal(got, q.astype(np.float32))
        elif outgrid == 2:
            np.testing.assert_allclose(got, [q[0] + 0.1 * v[0], q[1], q[2]], atol=2e-7)
        else:
            # the velocity across x is held at the first centre
            vv = _v(np.array([c[0][0], q[1], q[2]]))
            np.testing.assert_allclose(got, [q[0] + 0.1 * v[0], q[1] + 0.1 * vv[1], q[2] + 0.1 * vv[2]], atol=2e-7)
