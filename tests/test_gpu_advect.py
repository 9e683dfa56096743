"""Marker advection on the device (stk_move_markers; movemark, vxyzcalc1,
epscalc1, parallel/mark3mg.c:174-629, 2963, 3244) against the oracle:

- through a solved sinker, random markers inside, on and beyond the walls,
  with accumulated strain, types and markk, RK4 and Euler, every outgrid
  mode: positions, strain, markk and the invariant are bit-identical to the
  oracle's step on the same velocities and strain rates;
- through a linear velocity field on a non-uniform grid, one RK4 step
  lands on the closed form (the oracle's own pin, repeated on the device).
"""
import sys

import numpy as np
import pytest

from pardiso2_amd import _lib as S
from pardiso2_amd import setups

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[advect] {msg}", flush=True)
    sys.stdout.flush()


def _markers(rng, n, size, lines=None):
    p = rng.uniform(-0.03, 1.03, (n, 3)) * size
    p[: n // 8] = rng.random((n // 8, 3)) * size * 0.02  # near the lower corner
    p[n // 8: n // 4, 0] = 0.0                            # on a wall
    if lines is not None:                                 # on grid lines (as float32 holds them)
        k = slice(n // 4, n // 4 + n // 8)
        for a in range(3):
            p[k, a] = rng.choice(lines[a], n // 8)
    return [np.ascontiguousarray(p[:, a], dtype=np.float32) for a in range(3)]


@pytest.mark.parametrize("markmod,outgrid", [(0, 0), (1, 0), (0, 1), (0, 2)])
def test_step_is_bit_identical_to_the_oracle(oracle_mod, markmod, outgrid):
    n = 33
    pb = setups.sinker(n, setups.coarsest_levels(n), seed=1)
    g = S.from_problem(pb)
    try:
        g.hydrostatic_guess()
        st = g.solve(precision=0, tol=1e-8)
        assert st.status == 0
        x = g.solution()
        _, eps = g.stress_strain()
        o = oracle_mod.Oracle(pb.gx, pb.gy, pb.gz, pb.levels)
        size = np.array([pb.gx[-1], pb.gy[-1], pb.gz[-1]])
        rng = np.random.default_rng(10 + markmod + 3 * outgrid)
        nm = 20000
        lines = (pb.gx, pb.gy, pb.gz)
        dev = _markers(rng, nm, size, lines)
        marke = rng.uniform(-0.5, 1.0, nm).astype(np.float32)
        markk = np.full(nm, 300.0, np.float32)
        markt = rng.choice([0, 60, 120], nm).astype(np.int32)
        ref = [a.copy() for a in dev]
        marke_o, markk_o = marke.copy(), markk.copy()
        vmax = np.abs(x[1::4]).max()
        dt = 0.02 * (pb.gx[-1] / max(vmax, 1e-300))   # about 2 % of the box per step
        eii = g.move_markers(*dev, marke=marke, markk=markk, markt=markt, timestep=dt, markmod=markmod,
                             outgrid=outgrid)
        eii_o = o.move_markers(x, eps, *ref, marke=marke_o, markk=markk_o, markt=markt, timestep=dt,
                               markmod=markmod, outgrid=outgrid)
        for a, b, name in zip(dev, ref, "xyz"):
            np.testing.assert_array_equal(a, b, err_msg=name)
        np.testing.assert_array_equal(marke, marke_o)
        np.testing.assert_array_equal(markk, markk_o)
        np.testing.assert_array_equal(eii, eii_o)
        moved = np.abs(dev[0] - _markers(np.random.default_rng(10 + markmod + 3 * outgrid), nm, size, lines)[0]).max()
        assert moved > 0
        _say(f"markmod {markmod} outgrid {outgrid}: {nm} markers bit-identical, {int((markk == 0).sum())} left "
             f"the box, max eii {eii.max():.3e}")
    finally:
        g.close()


def test_rk4_step_of_a_linear_flow_on_the_device():
    N = 13
    A = np.array([[0.3, -0.2, 0.1], [0.25, -0.1, 0.05], [-0.15, 0.4, -0.2]])
    B = np.array([0.5, -0.3, 0.2])
    rng = np.random.default_rng(0)
    grids = []
    for s in (1.0, 0.8, 1.3):
        h = rng.uniform(0.7, 1.3, N - 1)
        grids.append(np.concatenate([[0.0], np.cumsum(h / h.sum() * s)]))
    g = S.Stokes(*grids, 1)
    try:
        g.knobs(p0koef=1.0, p1koef=1.0, p2koef=0.0, nubeg=1e-3, nuend=1e3, viscmod=0)
        g.bc_lines(setups.free_slip_lines(N, N, N, gauge=None))
        one = np.ones(g.n0, np.float32)
        g.set_fields(one, one * 0, one, one, one, one)
        g.prepare()
        c = [np.append(0.5 * (gg[:-1] + gg[1:]), gg[-1]) for gg in grids]
        m3, m1, m2 = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
        m1, m2, m3 = m1.ravel(), m2.ravel(), m3.ravel()
        pos = {1: (grids[0][m1], c[1][m2], c[2][m3]), 2: (c[0][m1], grids[1][m2], c[2][m3]),
               3: (c[0][m1], c[1][m2], grids[2][m3])}
        x = np.zeros(4 * g.n0)
        for comp in (1, 2, 3):
            x[comp::4] = A[comp - 1] @ np.stack(pos[comp]) + B[comp - 1]
        g.set_solution(x)
        lo = np.array([c[0][1], c[1][1], c[2][1]])
        hi = np.array([c[0][N - 3], c[1][N - 3], c[2][N - 3]])
        p0 = (lo + rng.random((256, 3)) * (hi - lo)).astype(np.float32)
        mx, my, mz = (np.ascontiguousarray(p0[:, a]) for a in range(3))
        dt = 0.08
        g.move_markers(mx, my, mz, timestep=dt)
        for i in range(len(p0)):
            p = p0[i].astype(np.float64)
            k = A @ p + B
            want = p + dt * (k + dt / 2 * (A @ k) + dt ** 2 / 6 * (A @ A @ k) + dt ** 3 / 24 * (A @ A @ A @ k))
            np.testing.assert_allclose([mx[i], my[i], mz[i]], want, rtol=0, atol=2e-6)
    finally:
        g.close()


def test_step_on_a_refined_grid_is_bit_identical_to_the_oracle(oracle_mod):
    """Lines 20x denser over the middle third of each axis: the cell search
    walks over many lines per bin; an arbitrary velocity field."""
    N = 29
    grids = []
    for s in (1.0, 0.8, 1.3):
        h = np.ones(N - 1)
        h[(N - 1) // 3: 2 * (N - 1) // 3] = 0.05
        grids.append(np.concatenate([[0.0], np.cumsum(h / h.sum() * s)]))
    g = S.Stokes(*grids, 1)
    try:
        g.knobs(p0koef=1.0, p1koef=1.0, p2koef=0.0, nubeg=1e-3, nuend=1e3, viscmod=0)
        g.bc_lines(setups.free_slip_lines(N, N, N, gauge=None))
        one = np.ones(g.n0, np.float32)
        g.set_fields(one, one * 0, one, one, one, one)
        g.prepare()
        rng = np.random.default_rng(5)
        x = rng.standard_normal(4 * g.n0)
        g.set_solution(x)
        _, eps = g.stress_strain()
        o = oracle_mod.Oracle(*grids, 1)
        size = np.array([gg[-1] for gg in grids])
        nm = 30000
        dev = _markers(rng, nm, size, grids)
        ref = [a.copy() for a in dev]
        marke = np.full(nm, 0.25, np.float32)
        marke_o = marke.copy()
        eii = g.move_markers(*dev, marke=marke, timestep=0.01)
        eii_o = o.move_markers(x, eps, *ref, marke=marke_o, timestep=0.01)
        for a, b, name in zip(dev, ref, "xyz"):
            np.testing.assert_array_equal(a, b, err_msg=name)
        np.testing.assert_array_equal(marke, marke_o)
        np.testing.assert_array_equal(eii, eii_o)
        # the velocity-only step (no strain rates read) moves them the same way
        dev2 = _markers(np.random.default_rng(6), nm, size, grids)
        ref2 = [a.copy() for a in dev2]
        assert g.move_markers(*dev2, timestep=0.01, want_eii=False) is None
        o.move_markers(x, eps, *ref2, timestep=0.01)
        for a, b, name in zip(dev2, ref2, "xyz"):
            np.testing.assert_array_equal(a, b, err_msg=name)
        _say(f"refined grid: {nm} markers bit-identical, with and without strain")
    finally:
        g.close()


def test_skipped_markers_keep_their_eii():
    """Markers the step does not move (outside fluid markers under outgrid 2) keep the caller's eii
    entries: the device buffer starts from the caller's array, not from whatever it last held."""
    n = 17
    pb = setups.sinker(n, setups.coarsest_levels(n), seed=1)
    g = S.from_problem(pb)
    try:
        g.hydrostatic_guess()
        assert g.solve(precision=0, tol=1e-8).status == 0
        m = 512
        rng = np.random.default_rng(3)
        p = [np.ascontiguousarray(rng.uniform(0.05, 0.95, m) * v, dtype=np.float32)
             for v in (pb.gx[-1], pb.gy[-1], pb.gz[-1])]
        p[2][: m // 2] = -0.01  # below the box
        t = np.full(m, 60, np.int32)  # fluid: not moved while outside
        sentinel = np.full(m, -7.0, np.float32)
        for _ in range(3):  # reused device memory must not leak into the skipped entries
            e = sentinel.copy()
            g.move_markers(*[a.copy() for a in p], markt=t, timestep=1e-3, outgrid=2, eii=e)
            np.testing.assert_array_equal(e[: m // 2], -7.0)
            assert np.all(e[m // 2:] >= 0)
    finally:
        g.close()
