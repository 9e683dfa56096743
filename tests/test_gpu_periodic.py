"""Periodic boxes on the device against the oracle (reference periodic
branches: parallel/move3mg.c:27-34, 400-429, 834-847, 1481-1483; partners
parallel/in3mg.c:339-341; marker folds parallel/mark3mg.c:893-955,
1032-1111; post-solve copies parallel/move3mg.c:2933-3044, 3321-3379).

Two setups:
  * the 29^3 run behind tests/golden/mg_coarsest_rows.npz (periodic in x
    and y, zsize 1.16, its own boundary lines, 3 levels), with random node
    fields for the residual and sweep parity, and isoviscous for the
    comparison with the reference's printed coarsest rows;
  * the simple-shear box of parallel/init_inclusion.t3c
    (setups.periodic_shear): periodic x and y, vx = +1/-1 on the z walls.

Sweeps run with tiles narrower than the periodic extent: a tile that held
both sides of the wrap would read the copies it staged where the oracle's
row store names the unknown itself.
"""
import os

import numpy as np
import pytest

from pardiso2_amd import _lib as S
from pardiso2_amd import setups

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "mg_coarsest_rows.npz")


def _mg_lines(golden):
    return [tuple(int(v) if i not in (7, 8) else float(v) for i, v in enumerate(row)) for row in golden["bc"]]


def _mg_pair(oracle_mod, seed=0, iso=False, p2=0.1, smooth=False):
    """Oracle and device on the 29^3 periodic hierarchy of the golden run."""
    golden = np.load(GOLD)
    g = golden["grid"]
    n, L = int(g[0]), int(g[3])
    gx = setups.uniform_grid(n, g[4])
    gz = setups.uniform_grid(n, g[6])
    per = (int(g[7]), int(g[8]), int(g[9]))
    N = n ** 3
    if iso:
        f = [np.ones(N, np.float32) for _ in range(6)]
        kn = dict(nubeg=1.0, nuend=1.0)
    elif smooth:
        # one decade of viscosity varying smoothly (periodic in x and y), density +-10 %
        x, y, z = np.meshgrid(gx, gx, gz, indexing="ij")
        le = 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.3 * np.sin(3.0 * z + seed)
        ev = np.transpose(10.0 ** le, (2, 0, 1)).reshape(-1).astype(np.float32)  # (m3, m1, m2) node order
        f = [ev.copy() for _ in range(6)]
        f[1] = (3000.0 * (1.0 + 0.1 * np.transpose(np.cos(2 * np.pi * (x + y)), (2, 0, 1)).reshape(-1))).astype(np.float32)
        kn = dict(nubeg=1e-3, nuend=1e3)
    else:
        rng = np.random.default_rng(seed)
        f = [(10.0 ** rng.uniform(-1.5, 1.5, N)).astype(np.float32) for _ in range(6)]
        f[1] = (3000.0 + 300.0 * rng.standard_normal(N)).astype(np.float32)
        kn = dict(nubeg=1e-3, nuend=1e3)
    knobs = dict(p0koef=0.1, p1koef=0.0, p2koef=p2, v0koef=1.0, v1koef=1.0, viscmod=0, **kn)
    grav = (0.0, 0.0, 0.0) if iso else (0.3, -0.2, -9.81)
    o = oracle_mod.Oracle(gx, gx.copy(), gz, L, periodic=per)
    o.knobs(knobs["p0koef"], knobs["p1koef"], knobs["p2koef"], knobs["v0koef"], knobs["v1koef"], knobs["nubeg"],
            knobs["nuend"], knobs["viscmod"], grav)
    o.bc_lines(_mg_lines(golden))
    o.set_fields(*f)
    o.ronurestrict_all()
    o.assemble()
    d = S.Stokes(gx, gx.copy(), gz, L, periodic=per)
    d.knobs(knobs["p0koef"], knobs["p1koef"], knobs["p2koef"], knobs["v0koef"], knobs["v1koef"], knobs["nubeg"],
            knobs["nuend"], knobs["viscmod"], grav)
    d.bc_lines(_mg_lines(golden))
    d.set_fields(*f)
    d.prepare()
    return golden, o, d


@pytest.fixture(scope="module")
def mg(oracle_mod):
    golden, o, d = _mg_pair(oracle_mod, seed=1)
    yield golden, o, d
    d.close()


def test_periodic_fields_restriction_and_nd(mg):
    _, o, d = mg
    f = d.fields()
    n0 = o.info[0][4]
    np.testing.assert_array_equal(f["nu"][n0:], o.arr("nu")[n0:])
    np.testing.assert_array_equal(f["ro"][n0:], o.arr("ro")[n0:])
    num0, bondm = o.arr("num0"), o.arr("bondm")
    cont = np.where((num0[0::4] > 0) & (bondm[0::4] == 0))[0]
    np.testing.assert_array_equal(f["nd"][cont], o.arr("nd")[cont])


@pytest.mark.parametrize("level", [0, 1, 2])
def test_periodic_residual_bit_identical(mg, level):
    """b - A x of every row of every level, wrapped stencils included, in the
    reference's summation order."""
    _, o, d = mg
    b4, n4 = 4 * o.info[level][5], 4 * o.info[level][4]
    rng = np.random.default_rng(10 + level)
    x = rng.standard_normal(n4)
    f = rng.standard_normal(n4) * (o.arr("num0")[b4:b4 + n4] > 0)
    o.arr("sol0")[b4:b4 + n4] = x
    o.arr("fre1")[b4:b4 + n4] = f
    o.residual(level)
    r_or = o.arr("fre0")[b4:b4 + n4].copy()
    d.vec_set(S.SOL0, level, x)
    d.vec_set(S.FRE1, level, f)
    d.smooth(level, 0, tile=(8, 2, 3))
    r_d = d.vec_get(S.FRE0, level)
    rows = o.arr("num0")[b4:b4 + n4] > 0
    np.testing.assert_allclose(r_d[rows], r_or[rows], rtol=1e-12, atol=1e-12 * np.abs(r_or).max())
    assert np.mean(r_d[rows] == r_or[rows]) > 0.999


def test_periodic_mg_solve_with_gravity(oracle_mod):
    """The golden run's periodic hierarchy and boundary lines, smooth random
    viscosity (one decade) and density, gravity with x and y parts (the
    rows on line 0 read the wrapped density node): the device solve reaches
    1e-8 by the oracle's own rows and right-hand side."""
    golden, o, d = _mg_pair(oracle_mod, seed=4, smooth=True)
    try:
        st = d.solve(tol=1e-8)
        assert st.status == 0, (st.status, st.iterations, st.relres)
        rel, _ = o.relres_rows(d.solution())
        assert rel <= 1.0000001e-8, rel
    finally:
        d.close()


@pytest.mark.parametrize("level,tile", [(0, (8, 2, 3)), (0, (4, 4, 4)), (1, (4, 2, 2)), (2, (4, 2, 2))])
def test_periodic_fp64_sweeps_match_oracle(oracle_mod, level, tile):
    """Three FP64 tile sweeps (coupled, with pressure rows) equal the oracle's
    tile sweep on the rows, periodic partners set by the chain pass."""
    _, o, d = _mg_pair(oracle_mod, seed=2 + level)
    try:
        b4, n4 = 4 * o.info[level][5], 4 * o.info[level][4]
        rng = np.random.default_rng(20 + level)
        rows = o.arr("num0")[b4:b4 + n4] > 0
        x = rng.standard_normal(n4) * rows
        f = rng.standard_normal(n4) * rows
        o.arr("sol0")[b4:b4 + n4] = x
        o.arr("fre1")[b4:b4 + n4] = f
        d.vec_set(S.SOL0, level, x)
        d.vec_set(S.FRE1, level, f)
        o.sweep_tiles(3, level, *tile)
        d.smooth(level, 3, tile=tile)
        xo = o.arr("sol0")[b4:b4 + n4]
        xd = d.vec_get(S.SOL0, level)
        np.testing.assert_allclose(xd[rows], xo[rows], rtol=1e-12, atol=1e-12 * np.abs(xo).max())
        # unknowns without a row hold zeros, not the copies the tiles staged
        assert not np.any(xd[~rows])
    finally:
        d.close()


def test_periodic_fp32_sweep_to_fp32_accuracy(oracle_mod):
    """The V-cycle's FP32 velocity sweep on the periodic level 0 makes the
    oracle's FP64 row updates to FP32 accuracy."""
    _, o, d = _mg_pair(oracle_mod, seed=7)
    o.set_vonly(1)
    try:
        n4 = 4 * o.info[0][4]
        rng = np.random.default_rng(8)
        rows = o.arr("num0")[:n4] > 0
        x = rng.standard_normal(n4) * rows
        x[0::4] = 0.0
        f = rng.standard_normal(n4) * rows
        o.arr("sol0")[:n4] = x
        o.arr("fre1")[:n4] = f
        d.vec_set(S.SOL0, 0, x)
        d.vec_set(S.FRE1, 0, f)
        o.sweep_tiles(3, 0, 8, 4, 4)
        d.smooth(0, 3, tile=(8, 4, 4), precision=0, velocity_only=True)
        xo = o.arr("sol0")[:n4]
        xd = d.vec_get(S.SOL0, 0)
        err = np.abs(xd[rows] - xo[rows]).max() / np.abs(xo).max()
        assert 0 < err <= 2e-5, err
    finally:
        d.close()


def test_periodic_operator_matches_reference_rows():
    """The device operator on the coarsest level of the golden run
    (isoviscous, p2koef 0.1) applied to a random vector equals the product
    with the 50,319 coefficients the reference printed for that level, to
    the dump's %e precision."""
    import oracle as O
    O.build()
    golden, o, d = _mg_pair(O, iso=True)
    try:
        L = 2
        X, Y, Z = o.info[L][:3]
        n4 = 4 * X * Y * Z
        coef, idx = golden["coef"], golden["idx"]
        # rows of the dump: re-derive boundaries from the oracle's row lengths (ghosts are 1-entry identity rows)
        num0 = o.arr("num0")[4 * o.info[L][5]:4 * o.info[L][5] + n4]
        lens = np.where(num0 > 0, num0, 1)
        assert lens.sum() == coef.size
        starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
        rng = np.random.default_rng(3)
        x = rng.standard_normal(n4)
        ref = np.add.reduceat(coef * x[idx], starts)
        y = d.apply(L, x)
        rows = num0 > 0
        scale = np.add.reduceat(np.abs(coef * x[idx]), starts)
        assert np.all(np.abs(y[rows] - ref[rows]) <= 2e-6 * scale[rows] + 1e-300)
    finally:
        d.close()


def _shear_pair(oracle_mod, n=17, count=1, **kw):
    pb = setups.periodic_shear(n, count=count, **kw)
    return pb, oracle_mod.from_problem(pb), S.from_problem(pb)


def test_periodic_marker_folds_bit_identical(oracle_mod):
    """Node fields of the periodic shear box from its markers: the last-line
    sums folded into line 0 and copied back, bit-identical to the oracle."""
    pb, o, d = _shear_pair(oracle_mod, n=17, count=2, r=0.25)
    try:
        f = d.fields()
        n0 = o.info[0][4]
        for k in ("nu", "ro", "nxx", "nxy", "nxz", "nyz"):
            np.testing.assert_array_equal(f[k][:n0], o.arr(k)[:n0], err_msg=k)
        np.testing.assert_array_equal(f["nu"][n0:], o.arr("nu")[n0:])
    finally:
        d.close()


def test_periodic_couette_solve_exact():
    """Isoviscous periodic shear box (exactly constant fields): the discrete
    solution is vx linear in z between the held cell centres, vy = vz = 0,
    P = 0; the device solve reproduces it."""
    pb = setups.periodic_shear(17, count=0)
    d = S.Stokes(pb.gx, pb.gy, pb.gz, pb.levels, periodic=pb.periodic)
    try:
        k = pb.knobs
        d.knobs(k["p0koef"], k["p1koef"], k["p2koef"], k["v0koef"], k["v1koef"], k["nubeg"], k["nuend"],
                k["viscmod"], pb.gravity)
        d.bc_lines(pb.bc)
        N = d.n0
        one = np.ones(N, np.float32)
        d.set_fields(one, one * 3000, one, one, one, one)
        d.prepare()
        st = d.solve(tol=1e-10)
        assert st.status == 0
        x = d.solution()
        X = Y = Z = 17
        zc = (pb.gz[:-1] + pb.gz[1:]) / 2
        ex = np.zeros_like(x)
        for m3 in range(Z - 1):
            v = 1 - 2 * (zc[m3] - zc[0]) / (zc[Z - 2] - zc[0])
            for m1 in range(X):
                ex[4 * ((m3 * X + m1) * Y + np.arange(Y - 1)) + 1] = v
        err = np.abs(x - ex)
        assert err[1::4].max() < 1e-9 and err[2::4].max() < 1e-9 and err[3::4].max() < 1e-9
        assert np.abs(x[0::4]).max() < 1e-8
    finally:
        d.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_periodic_shear_inclusion_solve(oracle_mod, precision):
    """The reference's periodic simple-shear box with a 1e3 inclusion solves
    to 1e-8 (the oracle's residual of the device solution), passes the
    reference's own stop test, and the norms equal the oracle's."""
    pb, o, d = _shear_pair(oracle_mod, n=33, count=1, r=0.2)
    try:
        st = d.solve(tol=1e-8, precision=precision)
        assert st.status == 0, st
        x = d.solution()
        rel, _ = o.relres_rows(x)
        assert rel <= 1.0000001e-8, rel
        nd = d.norms()
        out = o.post(x, stresses=False)[2]
        no = oracle_mod.post_norms(out)
        for k in ("stoksum", "stoksum1", "stoksum2", "contsum", "bondsum"):
            assert abs(nd[k] - no[k]) <= 1e-10 * max(abs(no[k]), 1e-300), (k, nd[k], no[k])
        assert nd["converged"] == 1
    finally:
        d.close()


def test_periodic_stresses_match_oracle(oracle_mod):
    """Stress and strain-rate planes of a solved periodic box: shear stresses
    on the last periodic line from line 0, normal ones on line 0 from the last
    cell, as the reference stores them."""
    pb, o, d = _shear_pair(oracle_mod, n=17, count=1, r=0.25)
    try:
        st = d.solve(tol=1e-8)
        assert st.status == 0
        x = d.solution()
        sig_d, eps_d = d.stress_strain()
        sig_o, eps_o, _ = o.post(x, stresses=True)
        for k in range(6):
            a, b = sig_d[k], sig_o[k]
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max(), err_msg=f"sig {k}")
            a, b = eps_d[k], eps_o[k]
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max(), err_msg=f"eps {k}")
        # x periodic: sxy on the last x line equals line 0's
        X, Y = 17, 17
        s = sig_d[3].reshape(17, X, Y)
        np.testing.assert_array_equal(s[:-1, X - 1, 1:-1], s[:-1, 0, 1:-1])
    finally:
        d.close()


@pytest.mark.parametrize("markmod,outgrid", [(0, 0), (1, 0), (0, 2)])
def test_periodic_advection_bit_identical(oracle_mod, markmod, outgrid):
    """Markers through the solved periodic shear box (movemark's periodic
    branches, mark3mg.c:337-485, 2992-3760): interpolations across the x and
    y wraps, RK stage points and moved markers wrapped by one period;
    positions, strain, markk and the invariant bit-identical to the oracle's
    step on the same velocities and strain rates."""
    pb, o, d = _shear_pair(oracle_mod, n=33, count=1, r=0.2)
    try:
        st = d.solve(tol=1e-8)
        assert st.status == 0
        x = d.solution()
        _, eps = d.stress_strain()
        size = np.array([pb.gx[-1], pb.gy[-1], pb.gz[-1]])
        rng = np.random.default_rng(40 + markmod + 3 * outgrid)
        nm = 20000
        p = rng.random((nm, 3)) * size
        p[: nm // 4, 0] = rng.choice([0.0, 1.0], nm // 4) * size[0] + rng.uniform(-0.02, 0.02, nm // 4) * size[0]
        p[: nm // 4, 0] = np.mod(p[: nm // 4, 0], size[0])                  # near the x wrap
        p[nm // 4: nm // 2, 1] = np.mod(rng.uniform(-0.03, 0.03, nm // 4) * size[1], size[1])  # near the y wrap
        k = slice(nm // 2, nm // 2 + nm // 8)                               # on grid lines
        for a, gl in enumerate((pb.gx, pb.gy, pb.gz)):
            p[k, a] = rng.choice(gl, nm // 8)
        if outgrid == 2:
            p[-nm // 16:, 2] = rng.uniform(-0.03, 0.0, nm // 16) * size[2]   # below the z wall
        dev = [np.ascontiguousarray(p[:, a], dtype=np.float32) for a in range(3)]
        ref = [a.copy() for a in dev]
        marke = rng.uniform(-0.5, 1.0, nm).astype(np.float32)
        markk = np.full(nm, 300.0, np.float32)
        markt = rng.choice([0, 60, 120], nm).astype(np.int32)
        marke_o, markk_o = marke.copy(), markk.copy()
        vmax = np.abs(x[1::4]).max()
        dt = 0.04 * size[0] / vmax
        eii = d.move_markers(*dev, marke=marke, markk=markk, markt=markt, timestep=dt, markmod=markmod,
                             outgrid=outgrid)
        eii_o = o.move_markers(x, eps, *ref, marke=marke_o, markk=markk_o, markt=markt, timestep=dt,
                               markmod=markmod, outgrid=outgrid)
        assert o.wraps_out == 0
        for a, b, name in zip(dev, ref, "xyz"):
            np.testing.assert_array_equal(a, b, err_msg=name)
        np.testing.assert_array_equal(marke, marke_o)
        np.testing.assert_array_equal(markk, markk_o)
        np.testing.assert_array_equal(eii, eii_o)
        crossed = np.abs(dev[0] - p[:, 0].astype(np.float32)) > 0.5 * size[0]
        assert crossed.sum() > 10, crossed.sum()
        print(f"[periodic advect] markmod {markmod} outgrid {outgrid}: {nm} markers bit-identical, "
              f"{int(crossed.sum())} crossed the x wrap", flush=True)
    finally:
        d.close()


def test_periodic_advection_past_one_period_is_refused(oracle_mod):
    """A step longer than the period leaves points outside after one wrap:
    the reference exits there; the device finishes the step and returns
    STK_EINVAL."""
    pb, o, d = _shear_pair(oracle_mod, n=17, count=0)
    try:
        st = d.solve(tol=1e-8)
        assert st.status == 0
        x = d.solution()
        vmax = np.abs(x[1::4]).max()
        m = [np.array([0.5 * pb.gx[-1]], np.float32), np.array([0.5], np.float32),
             np.array([0.9 * pb.gz[-1]], np.float32)]
        with pytest.raises(S.StkError, match="invalid"):
            d.move_markers(*m, timestep=3.0 * pb.gx[-1] / vmax, want_eii=False)
    finally:
        d.close()
