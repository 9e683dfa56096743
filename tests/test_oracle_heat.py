"""The oracle's heat-equation multigrid (parallel/heat3mg.c) pinned by
closed forms (CPU).

- Rows (heaterr ynerr = 1, heat3mg.c:195-249): on a uniform grid with
  constant conductivity every face conductance is k/h^2 and the diagonal
  rho cp/dt + 6k/h^2, with rho cp clamped below at 2e6 (heat3mg.c:46-47);
  the periodic x wrap (heat3mg.c:21, 36) gives the same row on line 0.
- A temperature linear in z with no source is a steady exact state of the
  discrete step (Dirichlet top and bottom, zero-flux or periodic sides).
- Layered conductivity k(z): the piecewise-linear T(z) whose flux
  (k_face dT/dz, k_face the mean of the two nodes) is constant is exact;
  the same profile with the face means shifted by one line is not.
- Transfers (heat3mg.c:954-1330): restriction of a constant residual is the
  constant (normalised), prolongation reproduces a linear coarse field on
  the fine nodes, property restriction reproduces a linear field.
- The titerate step (heat3mg.c:384-590) converges to the direct solution of
  its level-0 rows, in the reference's sweep order and in the device's.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from pardiso2_amd import setups


def _heat(oracle_mod, n, Lt, bc_text, periodic=(0, 0, 0), fields=None, timestep=1e4):
    g = setups.uniform_grid(n, 1.0)
    o = oracle_mod.Oracle(g, g.copy(), g.copy(), max(Lt, 1), periodic=periodic)
    h = oracle_mod.Heat(o, Lt)
    h.bc_lines(setups.parse_bc_lines(bc_text, n, n, n))
    N = n ** 3
    ro, cp, kt, tk, tk1 = fields if fields else (np.full(N, 3000.0), np.full(N, 1000.0), np.full(N, 3.0),
                                                  np.zeros(N), np.zeros(N))
    h.set_fields(ro, cp, kt, tk, tk1)
    h.set(timestep)
    return g, o, h


def test_row_closed_form_and_periodic_wrap(oracle_mod):
    n, k, dt = 9, 3.0, 1e4
    g, o, h = _heat(oracle_mod, n, 0, setups.heat_bc_text(0.0, 1.0, xperiodic=True), periodic=(1, 0, 0),
                    fields=(np.full(n ** 3, 1500.0), np.full(n ** 3, 1000.0), np.full(n ** 3, k), np.full(n ** 3, 0.25),
                            np.full(n ** 3, 2.0)))
    hh = g[1] - g[0]
    R = 2e6  # 1500 * 1000 is clamped up to 2e6
    for m1 in (0, 3):  # line 0 wraps across the periodic x axis
        idx, val, rhs = h.row(0, m1, 4, 4)
        assert len(idx) == 7
        np.testing.assert_allclose(val[0], R / dt + 6 * k / hh ** 2, rtol=1e-13)
        np.testing.assert_allclose(val[1:], -k / hh ** 2, rtol=1e-12)
        np.testing.assert_allclose(rhs, 2.0 + 0.25 * R / dt, rtol=1e-13)
        nb = {int(v) for v in idx[1:]}
        xm = (n - 2) if m1 == 0 else m1 - 1
        assert 4 * n * n + xm * n + 4 in nb  # x- neighbour (wrapped on line 0)


@pytest.mark.parametrize("xper", [False, True])
def test_linear_temperature_is_steady_exact(oracle_mod, xper):
    n = 9  # spacing 1/8: T = z is exact in FP32
    N = n ** 3
    g = setups.uniform_grid(n, 1.0)
    tk = np.broadcast_to(g[:, None, None], (n, n, n)).reshape(-1)  # T = z
    kxy = np.random.default_rng(2).uniform(1.0, 4.0, (n, n)).astype(np.float32)
    kt = np.broadcast_to(kxy[None], (n, n, n)).reshape(-1)  # varies across z lines only
    fields = (np.full(N, 3100.0), np.full(N, 1100.0), kt, tk, np.zeros(N))
    g, o, h = _heat(oracle_mod, n, 0, setups.heat_bc_text(0.0, 1.0, xperiodic=xper),
                    periodic=(int(xper), 0, 0), fields=fields)
    h.assemble()
    h.arr("sol0")[:N] = tk
    h.residual(0)
    r = h.arr("fre0")[:N]
    b = h.arr("fre1")[:N]
    assert np.abs(r).max() < 1e-12 * np.abs(b).max()


def _layered(n, shift):
    rng = np.random.default_rng(3)
    kz = rng.uniform(1.0, 6.0, n).astype(np.float32)  # conductivity per z node line
    g = setups.uniform_grid(n, 1.0)
    h = g[1] - g[0]
    kf = (kz[:-1] + kz[1:]).astype(np.float64) / 2.0  # face between lines m3 and m3+1 (FP32 sum)
    if shift:
        kf = np.roll(kf, 1)
    q = 0.37
    T = np.concatenate([[0.0], np.cumsum(q * h / kf)])
    return kz, T


@pytest.mark.parametrize("shift", [False, True])
def test_layered_conductivity_flux_form(oracle_mod, shift):
    n = 13
    N = n ** 3
    kz, T = _layered(n, shift)
    kt = np.broadcast_to(kz[:, None, None], (n, n, n)).reshape(-1)
    tk = np.broadcast_to(T[:, None, None], (n, n, n)).reshape(-1)
    bc = setups.heat_bc_text(float(np.float32(T[0])), float(np.float32(T[-1])))
    # a step so long that the rho cp / dt term vanishes: the rows are the steady flux balance
    g, o, h = _heat(oracle_mod, n, 0, bc, fields=(np.full(N, 3000.0), np.full(N, 1000.0), kt, tk, np.zeros(N)),
                    timestep=1e30)
    h.assemble()
    h.arr("sol0")[:N] = tk
    h.residual(0)
    r = h.arr("fre0")[:N].reshape(n, n, n)[1:-1, 1:-1, 1:-1]
    scale = np.abs(h.arr("val1")).max() * np.abs(T).max()
    if shift:
        assert np.abs(r).max() > 1e-6 * scale
    else:
        assert np.abs(r).max() < 1e-12 * scale


def test_transfers_reproduce_constants_and_linear_fields(oracle_mod):
    n = 17
    N = n ** 3
    g, o, h = _heat(oracle_mod, n, 1, setups.heat_bc_text(0.0, 1.0))
    h.assemble()
    i1 = o.info[1]
    b1, n1 = i1[5], i1[4]
    # restriction: a constant fine residual gives the constant on every coarse node without a T equation
    h.arr("fre0")[:N] = 2.5
    h.trestrict(1)
    tb = h.arr("tbondm")[b1:b1 + n1]
    f1 = h.arr("fre1")[b1:b1 + n1]
    np.testing.assert_allclose(f1[tb == 0], 2.5, rtol=1e-14)
    assert np.all(f1[tb != 0] == 0)
    # prolongation of a coarse field linear in x, y, z onto fine nodes without a T equation
    X, Y, Z = i1[:3]
    gc = [o.grid(1, a) for a in range(3)]
    cz, cx, cy = np.meshgrid(gc[2], gc[0], gc[1], indexing="ij")
    lin = lambda x, y, z: 0.3 + 1.5 * x - 0.7 * y + 2.0 * z
    h.arr("sol0")[b1:b1 + n1] = lin(cx, cy, cz).reshape(-1)
    h.arr("sol0")[:N] = 0.0
    h.tprolong(1, 0)
    fz, fx, fy = np.meshgrid(g, g, g, indexing="ij")
    fine = h.arr("sol0")[:N]
    tbf = h.arr("tbondm")[:N]
    np.testing.assert_allclose(fine[tbf == 0], lin(fx, fy, fz).reshape(-1)[tbf == 0], rtol=1e-12, atol=1e-12)
    # property restriction of a linear conductivity: exact where the coarse node's two spacings are equal
    # (the 17-line axis coarsens to lines 0 1 2 4 6 ... 14 15 16: coarse lines 3..7 are symmetric)
    h.arr("kt")[:N] = lin(fx, fy, fz).reshape(-1)
    h.rocprestrict(1)
    kc = h.arr("kt")[b1:b1 + n1].reshape(Z, X, Y)
    s = slice(3, -3)
    np.testing.assert_allclose(kc[s, s, s], lin(cx, cy, cz)[s, s, s], rtol=2e-6)
    assert np.abs(kc[2, 5, 5] - lin(cx, cy, cz)[2, 5, 5]) > 1e-3  # uneven spacings: the hat mean shifts


def _direct(h, N):
    num0, lin0, val1 = h.arr("num0")[:N], h.arr("lin0"), h.arr("val1")
    rows, cols, vals = [], [], []
    for r in range(N):
        for k in range(num0[r]):
            rows.append(r)
            cols.append(lin0[7 * r + k])
            vals.append(val1[7 * r + k])
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    return spl.spsolve(A.tocsc(), h.arr("fre1")[:N].copy())


@pytest.mark.parametrize("colored,xper", [(False, False), (True, False), (True, True)])
def test_multigrid_step_converges_to_direct_solution(oracle_mod, colored, xper):
    n, Lt = 17, 2
    hp = setups.heat_box(n, Lt, xperiodic=xper)
    g = hp.gx
    o = oracle_mod.Oracle(g, g.copy(), g.copy(), Lt, periodic=hp.periodic)
    h = oracle_mod.Heat(o, Lt)
    h.bc_lines(hp.bc)
    h.set_fields(hp.ro, hp.cp, hp.kt, hp.tk, hp.tk1)
    h.set(hp.timestep)
    it, hs, bs = h.step([3, 3, 3], [0, 3, 3], heatmin=1e-6, maxit=60, colored=colored)
    assert 0 < it < 60, (it, hs)
    N = n ** 3
    x = _direct(h, N)
    np.testing.assert_allclose(h.arr("tk"), x.astype(np.float32), rtol=0, atol=1e-5)


def test_titerate_substep_count(oracle_mod):
    # the sub-step limit maxtkstep / max|tk1 / rho cp| (heat3mg.c:331-382, 401-405): a step 2.5 limits
    # long is cut into 0.4 + 0.4 + 0.2; no source: one sub-step; a zero step: nothing done
    hp = setups.heat_box(17, 2)
    g = hp.gx

    def run(tk1, maxtk, timestep):
        o = oracle_mod.Oracle(g, g.copy(), g.copy(), 2)
        h = oracle_mod.Heat(o, 2)
        h.bc_lines(hp.bc)
        h.set_fields(hp.ro, hp.cp, hp.kt, hp.tk, tk1)
        h.set(hp.timestep)
        return h, h.titerate(timestep, maxtk, [3, 3, 3], [0, 3, 3], heatmin=1e-5, maxit=100)

    tk1 = (hp.tk1 * 2e4).astype(np.float32)
    heat_rows = run(tk1, 25.0, 0.0)[0].arr("tbondm")[: 17 ** 3] == 0
    rcp = np.maximum(hp.ro.astype(np.float32) * hp.cp.astype(np.float32), np.float32(2e6))
    rate = np.max(np.abs(tk1 / rcp.astype(np.float64))[heat_rows])
    h, r = run(tk1, float(rate * hp.timestep / 2.5), hp.timestep)
    assert r["substeps"] == 3 and r["converged"]
    np.testing.assert_allclose(r["tkstep"], hp.timestep / 2.5, rtol=1e-6)
    h, r = run(np.zeros_like(tk1), 25.0, hp.timestep)
    assert r["substeps"] == 1 and r["tkstep"] == 1e30
    h, r = run(tk1, 25.0, 0.0)
    assert r["substeps"] == 0
    np.testing.assert_array_equal(h.arr("tk"), hp.tk.astype(np.float32))


def _src_setup(oracle_mod, n=11, seed=5):
    rng = np.random.default_rng(seed)
    h = rng.uniform(0.7, 1.3, n - 1)
    g = np.concatenate([[0.0], np.cumsum(h / h.sum())])
    o = oracle_mod.Oracle(g, g.copy(), g.copy(), 0)
    H = oracle_mod.Heat(o, 0)
    H.bc_lines(setups.parse_bc_lines(setups.heat_bc_text(0.0, 1.0), n, n, n))
    N = n ** 3
    ro = np.full(N, 3300.0, np.float32)
    ro[: N // 3] = 1500.0  # below the 2e3 threshold: no heating terms there
    tk = rng.uniform(0.5, 1.5, N).astype(np.float32)
    H.set_fields(ro, np.full(N, 1000.0), np.full(N, 3.0), tk, np.zeros(N))
    return g, o, H, ro, tk, N


def test_sources_closed_forms(oracle_mod):
    """heaterr's right part (heat3mg.c:54-187): with a constant velocity, a pressure linear in x, y, z at
    the cell centres and constant stresses and strain rates, the adiabatic terms are et tk0 v.grad P
    (adiabyn 2) and et tk0 ro v.g (adiabyn 1), the frictional term the full contraction sig : eps,
    all where ro > 2e3 only."""
    g, o, H, ro, tk, N = _src_setup(oracle_mod)
    n = len(g)
    v0 = np.array([0.3, -0.2, 0.5])
    a = np.array([2.0, -1.0, 0.5])
    c = np.concatenate([[0.0], 0.5 * (g[:-1] + g[1:])])  # centre of cell m (upper corner m)
    m3, m1, m2 = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x = np.zeros(4 * N)
    x[0::4] = (a[0] * c[m1] + a[1] * c[m2] + a[2] * c[m3]).ravel()
    for k in range(3):
        x[1 + k::4] = v0[k]
    sv = np.array([1.5, -0.5, 2.0, 0.25, -0.75, 0.5])
    ev = np.array([0.2, 0.1, -0.3, 0.05, 0.4, -0.15])
    sig = np.repeat(sv[:, None], N, axis=1)
    eps = np.repeat(ev[:, None], N, axis=1)
    ht = np.random.default_rng(6).uniform(0.0, 1e-6, N).astype(np.float32)
    et = np.full(N, 3e-5, np.float32)
    grav = (0.0, 0.0, 9.81)
    heat = H.arr("tbondm")[:N] == 0
    hot = heat & (ro > 2e3)
    v32 = v0.astype(np.float32).astype(np.float64)
    f32 = lambda z: z.astype(np.float32).astype(np.float64)
    contraction = float(np.sum(f32(sv[:3]) * f32(ev[:3])) + 2 * np.sum(f32(sv[3:]) * f32(ev[3:])))
    for adiab, frict in ((0, 0), (1, 0), (2, 0), (0, 1), (2, 1)):
        H.sources(x, sig, eps, ht, et, adiab, frict, grav)
        tk1 = H.arr("tk1").astype(np.float64)
        want = ht.astype(np.float64).copy()
        etk = (et * tk).astype(np.float64)
        if adiab == 1:
            want[hot] += (et * tk * ro).astype(np.float64)[hot] * float(v32 @ np.array(grav))
        if adiab == 2:
            want[hot] += etk[hot] * float(v32 @ a)
        if frict:
            want[hot] += contraction
        np.testing.assert_allclose(tk1[heat], want[heat].astype(np.float32), rtol=2e-6, atol=1e-12,
                                   err_msg=f"adiabyn {adiab} frictyn {frict}")
        np.testing.assert_array_equal(tk1[~heat], ht[~heat])


def _mk_setup(oracle_mod, n=9, seed=9):
    rng = np.random.default_rng(seed)
    g = setups.uniform_grid(n, 1.0)
    o = oracle_mod.Oracle(g, g.copy(), g.copy(), 0)
    H = oracle_mod.Heat(o, 0)
    N = n ** 3
    zz, xx, yy = np.meshgrid(g, g, g, indexing="ij")
    T0 = (0.2 + 0.5 * xx - 0.3 * yy + 0.8 * zz).ravel()   # linear: interpolated exactly
    dT = 0.05
    H.set_fields(np.full(N, 3000.0), np.full(N, 1000.0), np.full(N, 3.0), T0, np.zeros(N))
    H.arr("tk")[:] = (T0 + dT).astype(np.float32)  # the step's result: a uniform change
    return g, o, H, T0, dT, rng


def _mk_markers(g, rng, per_cell=2):
    n = len(g)
    c = []
    for k in range(n - 1):
        for i in range(n - 1):
            for j in range(n - 1):
                for _ in range(per_cell):
                    c.append((g[i] + rng.uniform(0.1, 0.9) * (g[i + 1] - g[i]), g[j] + rng.uniform(0.1, 0.9) * (g[j + 1] - g[j]),
                              g[k] + rng.uniform(0.1, 0.9) * (g[k + 1] - g[k])))
    p = np.array(c)
    return [np.ascontiguousarray(p[:, a], dtype=np.float32) for a in range(3)]


def _lin(p):
    return 0.2 + 0.5 * p[0].astype(np.float64) - 0.3 * p[1] + 0.8 * p[2]


def test_marker_temperatures_closed_forms(oracle_mod):
    """titerate's marker update (heat3mg.c:618-950) with linear node temperatures: markers in
    equilibrium with the grid gain the nodal change; fluid markers and new markers (markk <= 0) take
    the interpolated new / old temperature (tdeep below zdeep); a uniform offset of all markers
    survives the numerical diffusion (the antidiffusion returns the mean of what diffused); the
    boundary cells cap a marker at the nodes' largest temperature."""
    g, o, H, T0, dT, rng = _mk_setup(oracle_mod)
    mx, my, mz = _mk_markers(g, rng)
    M = len(mx)
    Told = _lin((mx, my, mz))
    # in equilibrium, with and without diffusion
    for hd in (0.0, 1.0):
        markk = Told.astype(np.float32)
        H.markers(mx, my, mz, np.zeros(M, np.int32), markk, heatdif=hd, timestep=1e9)
        np.testing.assert_allclose(markk, Told + dT, rtol=0, atol=2e-6)
    # fluid markers take the new temperature, new markers the old one (or tdeep below zdeep)
    t = np.where(np.arange(M) % 3 == 0, 60, 0).astype(np.int32)
    markk = np.where(np.arange(M) % 3 == 1, -1.0, Told).astype(np.float32)
    H.markers(mx, my, mz, t, markk, heatdif=0.0, zdeep=0.5, tdeep=0.9)
    fluid, new = t == 60, (np.arange(M) % 3 == 1)
    np.testing.assert_allclose(markk[fluid], Told[fluid] + dT, atol=2e-6)
    deep = new & (my > 0.5) & (Told < 0.9)
    np.testing.assert_allclose(markk[deep], 0.9 + dT, atol=2e-6)
    np.testing.assert_allclose(markk[new & ~deep], Told[new & ~deep] + dT, atol=2e-6)
    # a uniform offset survives diffusion + antidiffusion; an alternating one decays
    for off, keeps in ((0.01, True), (None, False)):
        d = np.full(M, 0.01) if off else np.where(np.arange(M) % 2 == 0, 0.01, -0.01)
        markk = (Told + d).astype(np.float32)
        H.markers(mx, my, mz, np.zeros(M, np.int32), markk, heatdif=1.0, timestep=1e3)
        got = markk - (Told + dT)
        # (markers whose nodes no boundary-cell marker reaches: those are capped, see below)
        inner = (mx > g[2]) & (mx < g[-3]) & (my > g[2]) & (my < g[-3]) & (mz > g[2]) & (mz < g[-3])
        if keeps:
            np.testing.assert_allclose(got[inner], 0.01, atol=1e-5)
        else:
            # subgrid differences relax by exp(-K), K = heatdif kt / (rho cp) 2 (3 / h^2) dt
            h = g[1] - g[0]
            K = 1.0 * 3.0 / 3e6 * 2.0 * (3.0 / h ** 2) * 1e3
            r = np.abs(got[inner]) / 0.01
            assert np.all(np.abs(r - np.exp(-K)) < 0.1), (r.min(), r.max(), np.exp(-K))
    # boundary cap
    markk = Told.astype(np.float32)
    k = int(np.argmin(mx))  # a marker in the first x cell
    markk[k] = 50.0
    H.markers(mx, my, mz, np.zeros(M, np.int32), markk, heatdif=1.0, timestep=1e-30)
    assert markk[k] < T0.max() + dT + 1e-5
    markk = Told.astype(np.float32)
    markk[k] = 50.0
    H.markers(mx, my, mz, np.zeros(M, np.int32), markk, heatdif=0.0)
    assert markk[k] > 49.0  # no cap without diffusion
