"""The Schur approximation of the block preconditioner against a numpy
restatement of the same algorithm built from the oracle's assembled rows:

    K = B D^-1 G  (continuity rows x 1/|momentum diagonal| x pressure columns),
    K_l = P^T K_{l-1} P  (piecewise-constant aggregation along the hierarchy),
    one K-cycle: red/black Gauss-Seidel, summed restriction, scaled
    injection, regularised dense inverse on the coarsest level,
    z_p = -K^-1 (B D^-1 A D^-1 G) K^-1 r_p.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from pardiso2_amd import _lib as S
from pardiso2_amd import setups

pytestmark = pytest.mark.gpu


def _assembled(o, level=0):
    info = o.info[level]
    base, n4 = 4 * info[5], 4 * info[4]
    num0, pos0, lin0, val1 = (o.arr(k) for k in ("num0", "pos0", "lin0", "val1"))
    cnt = num0[base:base + n4]
    act = cnt > 0
    rows = np.repeat(np.arange(n4)[act], cnt[act])
    offs = np.concatenate([np.arange(k) for k in cnt[act]])
    pp = np.repeat(pos0[base:base + n4][act], cnt[act]) + offs
    A = sp.csr_matrix((val1[pp], (rows, lin0[pp] - base)), shape=(n4, n4))
    eq = act & (o.arr("bondm")[base:base + n4] == 0)
    return A, eq


class Mirror:
    def __init__(self, o, L, nu=2, scale=2.0):
        self.L, self.nu, self.scale = L, nu, scale
        A, eq = _assembled(o)
        n0 = o.info[0][4]
        isP = np.arange(4 * n0) % 4 == 0
        diag = np.abs(A.diagonal())
        dinv = np.where(eq & ~isP, 1.0 / np.where(diag > 0, diag, 1.0), 0.0)
        ip = np.where(isP)[0]
        self.Bm = A[ip][:, :]                               # cell rows (all P unknowns)
        self.Bm = sp.diags((eq[ip]).astype(float)) @ self.Bm  # continuity equations only
        self.G = A[:, ip]                                   # pressure columns
        self.Dinv = sp.diags(dinv)
        velmask = sp.diags((~isP).astype(float))
        self.Avv = velmask @ A @ velmask                    # velocity block, all velocity rows
        self.ip = ip
        K = (self.Bm @ self.Dinv @ self.G).tocsr()
        self.K = [K]
        self.P = []
        self.dims = [o.info[0][:3]]
        for l in range(1, L + 1):
            par = []
            for a in range(3):
                gf, gc = o.grid(l - 1, a), o.grid(l, a)
                lines = np.searchsorted(gf, gc)
                assert np.all(gf[lines] == gc)
                p = np.zeros(len(gf), int)
                for m in range(1, len(gc)):
                    p[lines[m - 1] + 1:lines[m] + 1] = m
                par.append(p)
            X, Y, Z = self.dims[-1]
            Xc, Yc, Zc = o.info[l][:3]
            m3, m1, m2 = np.meshgrid(np.arange(Z), np.arange(X), np.arange(Y), indexing="ij")
            cidx = (par[2][m3] * Xc * Yc + par[0][m1] * Yc + par[1][m2]).ravel()
            nf = X * Y * Z
            P = sp.csr_matrix((np.ones(nf), (np.arange(nf), cidx)), shape=(nf, Xc * Yc * Zc))
            self.P.append(P)
            self.K.append((P.T @ self.K[-1] @ P).tocsr())
            self.dims.append((Xc, Yc, Zc))
        Kc = self.K[L]
        d = Kc.diagonal()
        self.cells = np.where(d != 0)[0]
        Ad = Kc[self.cells][:, self.cells].toarray()
        N = len(self.cells)
        Ad += d[self.cells].mean() / N
        self.inv = np.linalg.inv(Ad)

    def colour(self, l):
        X, Y, Z = self.dims[l]
        m3, m1, m2 = np.meshgrid(np.arange(Z), np.arange(X), np.arange(Y), indexing="ij")
        return ((m1 + m2 + m3) & 1).ravel()

    def kcycle(self, l, b):
        K = self.K[l]
        if l == self.L:
            x = np.zeros_like(b)
            x[self.cells] = self.inv @ b[self.cells]
            return x
        d = K.diagonal()
        act = d != 0
        col = self.colour(l)
        x = np.zeros_like(b)

        def half(c):
            m = act & (col == c)
            r = b - K @ x
            x[m] += r[m] / d[m]
        for _ in range(self.nu):
            half(0); half(1)
        r = np.where(act, b - K @ x, 0.0)
        xc = self.kcycle(l + 1, self.P[l].T @ r)
        x += np.where(act, self.scale * (self.P[l] @ xc), 0.0)
        for _ in range(self.nu):
            half(1); half(0)
        return x

    def schur(self, r):
        rp = np.zeros(len(self.ip))
        eqp = np.asarray(self.Bm.sum(axis=1)).ravel() != 0
        rp = r[self.ip] * (np.abs(self.Bm).sum(axis=1).A.ravel() > 0)
        p1 = self.kcycle(0, rp)
        y = self.Dinv @ (self.G @ p1)
        w = self.Avv @ y
        t = self.Bm @ (self.Dinv @ w)
        return -self.kcycle(0, t)


@pytest.fixture(scope="module", params=["sinker17", "randfield25"])
def case(request, oracle_mod):
    if request.param == "sinker17":
        n, L = 17, 2
        pb = setups.sinker(n, L, eta_in=1e22, eta_out=1e20)
    else:
        n, L = 25, 2
        pb = setups.random_field(n, L, lo=19.0, hi=23.0, field_n=33)
    pb.bc = setups.free_slip_lines(n, n, n, gauge=None)
    o = oracle_mod.from_problem(pb)
    g = S.from_problem(pb)
    d = S.default_opts(L)  # the K-cycle sweeps and coarse-correction scale the solve runs
    yield pb, o, g, Mirror(o, L, nu=d.ksweeps, scale=d.kscale)
    g.close()


def test_k_operator_and_aggregated_levels(case):
    pb, o, g, m = case
    rng = np.random.default_rng(1)
    for l in range(pb.levels + 1):
        x = rng.standard_normal(o.info[l][4])
        y = g.kop_apply(l, x)
        ye = m.K[l] @ x
        np.testing.assert_allclose(y, ye, rtol=1e-12, atol=1e-12 * np.abs(ye).max(), err_msg=f"level {l}")
        # constant cell fields are in the kernel of every level's operator (walls are Neumann faces)
        act = m.K[l].diagonal() != 0
        y1 = g.kop_apply(l, act.astype(float))
        assert np.abs(y1).max() <= 1e-10 * np.abs(m.K[l].diagonal()).max()


def test_kcycle_matches_restatement_and_contracts(case):
    pb, o, g, m = case
    rng = np.random.default_rng(2)
    act = m.K[0].diagonal() != 0
    b = rng.standard_normal(o.info[0][4]) * act
    x = g.kcycle_apply(b, precision=1, ksweeps=m.nu, kscale=m.scale)
    xe = m.kcycle(0, b)
    np.testing.assert_allclose(x, xe, rtol=1e-9, atol=1e-9 * np.abs(xe).max())
    b0 = b - b[act].mean() * act
    x0 = g.kcycle_apply(b0, precision=1, ksweeps=m.nu, kscale=m.scale)
    assert np.linalg.norm(b0 - m.K[0] @ x0) < 0.5 * np.linalg.norm(b0)
    x32 = g.kcycle_apply(b, precision=0, ksweeps=m.nu, kscale=m.scale)
    np.testing.assert_allclose(x32, xe, rtol=1e-3, atol=1e-4 * np.abs(xe).max())


def test_schur_block_matches_restatement(case):
    pb, o, g, m = case
    rng = np.random.default_rng(3)
    r = rng.standard_normal(4 * o.info[0][4])
    z = g.schur_apply(r, precision=1, schur=1)
    ze = m.schur(r)
    zp = z[0::4]
    np.testing.assert_allclose(zp, ze, rtol=1e-8, atol=1e-8 * np.abs(ze).max())
    assert not np.any(z[1::4]) and not np.any(z[2::4]) and not np.any(z[3::4])
