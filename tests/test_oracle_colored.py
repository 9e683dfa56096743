"""The oracle's multicolour gausmat4 over the assembled rows (the order the
device uses on a caller's row store): the exact solution of the assembled
system is its fixed point on every level (a wrong sign, scale or index in
the row update moves it), and one sweep lowers the residual of a smooth
error in the velocity rows."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from pardiso2_amd import setups


def _system(o, l):
    info = o.info[l]
    base, n4 = 4 * info[5], 4 * info[4]
    num0, pos0, lin0, val1 = (o.arr(k) for k in ("num0", "pos0", "lin0", "val1"))
    cnt = num0[base:base + n4]
    act = cnt > 0
    rows = np.repeat(np.arange(n4)[act], cnt[act])
    offs = np.concatenate([np.arange(k) for k in cnt[act]])
    pp = np.repeat(pos0[base:base + n4][act], cnt[act]) + offs
    A = sp.csr_matrix((val1[pp], (rows, lin0[pp] - base)), shape=(n4, n4))
    A = A + sp.diags((~act).astype(float))
    return A, base, n4, act


def test_exact_solution_is_a_fixed_point(oracle_mod):
    n = 11
    L = setups.coarsest_levels(n)
    # unit-scale viscosities: the sparse direct solve is then exact to rounding
    pb = setups.sinker(n, L, eta_in=10.0, eta_out=1.0, knobs=dict(nubeg=0.1, nuend=100.0))
    pb.bc = setups.free_slip_lines(n, n, n, gauge=(2, 2, 2))
    o = oracle_mod.from_problem(pb)
    rng = np.random.default_rng(3)
    for l in range(L + 1):
        A, base, n4, act = _system(o, l)
        f = o.arr("fre1")
        f[base:base + n4] = np.where(act, rng.standard_normal(n4), 0.0)
        x = spla.spsolve(A.tocsc(), f[base:base + n4])
        o.arr("sol0")[base:base + n4] = x
        o.gausmat4_colored(2, l)
        y = o.arr("sol0")[base:base + n4]
        assert np.abs(y - x).max() <= 1e-9 * np.abs(x).max(), l
        assert np.abs(o.arr("fre0")[base:base + n4]).max() <= 1e-9 * np.abs(f[base:base + n4]).max(), l
        # a planted error in the data moves the state: the check is not vacuous
        o.arr("sol0")[base:base + n4] = x
        f[base + 4 * (n4 // 8) + 1] += 1.0
        o.gausmat4_colored(1, l)
        assert np.abs(o.arr("sol0")[base:base + n4] - x).max() > 1e-6 * np.abs(x).max()
