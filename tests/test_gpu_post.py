"""The post-solve pass on the device against the oracle: stress and
strain-rate planes (parallel/move3mg.c:2908-3046) and the reference's
convergence norms (parallel/move3mg.c:3090-3238), on arbitrary states and on
the solution of a solve, whose stats carry the same norms.  The oracle's
pass is pinned by closed forms in test_post_norms.py.
"""
import sys

import numpy as np
import pytest

from pardiso2_amd import _lib as S
from pardiso2_amd import setups

pytestmark = pytest.mark.gpu

KEYS = ("stoksum", "stoksum1", "stoksum2", "contsum", "contsumb", "bondsum", "maxvxyz")


def _say(msg):
    print(f"[post] {msg}", flush=True)
    sys.stdout.flush()


def _oracle_on_device_fields(oracle_mod, pb, g):
    o = oracle_mod.Oracle(pb.gx, pb.gy, pb.gz, pb.levels)
    k = pb.knobs
    o.knobs(k["p0koef"], k["p1koef"], k["p2koef"], k["v0koef"], k["v1koef"], k["nubeg"], k["nuend"], k["viscmod"],
            pb.gravity)
    o.bc_lines(pb.bc)
    f = g.fields()
    o.set_fields(f["nu"], f["ro"], f["nxx"], f["nxy"], f["nxz"], f["nyz"])
    return o


def _compare_norms(dev, ref, rtol):
    for k in KEYS:
        assert dev[k] == pytest.approx(ref[k], rel=rtol, abs=1e-300), k
    for k in ("stoknum", "contnum", "contnumb", "bondnum", "minvx", "maxvx", "minvy", "maxvy", "minvz", "maxvz",
              "minpr", "maxpr", "minnu", "maxnu"):
        assert dev[k] == ref[k], k


@pytest.fixture(scope="module", params=[("sinker", 33, 0), ("randfield", 25, 1), ("spheres", 33, 2)])
def case(request, oracle_mod):
    wl, n, vm = request.param
    L = setups.coarsest_levels(n)
    knobs = dict(viscmod=vm)
    if wl == "sinker":
        pb = setups.sinker(n, L, seed=1, knobs=knobs)
    elif wl == "randfield":
        pb = setups.random_field(n, L, seed=3, knobs=knobs)
    else:
        pb = setups.random_spheres(n, L, seed=2, knobs=knobs)
    pb.bc = setups.free_slip_lines(n, n, n, gauge=None)
    _say(f"{wl} {n}^3 viscmod {vm}")
    g = S.from_problem(pb)
    o = _oracle_on_device_fields(oracle_mod, pb, g)
    yield wl, pb, g, o
    g.close()


def test_stress_planes_and_norms_of_an_arbitrary_state(case, oracle_mod):
    wl, pb, g, o = case
    rng = np.random.default_rng(5)
    x = rng.standard_normal(4 * g.n0) * 1e-9
    x[0::4] *= 1e11  # pressures of the gravity scale, velocities of a slow flow
    g.set_solution(x)
    sig, eps = g.stress_strain()
    sig_o, eps_o, out = o.post(x)
    for k in range(6):
        np.testing.assert_allclose(eps[k], eps_o[k], rtol=1e-12, atol=1e-13 * np.abs(eps_o[k]).max(), err_msg=str(k))
        np.testing.assert_allclose(sig[k], sig_o[k], rtol=1e-12, atol=1e-13 * np.abs(sig_o[k]).max(), err_msg=str(k))
    dev = g.norms()
    ref = oracle_mod.post_norms(out)
    # viscmod 1/2 take exp/log of the device, the oracle's are the host C library's
    _compare_norms(dev, ref, 1e-11 if wl != "sinker" else 1e-12)
    _say(f"{wl}: stoksum {dev['stoksum']:.3e} contsum {dev['contsum']:.3e} bondsum {dev['bondsum']:.3e}")


def test_solve_reports_the_reference_norms_of_its_solution(case, oracle_mod):
    wl, pb, g, o = case
    g.hydrostatic_guess()
    st = g.solve(precision=0, tol=1e-8)
    assert st.status == 0
    nrm = st.norms.as_dict()
    x = g.solution()
    _, _, out = o.post(x, stresses=False)
    ref = oracle_mod.post_norms(out)
    _compare_norms(nrm, ref, 1e-11)
    # the same as evaluating them afterwards
    _compare_norms(g.norms(), nrm, 0.0)
    # with the grid's own velocity scale and DIVVMIN = 9e-3 the reference's stop test passes at 1e-8
    assert nrm["converged"] == 1
    _say(f"{wl}: {st.iterations} iterations relres {st.relres:.2e}: stoksum {nrm['stoksum']:.3e} "
         f"stoksum1 {nrm['stoksum1']:.3e} contsum {nrm['contsum']:.3e} bondsum {nrm['bondsum']:.3e}")
