"""The allocation-free marker averaging (stk_markers_to_nodes_noalloc) gives
exactly the node fields of stk_markers_to_nodes, on two calls in a row on the
same GPU (the second smaller than the first, so stale scratch contents from
the first call would show), and refuses inputs beyond its scratch limit."""
import numpy as np
import pytest

from pardiso2_amd import _lib as S
from pardiso2_amd import setups

pytestmark = pytest.mark.gpu

NAMES = ("nu", "ro", "nxx", "nxy", "nxz", "nyz")


def _fields_both_ways(pb):
    a = S.Stokes(pb.gx, pb.gy, pb.gz, pb.levels)
    b = S.Stokes(pb.gx, pb.gy, pb.gz, pb.levels)
    try:
        for g in (a, b):
            k = pb.knobs
            g.knobs(k["p0koef"], k["p1koef"], k["p2koef"], k["v0koef"], k["v1koef"], k["nubeg"], k["nuend"],
                    k["viscmod"], pb.gravity)
        a.markers_to_nodes(pb.mx, pb.my, pb.mz, pb.mnu, pb.mro)
        rc = b.markers_to_nodes_noalloc(pb.mx, pb.my, pb.mz, pb.mnu, pb.mro)
        assert rc == 0, rc
        fa, fb = a.fields(), b.fields()
        return fa, fb, b
    finally:
        a.close()


def test_noalloc_markers_equal_allocating_entry_point_twice():
    # first call: 49^3 nodes, 884,736 markers; second: 33^3, 262,144 markers
    for n, seed in ((49, 3), (33, 4)):
        pb = setups.random_field(n, setups.coarsest_levels(n), seed=seed)
        print(f"noalloc: {n}^3, {pb.mx.size} markers", flush=True)
        fa, fb, b = _fields_both_ways(pb)
        try:
            for k in NAMES:
                np.testing.assert_array_equal(fb[k][: b.n0], fa[k][: b.n0], err_msg=k)
            # the noalloc fields give the same prepared solve
            b.bc_lines(pb.bc)
            b.prepare()
            b.hydrostatic_guess()
            st = b.solve(tol=1e-8)
            assert st.status == 0
        finally:
            b.close()


def test_noalloc_rejects_inputs_beyond_the_scratch_limit():
    n = 67  # 66^3 cells > 2^18
    g = setups.box_grid(n)
    ctx = S.Stokes(g, g, g, setups.coarsest_levels(n))
    try:
        m = np.full(8, 0.5, np.float32)
        assert ctx.markers_to_nodes_noalloc(m, m, m, m + 1e21, m + 3300) == -1
    finally:
        ctx.close()
    n = 33
    g = setups.box_grid(n)
    ctx = S.Stokes(g, g, g, setups.coarsest_levels(n))
    try:
        m = np.full((1 << 20) + 1, 0.5, np.float32)
        assert ctx.markers_to_nodes_noalloc(m, m, m, m + 1e21, m + 3300) == -1
    finally:
        ctx.close()
