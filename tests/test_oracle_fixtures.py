"""Oracle pinned against the reference's own dump of the coarsest-level rows.

mg/coef_MG_finest.txt holds every (coefficient, column) pair of the coarsest
level of a 29^3 run (zsize 1.16, periodic in x and y, isoviscous through
nuend=1, p2koef=0.1, two coarse levels), written by the reference's row store
(mg/gaus3mg_MG.c:373-382) in row order; ghost unknowns appear as identity
rows.  Reproducing all 50,319 pairs pins the hierarchy (gridcheck), the
boundary grammar projected on coarse levels, every stress stencil including
the periodic branches, the grad-div term and the slot order of the rows.
"""
import os

import numpy as np
import pytest

from pardiso2_amd import setups

GOLD = os.path.join(os.path.dirname(__file__), "golden", "mg_coarsest_rows.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLD)


@pytest.fixture(scope="module")
def coarse29(oracle_mod, golden):
    g = golden["grid"]
    n = int(g[0])
    gx = setups.uniform_grid(n, g[4])
    gz = setups.uniform_grid(n, g[6])
    o = oracle_mod.Oracle(gx, gx.copy(), gz, int(g[3]), periodic=(int(g[7]), int(g[8]), int(g[9])))
    o.knobs(p0koef=0.1, p1koef=0.0, p2koef=0.1, nubeg=1.0, nuend=1.0, viscmod=0)
    o.bc_lines([tuple(int(v) if i not in (7, 8) else float(v) for i, v in enumerate(row)) for row in golden["bc"]])
    o.assemble()
    return o


def _rows(o, L):
    info = o.info[L]
    base = 4 * info[5]
    num0 = o.arr("num0")
    vals, idxs = [], []
    for i in range(4 * info[4]):
        r = base + i
        if num0[r] == 0:
            vals.append(1.0)
            idxs.append(i)
            continue
        idx, val, _ = o.row(r)
        vals.extend(val)
        idxs.extend(idx - base)
    return np.array(vals), np.array(idxs)


def test_hierarchy_sizes(coarse29):
    assert [inf[0] for inf in coarse29.info] == [29, 17, 11]
    assert coarse29.info[2][5] == 29 ** 3 + 17 ** 3  # mgp of the coarsest level


def test_coarsest_rows_match_reference_dump(coarse29, golden):
    vals, idxs = _rows(coarse29, 2)
    ref_c, ref_i = golden["coef"], golden["idx"]
    assert vals.size == ref_c.size == 50319
    np.testing.assert_array_equal(idxs, ref_i)
    # the dump is printed with %e (7 significant digits)
    np.testing.assert_allclose(vals, ref_c, rtol=6e-7, atol=0)


def test_continuity_rows_match_pcoefs(coarse29, golden):
    L = 2
    base = 4 * coarse29.info[L][5]
    got = []
    num0 = coarse29.arr("num0")
    bondm = coarse29.arr("bondm")
    for i in range(0, 4 * coarse29.info[L][4], 4):
        r = base + i
        if num0[r] and not bondm[r]:
            idx, val, _ = coarse29.row(r)
            got.extend(zip(val, idx, idx - base))
    got = np.array(got)
    assert got.shape[0] == golden["pcoef"].size
    np.testing.assert_allclose(got[:, 0], golden["pcoef"], rtol=6e-7)
    np.testing.assert_array_equal(got[:, 1].astype(np.int64), golden["pabs"])
    np.testing.assert_array_equal(got[:, 2].astype(np.int32), golden["prel"])


def test_planted_slot_swap_is_detected(coarse29, golden):
    # a transposed pair of slots (a plausible assembly mistake) fails the check
    vals, idxs = _rows(coarse29, 2)
    k = 7420
    vals[k], vals[k + 1] = vals[k + 1], vals[k]
    assert not np.allclose(vals, golden["coef"], rtol=6e-7, atol=0)
