"""Host-side problem generators and the boundary grammar parser."""
import numpy as np

from pardiso2_amd import setups


def test_parse_reference_grammar_line():
    txt = "/comment\n Vx    x   x     0   y-1   1   z-2   0       1  -100 +0 +0 0 0\n~\n"
    (ln,) = setups.parse_bc_lines(txt, 29, 29, 29)
    assert ln == (1, 28, 28, 0, 27, 1, 26, 0.0, 1.0, -100, 0, 0)


def test_markers_stay_in_their_cells_and_are_cell_major():
    g = setups.box_grid(9)
    mx, my, mz = setups.markers_regular(g, g, g, jitter=0.45, seed=1)
    assert mx.size == 8 ** 3 * 8
    cx = np.searchsorted(g, mx) - 1
    cy = np.searchsorted(g, my) - 1
    cz = np.searchsorted(g, mz) - 1
    cell = (cz * 8 + cx) * 8 + cy
    assert np.all(np.diff(cell) >= 0)
    assert np.bincount(cell).min() == 8


def test_free_slip_walls_on_all_levels(oracle_mod):
    n = 19
    pb = setups.sinker(n, 2)
    pb.bc = setups.free_slip_lines(n, n, n, gauge=None)
    o = oracle_mod.from_problem(pb)
    bondm = o.arr("bondm")
    for l, (X, Y, Z, XY, N, P) in enumerate(o.info):
        # normal velocities on the wall lines 1 and X-2 are boundary rows with value 0
        for m3 in range(Z - 1):
            for m2 in range(Y - 1):
                for m1 in (1, X - 2):
                    r = 4 * (P + m3 * XY + m1 * Y + m2) + 1
                    b = bondm[r]
                    assert b and o.arr("bondv1")[2 * b] == 0 and o.arr("bondv1")[2 * b + 1] == 0
