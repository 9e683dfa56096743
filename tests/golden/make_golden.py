"""Write the golden fixtures of tests/golden from the reference's own data dumps.

Run in the development container only (it reads the reference tree):
    python tests/golden/make_golden.py /root/reference

Fixtures (data only, inputs and expected outputs):
  mg_coarsest_rows.npz
    coef, idx   all 50,319 (coefficient, level-relative column) pairs that
                mg/gaus3mg_MG.c:373-382 printed for the coarsest level of the
                run that produced mg/coef_MG_finest.txt, in row order
    pcoef, pabs, prel   the continuity-row pairs of mg/Pcoefs_gaus.txt
    bc          boundary lines of that run: the block of mg/init_sphere.t3c
                parsed with the reference grammar, plus the pressure cell of
                parallel/init_inclusion.t3c:53 (numbers only)
    grid        (nx, ny, nz, levels, xsize, ysize, zsize, xper, yper, zper)
  mg_vprestrict_pair.npz
    fre0_l1     level-1 residual fre0 of the same run after the up-cycle
                smoothing (mg/sol0_prerestrict.txt column 3, written by
                mg/move3mgold.c:2507-2515 after vprestrict(2))
    fre1_l2     the restricted level-2 right-hand side fre1 of that cycle
                (mg/sol0_gaus_precoarsest.txt column 2, move3mgold.c:2521-2527)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main(ref):
    from pardiso2_amd import setups
    c = np.loadtxt(os.path.join(ref, "mg", "coef_MG_finest.txt"))
    p = np.loadtxt(os.path.join(ref, "mg", "Pcoefs_gaus.txt"))
    txt = open(os.path.join(ref, "mg", "init_sphere.t3c")).read()
    blk = txt[txt.index("/BONDARY_CONDITIONS_DESCRIPTIONS"):]
    n = 29
    lines = setups.parse_bc_lines(blk, n, n, n) + [(0, 2, 2, 2, 2, 2, 2, 0.0, 0.0, 0, 0, 0)]
    np.savez_compressed(os.path.join(HERE, "mg_coarsest_rows.npz"),
                        coef=c[:, 0], idx=c[:, 1].astype(np.int32),
                        pcoef=p[:, 0], pabs=p[:, 1].astype(np.int64), prel=p[:, 2].astype(np.int32),
                        bc=np.array(lines, dtype=np.float64),
                        grid=np.array([n, n, n, 2, 1.0, 1.0, 1.16, 1, 1, 0], dtype=np.float64))
    print("wrote", os.path.join(HERE, "mg_coarsest_rows.npz"))
    pre = np.loadtxt(os.path.join(ref, "mg", "sol0_prerestrict.txt"))
    pc = np.loadtxt(os.path.join(ref, "mg", "sol0_gaus_precoarsest.txt"))
    np.savez_compressed(os.path.join(HERE, "mg_vprestrict_pair.npz"), fre0_l1=pre[:, 2], fre1_l2=pc[:, 1])
    print("wrote", os.path.join(HERE, "mg_vprestrict_pair.npz"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
