"""MI355X-native variable-viscosity 3D Stokes multigrid (gfx950 HIP).

The solver lives in libpardiso2_amd.so (C ABI, include/pardiso2_amd.h); this
package binds it (``_lib``) and provides the benchmark problem generators
(``setups``).
"""
from . import setups  # noqa: F401

__all__ = ["setups"]
