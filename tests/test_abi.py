"""The C-ABI library builds for gfx950, loads without a GPU, and exports
every function include/pardiso2_amd.h declares (no compute calls here)."""
import os
import re

from pardiso2_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "pardiso2_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:int|void|double)\s+\*?\s*(\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIBPATH):
        from pardiso2_amd import build
        build.build()
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= set(names)


def test_reference_entry_points_keep_reference_names():
    names = set(declared_functions())
    for ref in ("matclear", "gausmat", "gausmat2", "pardisomat", "gausmat4", "vprestrict", "vpprolong",
                "ronurestrict", "vpiterate"):
        assert ref in names
