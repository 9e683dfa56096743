"""A C program built against include/pardiso2_amd.h and linked with
libpardiso2_amd.so, declaring the reference's own prototypes of the replaced
routines beside the header (tests/native/ref_caller.c): it compiles with
-Werror, links, and its host-side checks pass without a device."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "pardiso2_amd")


def test_c_caller_compiles_links_and_runs(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    if not os.path.exists(os.path.join(LIBDIR, "libpardiso2_amd.so")):
        pytest.skip("library not built")
    exe = str(tmp_path / "ref_caller")
    cmd = [cc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "native", "ref_caller.c"), "-L", LIBDIR, "-lpardiso2_amd",
           f"-Wl,-rpath,{LIBDIR}", "-lm", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stdout


def test_conflicting_prototype_is_rejected(tmp_path):
    """The check has teeth: a prototype that disagrees with the header does
    not compile."""
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    src = tmp_path / "bad.c"
    src.write_text('#include "pardiso2_amd.h"\nint gausmat4(int, long int, long int, double);\nint main(void){return 0;}\n')
    r = subprocess.run([cc, "-std=c99", "-c", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(tmp_path / "bad.o")], capture_output=True, text=True)
    assert r.returncode != 0 and "conflicting types" in r.stderr
