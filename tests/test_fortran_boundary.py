"""The Fortran module (tt-fort_amd/fortran/ttfort_amd.F90) and a caller
that uses the reference's routine names compile with the image's flang and
link against both library kinds: the i4 preset with default INTEGER(4) and
the i8 preset with -fdefault-integer-8, as the reference's Makefile presets
build tt-fort programs.  The linked programs run their host-only checks
here (no GPU); the GPU checks of the same programs are in
test_gpu_fortran.py.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tt-fort_amd"))
import build as B  # noqa: E402


@pytest.mark.skipif(not os.path.exists(B.FLANG), reason="flang not in this image")
@pytest.mark.parametrize("kind", ["i4", "i8"])
def test_fortran_module_compiles_links_and_runs(kind, tmp_path):
    lib = os.path.join(B.LIB, "libttfort_amd%s.so" % ("" if kind == "i4" else "_i8"))
    assert os.path.exists(lib), "library not built: " + lib
    exe, cmds = B.fortran_commands(kind, str(tmp_path))
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=str(tmp_path))
        assert r.returncode == 0, (" ".join(cmd), r.stderr[-3000:])
    r = subprocess.run([exe, "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "my_chop3 ok" in r.stdout and "ttfort_amd" in r.stdout
    # the module's default-integer kind matches the library's
    nm = subprocess.run(["nm", "-D", lib], capture_output=True, text=True).stdout
    for sym in ("dbfun3", "dqr", "djac_gen", "djac_apply", "duchol_fort", "dgmresr_hh_fort", "drow_add",
                "drow_cut", "dtransp", "dperm1324", "ttf_primme_matvec", "ttf_zmatvec"):
        assert (" T %s\n" % sym) in nm, sym
