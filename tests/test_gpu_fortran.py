"""The Fortran boundary check programs (built by build() from
tt-fort_amd/fortran/, one per integer kind) run their GPU checks: dbfun3
through the reference's routine name against a direct loop over its index
definition, and dqr by orthonormality and reconstruction, each to 1e-12."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tt-fort_amd", "lib")


@pytest.mark.parametrize("kind", ["i4", "i8"])
def test_fortran_caller_on_gpu(kind):
    exe = os.path.join(LIB, "fortran_%s" % kind, "check_boundary_%s" % kind)
    assert os.path.exists(exe), "Fortran check not built (run build()): " + exe
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "boundary checks passed" in r.stdout
