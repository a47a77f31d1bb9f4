"""The Fortran bind(C) module driven from a Fortran program (the reference-side binding): the
generalized eigensolver and density matrix by residuals, the neighbour list by its invariants."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

PROGRAM = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "dftbplus_amd", "fortran", "test_dftbhip")


def test_fortran_module_program():
    if not os.path.exists(PROGRAM):
        pytest.skip("the Fortran test program is not built (no amdflang at build time)")
    r = subprocess.run([PROGRAM], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "PASS" in r.stdout
