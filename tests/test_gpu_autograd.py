"""The torch.autograd interface: backward (hand-written adjoint gradient kernel) against
central finite differences of the energy, the finite-difference Slater-Koster derivative path
and the CPU oracle. PyTorch runs in a child process (it brings its own HIP runtime)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed_refined"])
def test_autograd_gradients(precision):
    proc = subprocess.run([sys.executable, os.path.join(HERE, "autograd_driver.py"), "water32",
                           "48", str(precision)], capture_output=True, text=True, timeout=600)
    print(proc.stderr[-3000:])
    assert proc.returncode == 0, proc.stderr[-2000:]
    r = json.loads(proc.stdout.strip().splitlines()[-1])
    print(r)
    # central differences with h = 1e-4 bohr: truncation error ~ h^2 |E'''| ~ 1e-8
    assert r["fd_max_err"] < 1e-6
    # analytic versus finite-difference Slater-Koster derivatives (step eps^(1/4))
    assert r["sk_fd_path_max_diff"] < 1e-7
    assert r["oracle_max_diff"] < 1e-7
    assert r["oracle_energy_diff"] / r["natom"] < 1e-8
