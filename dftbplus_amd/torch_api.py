"""torch.autograd interface to the SCC-DFTB energy.

``SccEnergy.apply(coords, calc)`` returns the Mermin free energy (Ha) of the geometry
``coords`` (nAtom x 3, bohr) as a 0-d tensor. Its backward is the library's hand-written
adjoint of the energy (``dftbhip_ctx_gradients`` with the analytic Slater-Koster derivative
kernel): the energy-weighted density matrix is rebuilt on the device from the eigenvectors that
stay resident after the forward pass, so no eigenvector copies are stored on the autograd tape.

The library runs in a worker process (``python -m dftbplus_amd.worker``) because PyTorch
bundles a second HIP runtime. The role of this module is that of tools/pythonapi's
``dftbplus.DftbPlus`` (get_energy / get_gradients, tools/pythonapi/src/dftbplus/dftbplus.py).
"""
import os
import pickle
import subprocess
import sys

import numpy as np
import torch

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class Calculator:
    """Handle to a resident SCC context in a worker process."""

    def __init__(self, params, geometry, options, force_mode: int = 1, precision: int = 0,
                 **kwargs):
        env = dict(os.environ)
        env["PYTHONPATH"] = _ROOT + os.pathsep + env.get("PYTHONPATH", "")
        self._p = subprocess.Popen([sys.executable, "-m", "dftbplus_amd.worker"],
                                   stdin=subprocess.PIPE, stdout=subprocess.PIPE, env=env,
                                   cwd=_ROOT)
        self.natom = geometry.natom
        kw = dict(force_mode=force_mode, precision=precision, **kwargs)
        self.norb = self._call("init", (params, geometry, options, kw))
        self._coords = None
        self.n_evals = 0

    def _call(self, cmd, payload):
        pickle.dump((cmd, payload), self._p.stdin)
        self._p.stdin.flush()
        status, res = pickle.load(self._p.stdout)
        if status != "ok":
            raise RuntimeError(f"dftbhip worker: {res}")
        return res

    def energy(self, coords: np.ndarray) -> float:
        c = np.ascontiguousarray(coords, dtype=np.float64).reshape(self.natom, 3)
        e, q, nit, conv = self._call("energy", c)
        if not conv:
            raise RuntimeError(f"SCC did not converge in {nit} iterations")
        self._coords = c.copy()
        self.charges = q
        self.n_evals += 1
        return e

    def gradients(self, coords: np.ndarray) -> np.ndarray:
        c = np.ascontiguousarray(coords, dtype=np.float64).reshape(self.natom, 3)
        if self._coords is None or not np.array_equal(c, self._coords):
            self.energy(c)
        return self._call("gradients", None)

    def close(self):
        if self._p is not None and self._p.poll() is None:
            try:
                self._call("close", None)
            finally:
                self._p.wait(timeout=60)
        self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SccEnergy(torch.autograd.Function):
    """E(R): forward runs the SCC cycle, backward the adjoint gradient kernel."""

    @staticmethod
    def forward(ctx, coords: torch.Tensor, calc: Calculator) -> torch.Tensor:
        c = coords.detach().to("cpu", torch.float64).numpy()
        e = calc.energy(c)
        ctx.calc = calc
        ctx.coords = c
        ctx.dtype, ctx.device = coords.dtype, coords.device
        return torch.tensor(e, dtype=coords.dtype, device=coords.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        g = ctx.calc.gradients(ctx.coords)
        gt = torch.from_numpy(np.asarray(g)).to(ctx.device, ctx.dtype)
        return grad_out * gt, None


def energy(coords: torch.Tensor, calc: Calculator) -> torch.Tensor:
    return SccEnergy.apply(coords, calc)
