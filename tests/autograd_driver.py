"""Gradient checks of the torch.autograd interface (run as a child process of the GPU test:
the test process holds the library's HIP runtime, this one loads PyTorch's).

Prints one JSON line: analytic-adjoint gradients from ``SccEnergy.backward`` against central
finite differences of the energy (h = 1e-4 bohr) on 48 random coordinates, against the
finite-difference Slater-Koster derivative path (the reference's default force evaluation) and
against the CPU oracle."""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import torch  # noqa: E402

from dftbplus_amd import synth, torch_api  # noqa: E402


def main(name: str, ncheck: int, precision: int):
    geo, params, opts = synth.workload(name)
    kw = {} if precision == 0 else dict(refine_below=float("inf"), refine_steps=2)
    calc = torch_api.Calculator(params, geo, opts, force_mode=1, precision=precision, **kw)
    t0 = time.time()
    x = torch.tensor(geo.coords, dtype=torch.float64, requires_grad=True)
    e = torch_api.energy(x, calc)
    e.backward()
    g = x.grad.numpy().copy()
    t_fwd_bwd = time.time() - t0
    h = 1e-4
    rng = np.random.default_rng(48)
    idx = rng.choice(3 * geo.natom, size=min(ncheck, 3 * geo.natom), replace=False)
    c0 = geo.coords.reshape(-1).copy()
    fd = []
    for k in idx:
        c = c0.copy()
        c[k] += h
        ep = calc.energy(c)
        c[k] -= 2 * h
        em = calc.energy(c)
        fd.append((ep - em) / (2 * h))
    fd = np.array(fd)
    calc.close()
    # the reference's default: finite-difference Slater-Koster derivatives
    ref = torch_api.Calculator(params, geo, opts, force_mode=0, precision=precision, **kw)
    gfd = ref.gradients(geo.coords)
    ref.close()
    out = {"system": name, "natom": geo.natom, "precision": precision,
           "energy": float(e.item()), "fwd_bwd_s": t_fwd_bwd,
           "max_abs_grad": float(np.max(np.abs(g))),
           "fd_max_err": float(np.max(np.abs(g.reshape(-1)[idx] - fd))),
           "sk_fd_path_max_diff": float(np.max(np.abs(g - gfd)))}
    if geo.natom <= 200:
        from oracle import dftb_ref as R
        oc = R.Calculator(params, geo, opts)
        res = oc.scc()
        og = oc.gradients(res)
        out["oracle_energy_diff"] = abs(float(e.item()) - res.energy_mermin)
        out["oracle_max_diff"] = float(np.max(np.abs(g - og)))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]), int(sys.argv[3]) if len(sys.argv) > 3 else 0)
