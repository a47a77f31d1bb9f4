"""Shared fixtures: synthetic workloads and their oracle solutions (computed once per session)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs the HIP library on an MI355X")


@pytest.fixture(scope="session")
def cluster216():
    """The 216-atom C/H/O cluster solved by the oracle (SCC, charges, FD gradients)."""
    from dftbplus_amd import synth
    from oracle import dftb_ref as R
    geo, params, opts = synth.workload("cluster216")
    calc = R.Calculator(params, geo, opts)
    res = calc.scc()
    calc.gradients(res)
    return geo, params, opts, calc, res


@pytest.fixture(scope="session")
def diamond2():
    """A 64-atom periodic diamond-like cell (kT = 0.01 Ha) solved by the oracle."""
    from dftbplus_amd import synth
    from oracle import dftb_ref as R
    geo, params, opts = synth.workload("diamond2")
    calc = R.Calculator(params, geo, opts)
    res = calc.scc()
    calc.gradients(res)
    return geo, params, opts, calc, res


def reference_arrays(calc):
    """The oracle's neighbour list in the reference's 1-based Fortran array conventions."""
    from dftbplus_amd import api
    nl, sp = calc.nl, calc.sp
    nat = len(nl.neigh)
    mneigh = max(len(x) for x in nl.neigh) - 1
    ineigh = np.zeros((mneigh + 1, nat), dtype=np.int32, order="F")
    ipair = np.zeros((mneigh + 1, nat), dtype=np.int32, order="F")
    for a in range(nat):
        ineigh[: len(nl.neigh[a]), a] = nl.neigh[a] + 1
    k = 0
    for a in range(nat):
        for j in range(sp.nneigh_sk[a] + 1):
            ipair[j, a] = sp.start[k]
            k += 1
    species_all = calc.geo.species[nl.img2cent] + 1
    return api.RefArrays(nat, len(nl.img2cent), nl.coords_all, species_all, nl.img2cent + 1,
                         ineigh, sp.nneigh_sk, ipair, calc.off + 1, sp.nnz)
