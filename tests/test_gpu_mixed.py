"""Mixed-precision SCC (FP32 MFMA bulk, FP64 Cholesky / divide and conquer / reductions) against
the CPU oracle, with and without the FP64 Ogita-Aishima refinement of the eigenpairs."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INF = float("inf")


@pytest.mark.parametrize("system", ["cluster216", "diamond2"])
def test_mixed_refined_matches_fp64_tolerances(system, request):
    """Two refinement steps on every iteration: the FP64 tolerances of the plain path hold
    (energy 1e-8 Ha/atom, charges 1e-8 e, gradients 1e-7 Ha/bohr)."""
    from dftbplus_amd import api
    geo, params, opts, calc, res = request.getfixturevalue(system)
    dp = api.DftbPlus(params, geo, opts, precision=1, refine_below=INF, refine_steps=2)
    e, q, g = dp.get_energy(), dp.get_gross_charges(), dp.get_gradients()
    nat = geo.natom
    de, dq, dg = abs(e - res.energy_mermin) / nat, np.max(np.abs(q - res.charges)), \
        np.max(np.abs(g - res.gradients))
    print(f"{system} mixed+OA2: it {dp.n_iter} dE/atom {de:.2e} dq {dq:.2e} dF {dg:.2e}")
    assert dp.converged
    assert de < 1e-8 and dq < 1e-8 and dg < 1e-7
    dp.close()


@pytest.mark.parametrize("system", ["cluster216", "diamond2"])
def test_mixed_fp32_error_bounds(system, request):
    """FP32 bulk without refinement: the SCC stalls at the FP32 noise floor of the charges
    (about 1e-5 e here), so the state after the iteration limit is compared: energy within
    1e-6 Ha/atom, charges within 1e-4 e, gradients within 1e-5 Ha/bohr of the FP64 oracle."""
    from dftbplus_amd import api
    geo, params, opts, calc, res = request.getfixturevalue(system)
    o = dataclasses.replace(opts, scc_tol=1e-5)
    dp = api.DftbPlus(params, geo, o, precision=1, refine_below=0.0)
    e, q, g = dp.get_energy(), dp.get_gross_charges(), dp.get_gradients()
    nat = geo.natom
    de, dq, dg = abs(e - res.energy_mermin) / nat, np.max(np.abs(q - res.charges)), \
        np.max(np.abs(g - res.gradients))
    print(f"{system} mixed FP32: it {dp.n_iter} dE/atom {de:.2e} dq {dq:.2e} dF {dg:.2e}")
    assert de < 1e-6 and dq < 1e-4 and dg < 1e-5
    dp.close()


def test_mixed_switch_converges_to_fp64(cluster216):
    """Default switch (FP32 iterations until the charge error drops below 1e-4, then two
    refinement steps per iteration): converges to the SCC tolerance and meets the FP64
    tolerances."""
    from dftbplus_amd import api
    geo, params, opts, calc, res = cluster216
    dp = api.DftbPlus(params, geo, opts, precision=1, refine_below=1e-4, refine_steps=2)
    e, q, g = dp.get_energy(), dp.get_gross_charges(), dp.get_gradients()
    t = dp.timings()
    nat = geo.natom
    print(f"switch: it {dp.n_iter} refine time {t['refine']:.3f}s")
    assert dp.converged
    assert abs(e - res.energy_mermin) / nat < 1e-8
    assert np.max(np.abs(q - res.charges)) < 1e-8
    assert np.max(np.abs(g - res.gradients)) < 1e-7
    assert t["refine"] > 0.0
    dp.close()


def test_mixed_then_fp64_converges_to_fp64(cluster216):
    """refine_steps = 0: FP32 iterations until the charge error drops below 1e-3, then plain
    FP64 iterations; the converged result meets the FP64 tolerances."""
    from dftbplus_amd import api
    geo, params, opts, calc, res = cluster216
    dp = api.DftbPlus(params, geo, opts, precision=1, refine_below=1e-3, refine_steps=0)
    e, q, g = dp.get_energy(), dp.get_gross_charges(), dp.get_gradients()
    t = dp.timings()
    nat = geo.natom
    print(f"FP32 then FP64: it {dp.n_iter} (oracle {res.n_iter})")
    assert dp.converged
    assert abs(e - res.energy_mermin) / nat < 1e-8
    assert np.max(np.abs(q - res.charges)) < 1e-8
    assert np.max(np.abs(g - res.gradients)) < 1e-7
    assert t["refine"] == 0.0
    dp.close()
