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


def _fp64_and_switch(system, **switch):
    """The plain FP64 path and a switching mixed path, both converged to 1e-10 (a cycle that
    takes a different path, here through FP32 iterations, agrees with another only to about
    its SCC tolerance); the FP64 path itself is pinned to the oracle by test_gpu_scc."""
    from dftbplus_amd import api
    geo, params, opts, calc, res = system
    o = dataclasses.replace(opts, scc_tol=1e-10, max_scc_iter=200)
    ref = api.DftbPlus(params, geo, o)
    e0, q0, g0 = ref.get_energy(), ref.get_gross_charges(), ref.get_gradients()
    ref.close()
    dp = api.DftbPlus(params, geo, o, precision=1, **switch)
    e, q, g = dp.get_energy(), dp.get_gross_charges(), dp.get_gradients()
    t = dp.timings()
    nat = geo.natom
    de, dq, dg = abs(e - e0) / nat, np.max(np.abs(q - q0)), np.max(np.abs(g - g0))
    print(f"switch {switch}: it {dp.n_iter} dE/atom {de:.2e} dq {dq:.2e} dF {dg:.2e} "
          f"refine {t['refine']:.3f}s")
    assert dp.converged
    assert de < 1e-8 and dq < 1e-8 and dg < 1e-7
    dp.close()
    return t


def test_mixed_switch_converges_to_fp64(cluster216):
    """FP32 iterations until the charge error drops below 1e-4, then two refinement steps per
    iteration: converges to the FP64 fixed point within the FP64 tolerances."""
    t = _fp64_and_switch(cluster216, refine_below=1e-4, refine_steps=2)
    assert t["refine"] > 0.0


def test_mixed_then_fp64_converges_to_fp64(cluster216):
    """refine_steps = 0: FP32 iterations until the charge error drops below 1e-3, then plain
    FP64 iterations; converges to the FP64 fixed point within the FP64 tolerances."""
    t = _fp64_and_switch(cluster216, refine_below=1e-3, refine_steps=0)
    assert t["refine"] == 0.0


def test_mixed_tail_refinement_converges_to_fp64(cluster216):
    """Three phases: FP32 iterations above a charge error of 1e-3, FP32 eigenpairs refined by
    two Ogita-Aishima steps down to 1e-5, and below that no eigensolve at all: one step refines
    the previous iteration's FP64 eigenpairs against the new H. Converges to the FP64 fixed
    point within the FP64 tolerances."""
    t = _fp64_and_switch(cluster216, refine_below=1e-3, refine_steps=2, tail_below=1e-5)
    assert t["refine"] > 0.0


def test_mixed_tail_refinement_periodic(diamond2):
    """The same three-phase cycle on the 64-atom periodic cell at kT = 0.01 Ha (many
    partially filled levels near the Fermi level) against the FP64 oracle."""
    from dftbplus_amd import api
    geo, params, opts, calc, res = diamond2
    dp = api.DftbPlus(params, geo, opts, precision=1, refine_below=1e-3, refine_steps=2,
                      tail_below=1e-5)
    e, q, g = dp.get_energy(), dp.get_gross_charges(), dp.get_gradients()
    nat = geo.natom
    de, dq, dg = abs(e - res.energy_mermin) / nat, np.max(np.abs(q - res.charges)), \
        np.max(np.abs(g - res.gradients))
    print(f"diamond2 tail: it {dp.n_iter}/{res.n_iter} dE/atom {de:.2e} dq {dq:.2e} dF {dg:.2e}")
    assert dp.converged
    assert de < 1e-8 and dq < 1e-8 and dg < 1e-7
    dp.close()


@pytest.mark.parametrize("system", ["cluster216", "diamond2"])
def test_mixed_early_tail_matches_fp64(system, request):
    """FP32 iterations above a charge error of 1e-3, one iteration whose FP32 eigenpairs are
    refined by two Ogita-Aishima steps, and every later iteration a step on the previous
    eigenpairs against the new sparse H and S (no eigensolve): the cycle reaches the FP64
    path's fixed point within the FP64 tolerances (both converged to 1e-10; cycles converged
    to 1e-8 along different paths agree only to about 1e-8 e). The periodic cell has atoms that
    neighbour their own images, so the block-sparse products must sum image blocks as the
    dense unpack does."""
    t = _fp64_and_switch(request.getfixturevalue(system), refine_below=1e-3, refine_steps=2,
                         tail_below=1e-3)
    assert t["refine"] > 0.0


@pytest.mark.parametrize("band", [1, -1])
@pytest.mark.parametrize("system", ["cluster216", "diamond2"])
def test_tail_residual_f16_matches_fp64(system, band, request):
    """The tail with the refinement in residual form (the default): X^T S X, the Rayleigh quotients and
    the residuals H X - S X diag(lam) in FP64, the products with the residuals and with the
    correction in FP16 (FP32 sums). The FP16 rounding only slows the contraction of a step, so
    the cycle reaches the FP64 path's fixed point within the FP64 tolerances. X^T S X is formed
    and X^T H X on their diagonal tile band with the in-band couplings in FP64 and removed from
    the residuals (band 1, the default), or X^T S X as the whole lower triangle (-1)."""
    from dftbplus_amd import api
    api.set_tuning(10, band)
    try:
        t = _fp64_and_switch(request.getfixturevalue(system), refine_below=1e-3, refine_steps=0,
                             tail_below=1e-3, tail_steps=2)
    finally:
        api.set_tuning(10, 1)
    assert t["refine"] > 0.0


def test_tail_residual_then_fp64_form(cluster216):
    """Residual form above a charge error of 1e-6 e, the FP64 form below (tuning parameter
    11 = 6): the FP64 path's fixed point within the FP64 tolerances."""
    from dftbplus_amd import api
    api.set_tuning(11, 6)
    try:
        t = _fp64_and_switch(cluster216, refine_below=1e-3, refine_steps=0, tail_below=1e-3,
                             tail_steps=2)
    finally:
        api.set_tuning(11, 0)
    assert t["refine"] > 0.0


@pytest.mark.parametrize("system", ["cluster216", "diamond2"])
def test_tail_fp64_seed_two_steps_matches_fp64(system, request):
    """The tail schedule the benchmark runs: FP32 iterations above 1e-3, one plain FP64
    iteration that keeps all eigenvectors, then two Ogita-Aishima steps per iteration on the
    sparse H and S, here with X^T H X and X^T S X as FP64 GEMMs and the correction product in
    FP32 (tuning parameter 8 = 1): the FP64 path's fixed point within the FP64 tolerances (both
    converged to 1e-10, since cycles along different paths agree only to about their SCC
    tolerance; the FP64 path is pinned to the oracle by test_gpu_scc)."""
    from dftbplus_amd import api
    api.set_tuning(8, 1)
    try:
        t = _fp64_and_switch(request.getfixturevalue(system), refine_below=1e-3, refine_steps=0,
                             tail_below=1e-3, tail_steps=2)
    finally:
        api.set_tuning(8, 2)
    assert t["refine"] > 0.0
