"""k-point SCC against the CPU oracle's complex path, at the FP64 tolerances: complex
Hermitian problems through their real 2n embedding, time-reversal-invariant points (2k
integral) as real n x n problems."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MESHES = {
    # Monkhorst-Pack 2x2x2 (k = +-1/4: complex phases)
    "mp222": np.array([[a, b, c] for a in (-0.25, 0.25) for b in (-0.25, 0.25)
                       for c in (-0.25, 0.25)]),
    # Gamma-centred 2x2x2 (k = 0, 1/2: the eight time-reversal-invariant points, real H(k))
    "gc222": np.array([[a, b, c] for a in (0, 0.5) for b in (0, 0.5) for c in (0, 0.5)]),
    # Gamma-centred 3x3x3 (k = 0, 1/3, 2/3: Gamma real, the others complex)
    "gc333": np.array([[a, b, c] for a in (0, 1 / 3, 2 / 3) for b in (0, 1 / 3, 2 / 3)
                       for c in (0, 1 / 3, 2 / 3)]),
}


@pytest.mark.parametrize("mesh", sorted(MESHES))
def test_kpoint_scc_matches_oracle(mesh):
    from dftbplus_amd import api, synth
    from oracle import dftb_ref as R
    geo, params, opts = synth.workload("diamond1")
    o = dataclasses.replace(opts, kpoints=MESHES[mesh])
    calc = R.Calculator(params, geo, o)
    res = calc.scc()
    og = calc.gradients(res)
    dp = api.DftbPlus(params, geo, o)
    e, q, g = dp.get_energy(), dp.get_gross_charges(), dp.get_gradients()
    nat = geo.natom
    print(f"{mesh}: it {dp.n_iter}/{res.n_iter} dE/atom {abs(e - res.energy_mermin) / nat:.2e} "
          f"dq {np.max(np.abs(q - res.charges)):.2e} dF {np.max(np.abs(g - og)):.2e}")
    assert dp.converged and res.converged
    assert abs(e - res.energy_mermin) / nat < 1e-8
    assert np.max(np.abs(q - res.charges)) < 1e-8
    assert np.max(np.abs(g - og)) < 1e-7
    for k, v in (("nonscc", res.e_nonscc), ("scc", res.e_scc), ("ts", res.ts),
                 ("eband", res.eband)):
        assert abs(dp.energies[k] - v) / nat < 1e-8, k
    dp.close()


def test_single_gamma_kpoint_equals_real_path():
    from dftbplus_amd import api, synth
    geo, params, opts = synth.workload("diamond1")
    d0 = api.DftbPlus(params, geo, opts)
    d1 = api.DftbPlus(params, geo, dataclasses.replace(opts, kpoints=np.zeros((1, 3))))
    assert abs(d0.get_energy() - d1.get_energy()) < 1e-10
    assert np.max(np.abs(d0.get_gross_charges() - d1.get_gross_charges())) < 1e-10
    assert np.max(np.abs(d0.get_gradients() - d1.get_gradients())) < 1e-9
    d0.close()
    d1.close()


def test_kpoints_together_equal_one_by_one():
    """Eight real points of 247 orbitals tridiagonalised in one pass (one launch per column
    kernel for all points) give the results of the point-by-point eigensolves."""
    from dftbplus_amd import api, synth
    geo, params, opts = synth.workload("diamond2")
    o = dataclasses.replace(opts, kpoints=MESHES["gc222"])
    out = {}
    try:
        for mode in (1, 0):
            api.set_tuning(2, mode)
            dp = api.DftbPlus(params, geo, o)
            out[mode] = (dp.get_energy(), dp.get_gross_charges(), dp.get_gradients(),
                         dp.n_iter, dp.converged)
            dp.close()
    finally:
        api.set_tuning(2, 1)
    (e1, q1, g1, n1, c1), (e0, q0, g0, n0, c0) = out[1], out[0]
    print(f"together {e1:.14f} ({n1} it), one by one {e0:.14f} ({n0} it)")
    assert c1 and c0 and n1 == n0
    assert abs(e1 - e0) / geo.natom < 1e-13
    assert np.max(np.abs(q1 - q0)) < 1e-12
    assert np.max(np.abs(g1 - g0)) < 1e-11


@pytest.fixture(scope="module")
def amorphous64():
    """A 64-atom cell of the amorphous C/Si/H/O network recipe (seed 4, 2.3 g/cm^3, 247
    orbitals): the config-5 workload at a size the oracle solves in seconds per mesh."""
    from dftbplus_amd import synth
    return synth.workload("amorphous64")


@pytest.mark.parametrize("mesh", ["gc222", "mp222"])
def test_amorphous_network_kpoints_match_oracle(amorphous64, mesh):
    """The config-5 recipe (C/Si/H/O, the only workload with Si) with the Gamma-centred 2x2x2
    mesh (eight real problems, tridiagonalised together) and the Monkhorst-Pack 2x2x2 mesh
    (k = +-1/4, genuinely complex) against the oracle's complex path at the FP64 bounds."""
    from dftbplus_amd import api
    from oracle import dftb_ref as R
    geo, params, opts = amorphous64
    o = dataclasses.replace(opts, kpoints=MESHES[mesh])
    calc = R.Calculator(params, geo, o)
    res = calc.scc()
    og = calc.gradients(res)
    dp = api.DftbPlus(params, geo, o)
    e, q, g = dp.get_energy(), dp.get_gross_charges(), dp.get_gradients()
    nat = geo.natom
    print(f"amorphous64 {mesh}: it {dp.n_iter}/{res.n_iter} dE/atom "
          f"{abs(e - res.energy_mermin) / nat:.2e} dq {np.max(np.abs(q - res.charges)):.2e} "
          f"dF {np.max(np.abs(g - og)):.2e}")
    assert dp.converged and res.converged
    assert abs(e - res.energy_mermin) / nat < 1e-8
    assert np.max(np.abs(q - res.charges)) < 1e-8
    assert np.max(np.abs(g - og)) < 1e-7
    dp.close()


def test_md_steps_warm_start_equal_cold(amorphous64):
    """MD-like steps (every coordinate moved by up to +-0.05 bohr per step) with the SCC
    warm-started from the previous step's input charges reach the fixed points of cold starts
    (to the SCC tolerance) in fewer iterations."""
    from dftbplus_amd import api
    geo, params, opts = amorphous64
    o = dataclasses.replace(opts, kpoints=MESHES["gc222"])
    rng = np.random.default_rng(4)
    warm = api.DftbPlus(params, geo, o)
    warm.set_charge_restart(True)
    warm.get_energy()
    coords = geo.coords.copy()
    it_w, it_c = [], []
    for step in range(3):
        coords = coords + rng.uniform(-0.05, 0.05, coords.shape)
        warm.set_geometry(coords, geo.latvecs)
        ew, qw, gw = warm.get_energy(), warm.get_gross_charges(), warm.get_gradients()
        cold = api.DftbPlus(params, type(geo)(coords, geo.species, geo.latvecs), o)
        ec, qc, gc = cold.get_energy(), cold.get_gross_charges(), cold.get_gradients()
        it_w.append(warm.n_iter)
        it_c.append(cold.n_iter)
        print(f"step {step}: warm {warm.n_iter} it, cold {cold.n_iter} it, dE {abs(ew - ec):.2e}")
        assert warm.converged and cold.converged
        assert abs(ew - ec) / geo.natom < 1e-9
        assert np.max(np.abs(qw - qc)) < 5e-8
        assert np.max(np.abs(gw - gc)) < 1e-7
        cold.close()
    assert sum(it_w) < sum(it_c)
    warm.close()


def test_amorphous_kpoint_tail_matches_oracle(amorphous64):
    """Mixed precision at k-points (Gamma-centred 2x2x2, every point time-reversal invariant):
    plain FP64 iterations down to a charge error of 1e-3, then per point two Ogita-Aishima steps
    on its previous eigenpairs against its sparse H(k) and S(k) (block phases +-1), no
    eigensolve: the oracle's complex-path fixed point at the FP64 bounds."""
    from dftbplus_amd import api
    from oracle import dftb_ref as R
    geo, params, opts = amorphous64
    o = dataclasses.replace(opts, kpoints=MESHES["gc222"], max_scc_iter=200)
    calc = R.Calculator(params, geo, o)
    res = calc.scc()
    og = calc.gradients(res)
    dp = api.DftbPlus(params, geo, o, precision=1, refine_below=1e-3, refine_steps=0,
                      tail_below=1e-3, tail_steps=2)
    e, q, g = dp.get_energy(), dp.get_gross_charges(), dp.get_gradients()
    t = dp.timings()
    nat = geo.natom
    print(f"amorphous64 k-point tail: it {dp.n_iter}/{res.n_iter} dE/atom "
          f"{abs(e - res.energy_mermin) / nat:.2e} dq {np.max(np.abs(q - res.charges)):.2e} "
          f"dF {np.max(np.abs(g - og)):.2e} refine {t['refine']:.3f}s")
    assert dp.converged and res.converged
    assert t["refine"] > 0.0
    assert abs(e - res.energy_mermin) / nat < 1e-8
    assert np.max(np.abs(q - res.charges)) < 1e-8
    assert np.max(np.abs(g - og)) < 1e-7
    dp.close()
