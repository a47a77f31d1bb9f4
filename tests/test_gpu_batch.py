"""Molecule batch (one block-diagonal system, batched Jacobi eigensolver) against a loop of
single-molecule calculations of the plain path and against the CPU oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch24():
    from dftbplus_amd import api, synth
    mols, params, opts = synth.workload("molecules24")
    b = api.MoleculeBatch(params, mols, opts)
    e = b.scc()
    g = b.gradients()
    yield mols, params, opts, b, e, g
    b.close()


def test_batch_matches_single_molecule_loop(batch24):
    from dftbplus_amd import api
    mols, params, opts, b, e, g = batch24
    nconv = 0
    for m, geo in enumerate(mols):
        dp = api.DftbPlus(params, geo, opts)
        es = dp.get_energy()
        qs = dp.get_gross_charges()
        gs = dp.get_gradients()
        print(f"mol {m}: {geo.natom} atoms, it {dp.n_iter}/{b.n_iter[m]}, dE {abs(es - e[m]):.2e}")
        assert dp.converged == b.converged[m]
        if not dp.converged:
            dp.close()
            continue
        nconv += 1
        nat = geo.natom
        assert abs(es - e[m]) / nat < 1e-8
        assert np.max(np.abs(qs - b.charges[m])) < 5e-8
        assert np.max(np.abs(gs - g[m])) < 1e-7
        dp.close()
    assert nconv >= 20


def test_batch_matches_oracle(batch24):
    from oracle import dftb_ref as R
    mols, params, opts, b, e, g = batch24
    for m in (3, 7, 17, 19):
        calc = R.Calculator(params, mols[m], opts)
        res = calc.scc()
        og = calc.gradients(res)
        assert res.converged and b.converged[m]
        nat = mols[m].natom
        assert abs(res.energy_mermin - e[m]) / nat < 1e-8
        assert np.max(np.abs(res.charges - b.charges[m])) < 1e-8
        assert np.max(np.abs(og - g[m])) < 1e-7
