"""End-to-end parity of the resident SCC context with the CPU oracle (FP64 path)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# tolerances of the plain FP64 path: energy per atom, charges, gradients
E_TOL, Q_TOL, F_TOL = 1e-8, 1e-8, 1e-7


@pytest.mark.parametrize("tridiag", [0, 1], ids=["two_stage", "one_stage"])
@pytest.mark.parametrize("system", ["cluster216", "diamond2"])
def test_scc_energy_charges_forces(system, tridiag, request):
    from dftbplus_amd import api
    geo, params, opts, calc, res = request.getfixturevalue(system)
    dp = api.DftbPlus(params, geo, opts, tridiag=tridiag)
    e = dp.get_energy()
    q = dp.get_gross_charges()
    g = dp.get_gradients()
    nat = geo.natom
    print(f"{system}: iterations {dp.n_iter} (oracle {res.n_iter}), E {e:.12f} (oracle "
          f"{res.energy_mermin:.12f})")
    assert dp.converged
    assert abs(e - res.energy_mermin) / nat < E_TOL
    assert np.max(np.abs(q - res.charges)) < Q_TOL
    assert np.max(np.abs(g - res.gradients)) < F_TOL
    for k, v in (("nonscc", res.e_nonscc), ("scc", res.e_scc), ("repulsive", res.e_rep),
                 ("ts", res.ts)):
        assert abs(dp.energies[k] - v) / nat < E_TOL, k
    dp.close()
