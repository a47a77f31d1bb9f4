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


def test_input_charges_restart(cluster216):
    """ReadInitialCharges: a cycle started from the input charges another cycle ended with (what
    the reference writes to its charge file) converges at once to the same point; charges that
    do not sum to the electron count are refused with the reference's message."""
    from dftbplus_amd import api
    geo, params, opts, calc, res = cluster216
    dp = api.DftbPlus(params, geo, opts)
    e = dp.get_energy()
    qin = dp.get_input_charges()
    nat = geo.natom
    assert dp.converged and dp.n_iter > 5
    dp2 = api.DftbPlus(params, geo, opts)
    dp2.set_input_charges(qin)
    e2 = dp2.get_energy()
    print(f"restart: {dp2.n_iter} iterations (cold {dp.n_iter}), dE/atom {abs(e2 - e) / nat:.2e}")
    assert dp2.converged and dp2.n_iter <= 3
    assert abs(e2 - e) / nat < E_TOL
    assert np.max(np.abs(dp2.get_gross_charges() - dp.get_gross_charges())) < Q_TOL
    bad = qin.copy()
    bad[0] += 0.01
    with pytest.raises(api.DftbHipError, match="External file of charges has a total charge"):
        dp2.set_input_charges(bad)
    dp.close()
    dp2.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_sytrd_graph_replay_matches_launches(diamond2, precision):
    """The tridiagonalisation replayed from its captured HIP graph (every SCC iteration after
    the first on the same buffers) against the same cycle launched kernel by kernel: the same
    iterations, energies, charges and gradients to rounding (FP64 and mixed paths)."""
    from dftbplus_amd import api
    geo, params, opts, calc, res = diamond2
    out = []
    for graphs in (1, 0):
        api.set_tuning(12, graphs)
        try:
            dp = api.DftbPlus(params, geo, opts, precision=precision)
            out.append((dp.get_energy(), dp.get_gross_charges(), dp.get_gradients(), dp.n_iter))
            dp.close()
        finally:
            api.set_tuning(12, 0)
    (e1, q1, g1, n1), (e0, q0, g0, n0) = out
    print(f"graphs: {n1} iterations, launches: {n0}; dE {abs(e1 - e0):.2e}")
    assert n1 == n0
    assert abs(e1 - e0) < 1e-10 * geo.natom
    assert np.max(np.abs(q1 - q0)) < 1e-10 and np.max(np.abs(g1 - g0)) < 1e-10
