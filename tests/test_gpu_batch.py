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


def _chain(norb, rng):
    """A molecule of exactly norb orbitals (H: 1, C/N/O: 4): a helical chain, bonds 2.4-2.6
    bohr, no two atoms closer than 1.9 bohr."""
    from dftbplus_amd.model import Geometry
    nheavy = norb // 4
    spec = [int(rng.integers(1, 4)) for _ in range(nheavy)] + [0] * (norb - 4 * nheavy)
    rng.shuffle(spec)
    k = np.arange(len(spec))
    pos = np.stack([2.3 * k, 0.9 * np.sin(0.9 * k), 0.9 * np.cos(0.9 * k)], axis=1)
    pos += rng.uniform(-0.05, 0.05, size=pos.shape)
    return Geometry(pos, np.array(spec, dtype=np.int64), None)


def test_batch_size_classes_one_iteration():
    """Molecules at the edges of the eigensolver's size classes (LDS A and C up to 64 and 112
    orbitals, LDS A up to 190, HBM above), single atoms and odd sizes included, after one SCC
    iteration: the batch equals the single-molecule runs of the plain path."""
    import dataclasses
    from dftbplus_amd import api, synth
    _, params, opts = synth.workload("molecules1")
    o = dataclasses.replace(opts, max_scc_iter=1)
    rng = np.random.default_rng(21)
    sizes = [1, 2, 5, 63, 64, 65, 111, 112, 113, 189, 190, 191, 223]
    mols = [_chain(n, rng) for n in sizes]
    b = api.MoleculeBatch(params, mols, o)
    e = b.scc()
    g = b.gradients()
    for m, geo in enumerate(mols):
        dp = api.DftbPlus(params, geo, o)
        es, qs, gs = dp.get_energy(), dp.get_gross_charges(), dp.get_gradients()
        dp.close()
        print(f"{sizes[m]} orbitals: dE {abs(es - e[m]):.2e} dq {np.max(np.abs(qs - b.charges[m])):.2e}")
        assert abs(es - e[m]) / geo.natom < 1e-10
        assert np.max(np.abs(qs - b.charges[m])) < 1e-9
        assert np.max(np.abs(gs - g[m])) < 1e-8
    b.close()


def test_batch_level_at_zero_converges():
    """A molecule with an eigenvalue at 0 Ha (every on-site energy shifted by the constant a
    secant search on the oracle finds) converges like any other: the Jacobi rotation test has a
    rounding-level floor, so a pair touching a diagonal entry near 0 does not rotate on noise
    until the sweep limit (which would now be reported as an error). The batch equals the
    oracle for that molecule and a neighbour."""
    import copy
    from dftbplus_amd import api, synth
    from oracle import dftb_ref as R
    mols, params, opts = synth.workload("molecules8")
    t = 2

    def shifted(c):
        p = copy.deepcopy(params)
        for sp in p.species:
            sp.onsite = [x + c for x in sp.onsite]
        return p

    def level(c, k=None):
        res = R.Calculator(shifted(c), mols[t], opts).scc()
        e = res.eigenvalues
        if k is None:
            k = int(np.argmin(np.abs(e)))
        return e[k], k

    f0, k = level(0.0)
    c0, c1 = 0.0, -f0
    for _ in range(8):
        f1, _ = level(c1, k)
        if abs(f1) < 1e-12:
            break
        c0, c1, f0 = c1, c1 - f1 * (c1 - c0) / (f1 - f0), f1
    assert abs(f1) < 1e-10
    p = shifted(c1)
    b = api.MoleculeBatch(p, mols, opts)
    e = b.scc()
    g = b.gradients()
    for m in (t, 3):
        calc = R.Calculator(p, mols[m], opts)
        res = calc.scc()
        og = calc.gradients(res)
        assert res.converged and b.converged[m]
        assert abs(res.energy_mermin - e[m]) / mols[m].natom < 1e-8
        assert np.max(np.abs(res.charges - b.charges[m])) < 1e-8
        assert np.max(np.abs(og - g[m])) < 1e-7
    b.close()


def _batch_scc(mode, params, mols, opts):
    from dftbplus_amd import api
    api.set_tuning(4, mode)
    try:
        b = api.MoleculeBatch(params, mols, opts)
        e = b.scc()
        g = b.gradients()
        out = (e.copy(), [q.copy() for q in b.charges], g, b.n_iter.copy(), b.converged.copy())
        b.close()
    finally:
        api.set_tuning(4, 2)
    return out


def test_batch_device_scc_equals_host_scc():
    """Fermi levels, energies, convergence tests and Broyden mixing as device kernels (modes 1
    and 2) give the results of the host-thread implementation (mode 0) to rounding: the Fermi
    sums are combined in another order."""
    from dftbplus_amd import synth
    mols, params, opts = synth.workload("molecules24")
    e0, q0, g0, n0, c0 = _batch_scc(0, params, mols, opts)
    for mode in (1, 2):
        e1, q1, g1, n1, c1 = _batch_scc(mode, params, mols, opts)
        assert np.array_equal(c0, c1)
        assert np.max(np.abs(n1 - n0)) <= 1
        for m in range(len(mols)):
            if not c0[m]:
                continue
            assert abs(e1[m] - e0[m]) / mols[m].natom < 1e-10, (mode, m)
            assert np.max(np.abs(q1[m] - q0[m])) < 1e-8, (mode, m)
            assert np.max(np.abs(g1[m] - g0[m])) < 1e-8, (mode, m)


def test_batch_graph_replays_equal_eager_launches():
    """Every replay of the captured SCC iteration (mode 2) equals the same iteration launched
    kernel by kernel (mode 1), with the host reading the state between replays. The Mulliken
    and force sums add blocks with floating-point atomics, whose order varies from run to run,
    so two runs differ by rounding, which the early (not yet contracting) SCC iterations
    amplify: every step agrees to 1e-7, the converged energies to 1e-10 (relative) and the
    iteration counts to one."""
    from dftbplus_amd import api, synth
    mols, params, opts = synth.workload("molecules24")
    try:
        api.set_tuning(4, 1)
        eager = api.MoleculeBatch(params, mols, opts)
        eager.scc_begin()
        api.set_tuning(4, 2)
        graph = api.MoleculeBatch(params, mols, opts)
        graph.scc_begin()
    finally:
        api.set_tuning(4, 2)
    steps = 0
    na_e = na_g = 1
    while True:
        # a batch whose molecules have all converged takes no further step
        na_e = eager.scc_step() if na_e > 0 else 0
        na_g = graph.scc_step() if na_g > 0 else 0
        steps += 1
        ee = eager.scc_end()
        eg = graph.scc_end()
        assert np.max(np.abs(ee - eg) / np.maximum(1.0, np.abs(ee))) < 1e-7, steps
        for qa, qb in zip(eager.charges, graph.charges):
            assert np.max(np.abs(qa - qb)) < 1e-7, steps
        assert np.max(np.abs(eager.n_iter - graph.n_iter)) <= 1
        if (na_e == 0 and na_g == 0) or steps >= opts.max_scc_iter:
            break
    assert steps > 5
    conv = eager.converged & graph.converged
    assert np.max(np.abs(ee - eg)[conv] / np.abs(ee)[conv]) < 1e-10
    ge, gg = eager.gradients(), graph.gradients()
    for m, (a, b) in enumerate(zip(ge, gg)):
        if conv[m]:
            assert np.max(np.abs(a - b)) < 1e-7
    eager.close()
    graph.close()
