"""The neighbour list built on the device against the oracle's (updateNeighbourList order), bit
for bit, and SCC results with the device-built pair and block lists against the host-built ones
(to rounding)."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _geometries():
    from dftbplus_amd import synth
    out = []
    for name in ("cluster27", "diamond1", "diamond2"):
        geo, _, _ = synth.workload(name)
        out.append((name, geo.coords, geo.latvecs))
    rng = np.random.default_rng(11)
    # a small skewed cell: many images per atom
    lat = np.array([[6.0, 0.4, 0.0], [1.1, 5.5, 0.3], [0.2, -0.7, 7.0]])
    frac = rng.uniform(size=(5, 3))
    out.append(("skewed5", frac @ lat, lat))
    return out


@pytest.mark.parametrize("cutoff", [4.0, 10.5, 16.0])
def test_device_list_equals_oracle(cutoff):
    from dftbplus_amd import api
    from oracle import dftb_ref as R
    for name, x, lat in _geometries():
        ref = R.neighbour_list(x, lat, cutoff)
        call, i2c, neigh, d2 = api.neighbour_list(x, lat, cutoff)
        print(f"{name} cutoff {cutoff}: nAll {len(i2c)}, pairs {sum(len(v) for v in neigh)}")
        assert np.array_equal(call, ref.coords_all), name
        assert np.array_equal(i2c, ref.img2cent), name
        assert len(neigh) == len(ref.neigh)
        for a in range(len(neigh)):
            assert np.array_equal(neigh[a], ref.neigh[a]), (name, a)
            assert np.array_equal(d2[a], ref.dist2[a]), (name, a)


def test_atoms_too_close():
    from dftbplus_amd import api
    x = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0], [3.0, 0.0, 0.005]])
    with pytest.raises(api.DftbHipError, match="Atoms 3 and 2 too close|Atoms 2 and 3 too close"):
        api.neighbour_list(x, None, 5.0)


@pytest.mark.parametrize("kind", ["cluster27", "diamond2_gc222"])
def test_scc_with_device_list_equals_host_list(kind):
    from dftbplus_amd import api, synth
    if kind == "cluster27":
        geo, params, opts = synth.workload("cluster27")
    else:
        geo, params, opts = synth.workload("diamond2")
        k = np.array([[a, b, c] for a in (0, 0.5) for b in (0, 0.5) for c in (0, 0.5)])
        opts = dataclasses.replace(opts, kpoints=k)
    out = {}
    try:
        for mode in (1, 0):
            api.set_tuning(3, mode)
            dp = api.DftbPlus(params, geo, opts)
            out[mode] = (dp.get_energy(), dp.get_gross_charges(), dp.get_gradients())
            dp.close()
    finally:
        api.set_tuning(3, 1)
    (e1, q1, g1), (e0, q0, g0) = out[1], out[0]
    print(f"{kind}: device {e1:.15f} host {e0:.15f}")
    # equal lists; the eigensolver's atomically accumulated panel products make two runs agree
    # to rounding, not bit for bit
    assert abs(e1 - e0) / geo.natom < 1e-13
    assert np.max(np.abs(q1 - q0)) < 1e-12
    assert np.max(np.abs(g1 - g0)) < 1e-11


def test_single_and_isolated_atoms():
    """One atom, and atoms farther apart than the cutoff: every list holds the atom alone."""
    from dftbplus_amd import api
    from oracle import dftb_ref as R
    for x in (np.zeros((1, 3)), np.array([[0.0, 0.0, 0.0], [20.0, 0.0, 0.0], [0.0, 25.0, 3.0]])):
        call, i2c, neigh, d2 = api.neighbour_list(x, None, 10.0)
        ref = R.neighbour_list(x, None, 10.0)
        assert np.array_equal(call, ref.coords_all) and np.array_equal(i2c, ref.img2cent)
        for a in range(len(x)):
            assert np.array_equal(neigh[a], [a]) and np.array_equal(d2[a], [0.0])


def test_neighbour_capacity_error():
    """More neighbours of one atom than the LDS sort holds (4096) is reported, not truncated."""
    from dftbplus_amd import api, synth
    geo, _, _ = synth.workload("diamond2")
    with pytest.raises(api.DftbHipError, match="more than 4096"):
        api.neighbour_list(geo.coords, geo.latvecs, 45.0)
