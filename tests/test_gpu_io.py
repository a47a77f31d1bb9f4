"""A DFTB+ input directory (dftb_in.hsd, geo.gen, SK files) through the input reader, the
resident SCC context and the tagged-output writers, against the same system run directly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(kind):
    from dftbplus_amd import synth
    from dftbplus_amd.model import Options
    if kind == "cluster27":
        geo, el = synth.cluster_216(0, 3)
        params, files = synth.make_parameters(el, 100)
        return geo, el, params, files, Options(kT=1.0e-8, scc_tol=1.0e-9)
    geo, el = synth.diamond_supercell(1, 1)
    params, files = synth.make_parameters(el, 101)
    k = np.array([[a, b, c] for a in (0, 0.5) for b in (0, 0.5) for c in (0, 0.5)])
    return geo, el, params, files, Options(kT=0.01, scc_tol=1.0e-9, kpoints=k,
                                           kweights=np.full(8, 1 / 8))


@pytest.mark.parametrize("kind", ["cluster27", "diamond1_gc222"])
def test_input_directory_single_point(kind, tmp_path):
    from dftbplus_amd import api, hsd, tagged
    geo, el, params, files, opts = _case(kind)
    hsd.export_input(str(tmp_path), geo, el, files, opts)
    res = hsd.run(str(tmp_path))
    dp = api.DftbPlus(params, geo, opts)
    e, q, g = dp.get_energy(), dp.get_gross_charges(), dp.get_gradients()
    dp.close()
    nat = geo.natom
    print(f"{kind}: E {res['mermin_energy']:.12f} direct {e:.12f}")
    assert res["converged"]
    assert abs(res["mermin_energy"] - e) / nat < 1e-10
    assert np.max(np.abs(res["gross_charges"] - q)) < 1e-9
    assert np.max(np.abs(res["forces"] + g)) < 1e-8
    with open(tmp_path / "autotest.tag") as fh:
        auto = dict(tagged.parse_tags(fh.read()))
    assert abs(auto["mermin_energy"] - res["mermin_energy"]) < 1e-14 * abs(e) + 1e-15
    np.testing.assert_allclose(auto["forces"].T, res["forces"], rtol=1e-14, atol=1e-18)
    np.testing.assert_allclose(auto["end_coords"].T, geo.coords, atol=1e-12)
    pops = auto["orbital_charges"]
    assert pops.shape == (4, nat)
    assert abs(pops.sum() - res["n_electrons"]) < 1e-8
    with open(tmp_path / "results.tag") as fh:
        resu = dict(tagged.parse_tags(fh.read()))
    nk = 1 if opts.kpoints is None else len(opts.kpoints)
    assert resu["eigenvalues"].shape == (dp_norb(params, geo), nk, 1)
    np.testing.assert_allclose(resu["gross_atomic_charges"], res["gross_charges"], rtol=1e-14,
                               atol=1e-15)
    assert abs(resu["number_of_electrons"][0] - res["n_electrons"]) == 0.0
    w = np.ones(1) if opts.kweights is None else opts.kweights
    assert abs(np.sum(resu["filling"][:, :, 0] @ w) - res["n_electrons"]) < 1e-8
    if geo.latvecs is not None:
        assert abs(auto["cell_volume"] - abs(np.linalg.det(geo.latvecs))) < 1e-10


def dp_norb(params, geo):
    return int(sum(params.species[s].norb for s in geo.species))


@pytest.mark.parametrize("ascii", [False, True])
def test_read_initial_charges(ascii, tmp_path):
    """WriteCharges / ReadInitialCharges through the input directory: the first run writes the
    charge file (charges.bin by default, charges.dat with WriteChargesAsText), a second run with
    ReadInitialCharges = Yes reads it and is converged within three iterations, at the same
    energy and charges."""
    from dftbplus_amd import hsd
    geo, el, params, files, opts = _case("cluster27")
    hsd.export_input(str(tmp_path), geo, el, files, opts)
    inp = tmp_path / "dftb_in.hsd"
    text = inp.read_text()
    extra = "  WriteChargesAsText = Yes\n  ReadChargesAsText = Yes\n" if ascii else ""
    assert "Options {" in text
    text = text.replace("Options {", "Options {\n" + extra, 1)
    inp.write_text(text)
    res = hsd.run(str(tmp_path))
    assert (tmp_path / ("charges.dat" if ascii else "charges.bin")).exists()
    inp.write_text(text.replace("SCC = Yes", "SCC = Yes\n  ReadInitialCharges = Yes", 1))
    res2 = hsd.run(str(tmp_path))
    print(f"restart ({'text' if ascii else 'binary'}): {res2['n_iter']} iterations "
          f"(cold {res['n_iter']})")
    assert res2["converged"] and res2["n_iter"] <= 3 < res["n_iter"]
    assert abs(res2["mermin_energy"] - res["mermin_energy"]) / geo.natom < 1e-10
    assert np.max(np.abs(res2["gross_charges"] - res["gross_charges"])) < 1e-9
