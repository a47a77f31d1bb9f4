"""DFTB+ input ingestion and tagged output against DFTB+'s own test files (CPU)."""
import os

import numpy as np
import pytest

from dftbplus_amd import hsd, tagged

GOLD = os.path.join(os.path.dirname(__file__), "golden", "dftbplus")


@pytest.mark.parametrize("name", ["si32c32_elpa1_autotest.tag", "h2o_geoopt_autotest.tag",
                                  "si2_dual_autotest.tag"])
def test_tagged_files_rewrite_byte_identical(name):
    """Parsing a tag file written by DFTB+ and writing it back reproduces it byte for byte
    (header padding, field widths, 15 digits, 3-digit exponents, 3 values per line)."""
    with open(os.path.join(GOLD, name)) as fh:
        text = fh.read()
    items = tagged.parse_tags(text)
    assert tagged.format_tags(items) == text


def test_tagged_real_field_and_shapes():
    assert tagged.format_real(-2.59522962096508) == " -0.259522962096508E+001"
    assert tagged.format_real(0.0) == "  0.000000000000000E+000"
    assert tagged.format_real(1.0) == "  0.100000000000000E+001"
    assert tagged.format_real(9.999999999999999e-8) == "  0.100000000000000E-006"
    txt = tagged.format_tag("forces", np.arange(6.0).reshape(2, 3).T)
    head, *rows = txt.splitlines()
    assert head == "forces              :real:2:3,2"
    assert len(rows) == 2 and rows[0].split()[1] == "0.100000000000000E+001"
    assert tagged.format_tag("mermin_energy", -1.5).splitlines()[0] == "mermin_energy       :real:0:"
    assert tagged.format_tag("n", np.array([3, 4, 5])).splitlines() == [
        "n                   :integer:1:3", " " * 32 + "3" + " " * 32 + "4", " " * 32 + "5"]


def test_hsd_tree_of_reference_input():
    root = hsd.parse_hsd_file(os.path.join(GOLD, "si32c32_elpa1.hsd"))
    ham = root.child("Hamiltonian").method()
    assert ham.name == "DFTB"
    assert ham.child("scc").text() == "Yes"
    assert float(ham.child("SCCTolerance").text()) == 1.0e-10
    assert ham.child("Eigensolver").method().name == "ELPA"
    assert ham.child("Filling").method().child("Temperature").modifier == "Kelvin"
    lmax = {c.name: c.text() for c in ham.child("MaxAngularMomentum").children}
    assert lmax == {"Si": "p", "C": "p"}
    sk = {c.name: c.data[0] for c in ham.child("SlaterKosterFiles").children}
    assert sk["C-Si"] == "./C-Si.skf" and len(sk) == 4
    assert [float(x) for x in ham.child("KPointsAndWeights").data] == [0, 0, 0, 1]
    assert root.child("Options").child("WriteAutotestTag").text() == "Yes"
    g = hsd.parse_gen(root.child("Geometry").method().data)
    assert g.names == ["Si", "C"] and len(g.species) == 64
    assert np.sum(g.species == 0) == 32
    np.testing.assert_allclose(g.latvecs, np.eye(3) * 8.683348 / hsd.BOHR_AA, rtol=1e-15)
    np.testing.assert_allclose(g.coords[1], np.array([2.170837, 0.0, 2.170837]) / hsd.BOHR_AA,
                               rtol=1e-15)
    # the final coordinates DFTB+ wrote for this input are the input ones (single point)
    with open(os.path.join(GOLD, "si32c32_elpa1_autotest.tag")) as fh:
        tags = dict(tagged.parse_tags(fh.read()))
    np.testing.assert_allclose(tags["end_coords"].T, g.coords, atol=1e-12)


def test_fractional_gen_and_supercell_folding():
    root = hsd.parse_hsd_file(os.path.join(GOLD, "si2_richardson.hsd"))
    g = hsd.parse_gen(root.child("Geometry").method().data)
    a = 2.713546 / hsd.BOHR_AA
    np.testing.assert_allclose(g.coords[1], [a / 2, a / 2, a / 2], rtol=1e-14)
    ham = root.child("Hamiltonian").method()
    k, w = hsd._kpoints(ham, True)
    # 4x4x4 shifted by 1/2 in each direction: 64 points, inversion pairs merged into 32
    assert len(k) == 32 and abs(w.sum() - 1.0) < 1e-15 and np.allclose(w, 1 / 32)
    frac = np.round(k * 8).astype(int)
    assert np.all(frac % 2 == 1)
    for kp in k:
        assert not any(np.allclose((kp + q) - np.round(kp + q), 0) for q in k)
    with pytest.raises(hsd.HsdError, match="Differentiation|SCC"):
        hsd.load_input(os.path.join(GOLD, "si2_richardson.hsd"))


def test_supercell_folding_with_shift_outside_unit_interval():
    """getSuperSampling (periodic.F90:1160-1280) folds the points into [0, 1) before the
    inversion reduction: 2 x 2 x 2 with shifts 1.5 gives the points (i + 1.5) / 2 mod 1 in
    (i1, i2, i3) order with i3 fastest, and each later point at modulo(-k, 1) merges into the
    first. The expected list below is worked out by hand from those steps."""
    text = """
    KPointsAndWeights = SupercellFolding {
      2 0 0
      0 2 0
      0 0 2
      1.5 1.5 1.5
    }
    """
    ham = hsd.parse_hsd(text)
    k, w = hsd._kpoints(ham, True)
    expected = np.array([[0.75, 0.75, 0.75], [0.75, 0.75, 0.25], [0.75, 0.25, 0.75],
                         [0.75, 0.25, 0.25]])
    np.testing.assert_array_equal(k, expected)
    np.testing.assert_array_equal(w, np.full(4, 0.25))
    # an odd folding with a shift below zero keeps the Gamma point and pairs the rest
    text3 = text.replace("2 0 0", "3 0 0").replace("0 2 0", "0 1 0").replace("0 0 2", "0 0 1")
    text3 = text3.replace("1.5 1.5 1.5", "-1.0 0.0 0.0")
    k3, w3 = hsd._kpoints(hsd.parse_hsd(text3), True)
    np.testing.assert_allclose(k3, [[2.0 / 3.0, 0, 0], [0.0, 0, 0]], atol=1e-15)
    np.testing.assert_allclose(w3, [2.0 / 3.0, 1.0 / 3.0], rtol=1e-15)


def test_hsd_syntax_forms(tmp_path):
    (tmp_path / "inc.hsd").write_text("Inner = 3\n")
    (tmp_path / "vals.txt").write_text("1 2\n3 4\n")
    text = """
    # comment
    A = 1.5  # trailing comment
    B [eV] = 2; C = "x y"
    D = Method [m] { E = Yes
      F { 7 8 9 } }
    G = { <<< "vals.txt" }
    H { <<+ "inc.hsd" }
    +I = 4
    """
    root = hsd.parse_hsd(text, str(tmp_path))
    assert root.child("a").text() == "1.5"
    assert root.child("B").modifier == "eV" and root.child("B").text() == "2"
    assert root.child("C").data == ["x y"]
    m = root.child("D").method()
    assert m.name == "Method" and m.modifier == "m" and m.child("E").text() == "Yes"
    assert m.child("F").data == ["7", "8", "9"]
    assert root.child("G").data == ["1", "2", "3", "4"]
    assert root.child("H").child("Inner").text() == "3"
    assert root.child("I").text() == "4"
    with pytest.raises(hsd.HsdError):
        hsd.parse_hsd("A = { 1 2")
