"""band.out and detailed.out writers (src/dftbp/dftbplus/mainio.F90:2459-2507, 2794-3734):
Fortran edit descriptors against hand-derived records, band.out records in the layout the
reference's own band.out reader accepts (tools/dptools bandout.py: optional index column,
energy, filling), and the detailed.out blocks of an SCC single point; charge files
(src/dftbp/dftb/sccinit.F90:185-905) against the reference's own charges.dat files."""
import os
import re

import numpy as np
import pytest

from dftbplus_amd import outfiles as O


def test_edit_descriptors():
    assert O.fmt_f(0.5, 16, 8) == "      0.50000000"
    assert O.fmt_f(-1.25, 9, 5) == " -1.25000"
    assert O.fmt_f(123456.0, 5, 2) == "*****"
    assert O.fmt_i(7, 5) == "    7"
    assert O.fmt_e(123.456, 18, 8) == "    0.12345600E+03"
    assert O.fmt_e(-0.000123456, 13, 6) == "-0.123456E-03"
    assert O.fmt_e(0.0, 13, 6) == " 0.000000E+00"
    assert O.fmt_e(0.99999999999, 13, 6) == " 0.100000E+01"   # rounding carries into the exponent
    assert O.fmt_e(1.0, 18, 8) == "    0.10000000E+01"
    # absolute tabs: value in columns 32-49, unit at 51, second value in 54-69, unit at 71
    line = O.format2u("Total energy", -1.5, "H", -40.8170768, "eV")
    assert line[:13] == "Total energy:"
    assert line[13:31] == " " * 18
    assert line[31:49] == "     -1.5000000000"
    assert line[50] == "H"
    assert line[53:69] == "        -40.8171"
    assert line[70:] == "eV"
    assert O.list_directed("KPT ", 1, " SPIN ", 1, " KWEIGHT ", 1.0) == \
        " KPT            1  SPIN            1  KWEIGHT    1.0000000000000000"
    assert O.list_directed(0.125).endswith("0.12500000000000000")


def test_band_out_layout(tmp_path):
    rng = np.random.default_rng(0)
    eig = np.sort(rng.uniform(-1.0, 0.5, (6, 2)), axis=0)
    fill = np.where(eig < 0.0, 2.0, 0.0)
    path = tmp_path / "band.out"
    O.write_band_out(str(path), eig, fill, [0.25, 0.75])
    text = path.read_text()
    blocks = text.strip("\n").split("\n\n")
    assert len(blocks) == 2
    for ik, blk in enumerate(blocks):
        lines = blk.split("\n")
        head = re.match(r"\s*KPT\s*(\d+)\s+SPIN\s*(\d+)\s+KWEIGHT\s*(\d+\.\d+)", lines[0])
        assert head and int(head.group(1)) == ik + 1 and int(head.group(2)) == 1
        assert abs(float(head.group(3)) - [0.25, 0.75][ik]) < 1e-15
        for il, ln in enumerate(lines[1:]):
            assert len(ln) == 6 + 12 + 9                      # (I6, F12.4, F9.5)
            i, e, f = ln[:6], ln[6:18], ln[18:]
            assert int(i) == il + 1
            assert abs(float(e) - O.HARTREE_EV * eig[il, ik]) <= 5e-5
            assert abs(float(f) - fill[il, ik]) <= 5e-6


def test_detailed_out_blocks():
    from dftbplus_amd import synth
    geo, params, opts = synth.workload("cluster27")
    nat = geo.natom
    norb = np.array([params.species[s].norb for s in geo.species])
    rng = np.random.default_rng(1)
    qorb = rng.uniform(0.5, 1.5, int(norb.sum()))
    en = {"mermin": -10.0, "total": -9.99, "nonscc": -12.0, "scc": 0.5, "repulsive": 1.51,
          "ts": 0.01, "eband": -11.0, "ef": -0.2}
    forces = rng.standard_normal((nat, 3)) * 1e-2
    res = {"energies": en, "orbital_populations": qorb, "norb_atom": norb, "converged": True,
           "forces": forces}
    txt = O.detailed_out(res, geo, params, 12, -3.0e-9, 4.0e-9, float(qorb.sum()))
    lines = txt.split("\n")
    i = lines.index(" Atomic gross charges (e)")
    off = np.concatenate([[0], np.cumsum(norb)])
    q0 = [sum(params.species[s].occ) for s in geo.species]
    for a in range(nat):
        idx, val = lines[i + 2 + a].split()
        assert int(idx) == a + 1
        assert abs(float(val) - (q0[a] - qorb[off[a]:off[a + 1]].sum())) < 5e-9
    assert any(ln.startswith("Total Mermin free energy:") and
               abs(float(ln[31:49]) - (en["total"] - en["ts"])) < 1e-10 for ln in lines)
    k = lines.index("Total Forces")
    for a in range(nat):
        parts = lines[k + 1 + a]
        assert len(parts) == 5 + 3 * 20                    # (I5, 3F20.12)
        assert np.allclose([float(parts[5 + 20 * c: 25 + 20 * c]) for c in range(3)],
                           forces[a], atol=1e-12)
    scc = lines[lines.index(" iSCC Total electronic   Diff electronic      SCC error    ") + 1]
    assert int(scc[:5]) == 12 and abs(float(scc[5:23]) - (en["nonscc"] + en["scc"])) < 1e-6
    assert "SCC converged" in lines
    orb = lines.index("Orbital populations (up)")
    first = lines[orb + 2].split()
    assert first[0] == "1" and first[-1] in ("s", "p_y")


def test_charge_files_of_the_reference(tmp_path):
    """initQFromFile on the reference's own charge files (format 4, text): the populations per
    atom and the checksum; a damaged value fails the checksum with the reference's message,
    and a system with fewer atoms than the file is refused."""
    import shutil
    from dftbplus_amd import outfiles
    gold = os.path.join(os.path.dirname(__file__), "golden", "dftbplus", "charges")
    shutil.copy(os.path.join(gold, "2H2O-3rdorder_read.charges.dat"), tmp_path / "charges.dat")
    q = outfiles.read_charges(str(tmp_path / "charges"), [4, 4, 1, 1, 1, 1], True)
    sums = np.array([x.sum() for x in q[0]])
    assert len(q) == 1 and np.allclose(sums[:2], [6.7277403366336541, 6.6870429048902755],
                                       rtol=0, atol=1e-15)
    assert abs(sums.sum() - 16.0) < 1e-7
    assert np.all(q[0][0][1:] == 0.0)
    # a change of 1e-6 e in one value: beyond the checksum tolerance (1e-7)
    text = (tmp_path / "charges.dat").read_text().replace("0.63772931009254918",
                                                           "0.63773031009254918")
    (tmp_path / "charges.dat").write_text(text)
    with pytest.raises(outfiles.ChargeFileError, match="internal checksum failure"):
        outfiles.read_charges(str(tmp_path / "charges"), [4, 4, 1, 1, 1, 1], True)
    shutil.copy(os.path.join(gold, "H3_skipcheck.charges.dat"), tmp_path / "charges.dat")
    q = outfiles.read_charges(str(tmp_path / "charges"), [1, 1, 1], True)
    assert abs(sum(x.sum() for x in q[0]) - 3.0000000000000009) < 1e-12
    # format 4 holds no atom count: read for two atoms, the third population is left out and
    # the checksum fails, as in the reference
    with pytest.raises(outfiles.ChargeFileError, match="internal checksum failure"):
        outfiles.read_charges(str(tmp_path / "charges"), [1, 1], True)
    with pytest.raises(outfiles.ChargeFileError, match="number of spins"):
        outfiles.read_charges(str(tmp_path / "charges"), [1, 1, 1], True, n_spin=2)
    # format 8 holds it
    outfiles.write_charges(str(tmp_path / "c8"), [np.array([1.0])] * 3, True)
    with pytest.raises(outfiles.ChargeFileError, match="more atoms"):
        outfiles.read_charges(str(tmp_path / "c8"), [1, 1], True)


@pytest.mark.parametrize("ascii", [True, False])
def test_charge_file_round_trip(tmp_path, ascii):
    """writeQToFile (format 8, text or unformatted records) read back by initQFromFile: the
    same populations, header flags all false, the checksum consistent."""
    import struct
    from dftbplus_amd import outfiles
    pops = [np.array([6.25, 0.0, 0.0, 0.0]), np.array([0.875]), np.array([0.875])]
    path = outfiles.write_charges(str(tmp_path / "charges"), pops, ascii)
    q = outfiles.read_charges(str(tmp_path / "charges"), [4, 1, 1], ascii)
    for a, b in zip(q[0], pops):
        assert np.array_equal(a, b)
    if ascii:
        lines = open(path).read().splitlines()
        assert lines[0].split() == ["8"] and lines[1].split()[:5] == ["F"] * 5
        assert lines[1].split()[5:7] == ["3", "1"]
    else:
        data = open(path, "rb").read()
        (n,) = struct.unpack_from("<i", data, 0)
        assert n == 4 and struct.unpack_from("<i", data, 4)[0] == 8
