"""Plain-text result files of DFTB+: band.out and detailed.out.

The records follow the reference's writers (src/dftbp/dftbplus/mainio.F90): ``writeBandOut``
(:2459-2507) and, for a spin-unpolarised SCC single point, ``writeDetailedOut1`` (:2794-2926),
``writeDetailedOut2`` (:2930-3229), ``writeDetailedOut3`` (:3330-3572) and ``writeDetailedOut4``
(:3576-3734), with the edit descriptors of ``io/commonformats.F90``. Fortran formatted output
is reproduced by the small formatter below (A, I, F, E edit descriptors, absolute tabs T, and
list-directed records as gfortran writes them).
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np

HARTREE_EV = 27.2113845          # Hartree__eV, common/constants.F90:34
SHELL_NAMES = "spdf"             # shellNames, common/constants.F90:159
SPIN_NAME = ["up", "down"]       # spinName (trimmed), common/constants.F90:163
# orbitalNames(m, l), type/orbitals.F90:59-63 (trimmed)
ORBITAL_NAMES = {0: [""], 1: ["y", "z", "x"], 2: ["xy", "yz", "z2", "xz", "x2-y2"],
                 3: ["y(3x2-y2)", "x2+y2+z2", "yz2", "z3", "xz2", "z(x2-y2)", "x(x2-3y2)"]}


# ------------------------------------------------------------------------------------------
# Fortran edit descriptors
# ------------------------------------------------------------------------------------------
def fmt_i(v: int, w: int) -> str:
    s = str(int(v))
    return s.rjust(w) if len(s) <= w else "*" * w


def fmt_f(x: float, w: int, d: int) -> str:
    """Fw.d (gfortran: a leading zero when it fits, asterisks on overflow)."""
    s = f"{x:.{d}f}"
    if len(s) > w and s.startswith(("0.", "-0.")):
        s = s.replace("0.", ".", 1)
    return s.rjust(w) if len(s) <= w else "*" * w


def fmt_e(x: float, w: int, d: int) -> str:
    """Ew.d: 0.ddddE+xx (mantissa in [0.1, 1)), exponent of two digits up to 99."""
    if x == 0.0 or not math.isfinite(x):
        mant, ex = 0.0, 0
    else:
        ex = math.floor(math.log10(abs(x))) + 1
        mant = x / 10.0 ** ex
        # rounding to d digits can carry the mantissa to 1.0
        if round(abs(mant), d) >= 1.0:
            ex += 1
            mant = x / 10.0 ** ex
    digits = f"{abs(mant):.{d}f}"          # "0.dddd"
    sign = "-" if (x < 0 or (x == 0 and math.copysign(1.0, x) < 0)) else ""
    exps = f"{'+' if ex >= 0 else '-'}{abs(ex):02d}"
    s = f"{sign}{digits}E{exps}" if abs(ex) <= 99 else f"{sign}{digits}{'+' if ex >= 0 else '-'}{abs(ex):03d}"
    return s.rjust(w) if len(s) <= w else "*" * w


class Record:
    """One output record with absolute tabbing (Tn moves to column n, 1-based; later writes
    overwrite, skipped positions are blanks)."""

    def __init__(self):
        self.buf: list = []
        self.pos = 0

    def put(self, s: str) -> "Record":
        end = self.pos + len(s)
        if len(self.buf) < end:
            self.buf.extend(" " * (end - len(self.buf)))
        self.buf[self.pos:end] = list(s)
        self.pos = end
        return self

    def tab(self, col: int) -> "Record":
        self.pos = col - 1
        return self

    def text(self) -> str:
        return "".join(self.buf)


def format2u(label: str, v: float, u1: str, v2: float, u2: str) -> str:
    """(A, ':', T32, F18.10, T51, A, T54, F16.4, T71, A)"""
    return (Record().put(label).put(":").tab(32).put(fmt_f(v, 18, 10)).tab(51).put(u1)
            .tab(54).put(fmt_f(v2, 16, 4)).tab(71).put(u2).text())


def format1ue(label: str, v: float, u: str) -> str:
    """(A, ':', T37, E13.6, T51, A)"""
    return Record().put(label).put(":").tab(37).put(fmt_e(v, 13, 6)).tab(51).put(u).text()


def list_directed(*items) -> str:
    """A list-directed record as gfortran writes it: a leading blank, strings as they are,
    integers in 12 columns, reals as G25.17 without trailing blanks (separated by a blank from
    a preceding number)."""
    out = " "
    prev_num = False
    for it in items:
        if isinstance(it, str):
            out += (" " if prev_num else "") + it
            prev_num = False
        elif isinstance(it, (int, np.integer)):
            out += fmt_i(int(it), 12)
            prev_num = True
        else:
            out += (" " if prev_num else "") + _g25_17(float(it))
            prev_num = True
    return out


def _g25_17(x: float) -> str:
    """gfortran's list-directed REAL(8): F format with 17 significant digits for
    0.1 <= |x| < 1e16 (right-justified in 21 columns plus the trailing blanks dropped here),
    else E format."""
    ax = abs(x)
    if ax == 0.0:
        return "0.0000000000000000".rjust(21)
    if 0.1 <= ax < 1e16:
        ip = math.floor(math.log10(ax)) + 1   # digits before the point (0 below 1)
        dec = 17 - ip
        s = f"{x:.{dec}f}"
        return s.rjust(21)
    return fmt_e(x, 25, 17).lstrip().rjust(21)


# ------------------------------------------------------------------------------------------
# band.out
# ------------------------------------------------------------------------------------------
def write_band_out(path: str, eigen: np.ndarray, filling: np.ndarray, kweights: Sequence[float],
                   append: bool = False) -> None:
    """writeBandOut: eigen and filling as (nLevel, nK, nSpin) (or (nLevel, nK)), per k-point and
    spin a list-directed header and one '(I6, F12.4, F9.5)' record per level (eV, filling)."""
    eigen = np.asarray(eigen, dtype=float)
    filling = np.asarray(filling, dtype=float)
    if eigen.ndim == 2:
        eigen, filling = eigen[:, :, None], filling[:, :, None]
    lines = []
    for isp in range(eigen.shape[2]):
        for ik in range(eigen.shape[1]):
            lines.append(list_directed("KPT ", ik + 1, " SPIN ", isp + 1, " KWEIGHT ",
                                       float(kweights[ik])))
            for il in range(eigen.shape[0]):
                lines.append(fmt_i(il + 1, 6) + fmt_f(HARTREE_EV * eigen[il, ik, isp], 12, 4)
                             + fmt_f(filling[il, ik, isp], 9, 5))
            lines.append("")
    with open(path, "a" if append else "w") as fh:
        fh.write("\n".join(lines) + "\n")


# ------------------------------------------------------------------------------------------
# detailed.out (spin-unpolarised SCC single point)
# ------------------------------------------------------------------------------------------
def _shell_labels(ang: Sequence[int]) -> list:
    """getShellNames (type/orbitals.F90:70-100): s, p, ... then s2, p2, ... for repeats."""
    names = []
    for a in ang:
        base = SHELL_NAMES[a]
        cnt = sum(1 for x in names if x[0] == base)
        names.append(base if cnt == 0 else f"{base}{cnt + 1}")
    return names


def detailed_out(res: dict, geo, params, scc_iter: int, diff_elec: float, scc_error: float,
                 q_input_sum: Optional[float] = None) -> str:
    """The text of detailed.out for the final state of an SCC single point.

    res: ``hsd.run``-style results (energies, orbital populations, convergence, optional
    forces); scc_iter, diff_elec, scc_error: the SCC block of the last iteration; q_input_sum:
    input electrons of the last iteration (defaults to the output ones)."""
    species = np.asarray(geo.species)
    nat = len(species)
    en = dict(res["energies"])
    en["elec"] = en["nonscc"] + en["scc"]
    qorb = np.asarray(res["orbital_populations"], dtype=float)
    norb = np.asarray(res["norb_atom"])
    off = np.concatenate([[0], np.cumsum(norb)])
    q0 = np.array([sum(params.species[s].occ) for s in species])
    qout = np.array([qorb[off[a]:off[a + 1]].sum() for a in range(nat)])
    qin_sum = float(qout.sum()) if q_input_sum is None else float(q_input_sum)
    L = []
    # writeDetailedOut1
    L.append(list_directed("Fermi distribution function"))
    L.append("")
    L.append("Calculation with static geometry")
    L.append("")
    L.append("")
    L.append("*" * 80)
    L.append("iSCC" .rjust(5) + " Total electronic " + "  Diff electronic " + "     SCC error    ")
    L.append(fmt_i(scc_iter, 5) + fmt_e(en["elec"], 18, 8) + fmt_e(diff_elec, 18, 8)
             + fmt_e(scc_error, 18, 8))
    L.append("*" * 80)
    L.append("")
    # writeDetailedOut2 (tPrintMulliken, nSpin = 1)
    L.append(" Total charge: " + fmt_f(float(np.sum(q0 - qout)), 14, 8))
    L.append("")
    L.append(" Atomic gross charges (e)")
    L.append(" Atom" + " " + " Charge".rjust(16))
    for a in range(nat):
        L.append(fmt_i(a + 1, 5) + " " + fmt_f(q0[a] - qout[a], 16, 8))
    L.append("")
    L.append("Nr. of electrons (" + SPIN_NAME[0] + "):" + fmt_f(float(qout.sum()), 16, 8))
    L.append("Atom populations (" + SPIN_NAME[0] + ")")
    L.append(" Atom" + " " + " Population".rjust(16))
    for a in range(nat):
        L.append(fmt_i(a + 1, 5) + " " + fmt_f(qout[a], 16, 8))
    L.append("")
    L.append("l-shell populations (" + SPIN_NAME[0] + ")")
    L.append(" Atom" + " " + "Sh." + " " + "  l" + " " + " Population".rjust(16))
    for a in range(nat):
        sp = params.species[species[a]]
        pos = off[a]
        for ish, ang in enumerate(sp.shells):
            nl = 2 * ang + 1
            L.append(fmt_i(a + 1, 5) + " " + fmt_i(ish + 1, 3) + " " + fmt_i(ang, 3) + " "
                     + fmt_f(float(qorb[pos:pos + nl].sum()), 16, 8))
            pos += nl
    L.append("")
    L.append("Orbital populations (" + SPIN_NAME[0] + ")")
    L.append(" Atom" + " " + "Sh." + " " + "  l" + " " + "  m" + " " + " Population".rjust(16)
             + " " + " Label")
    for a in range(nat):
        sp = params.species[species[a]]
        labels = _shell_labels(sp.shells)
        pos = off[a]
        for ish, ang in enumerate(sp.shells):
            pre = labels[ish] + ("_" if ang > 0 else "")
            for kk in range(2 * ang + 1):
                L.append(fmt_i(a + 1, 5) + " " + fmt_i(ish + 1, 3) + " " + fmt_i(ang, 3) + " "
                         + fmt_i(kk - ang, 3) + " " + fmt_f(float(qorb[pos + kk]), 16, 8) + "  "
                         + pre + ORBITAL_NAMES[ang][kk])
            pos += 2 * ang + 1
    L.append("")
    # writeDetailedOut3 (Fermi filling: chemical potential, band energy, entropy available)
    ef, eband, ts = en["ef"], en["eband"], en["ts"]
    e0 = eband - 0.5 * ts
    L.append(format2u("Fermi level", ef, "H", HARTREE_EV * ef, "eV"))
    L.append(format2u("Band energy", eband, "H", HARTREE_EV * eband, "eV"))
    L.append(format2u("TS", ts, "H", HARTREE_EV * ts, "eV"))
    L.append(format2u("Band free energy (E-TS)", eband - ts, "H", HARTREE_EV * (eband - ts), "eV"))
    L.append(format2u("Extrapolated E(0K)", e0, "H", HARTREE_EV * e0, "eV"))
    L.append("Input / Output electrons (q):" + fmt_f(qin_sum, 18, 10)
             + fmt_f(float(qout.sum()), 18, 10))
    L.append("")
    for label, key in (("Energy H0", "nonscc"), ("Energy SCC", "scc"),
                       ("Total Electronic energy", "elec"), ("Repulsive energy", "repulsive"),
                       ("Total energy", "total")):
        L.append(format2u(label, en[key], "H", en[key] * HARTREE_EV, "eV"))
    ezero = en["total"] - 0.5 * ts
    L.append(format2u("Extrapolated to 0", ezero, "H", ezero * HARTREE_EV, "eV"))
    L.append(format2u("Total Mermin free energy", en["total"] - ts, "H",
                      (en["total"] - ts) * HARTREE_EV, "eV"))
    efr = en["total"] - ts   # the force-related energy of the Fermi distribution (Mermin)
    L.append(format2u("Force related energy", efr, "H", efr * HARTREE_EV, "eV"))
    L.append("")
    # writeDetailedOut4
    L.append("Non-constrained calculation")
    L.append("")
    L.append("SCC converged" if res.get("converged", True)
             else "SCC is NOT converged, maximal SCC iterations exceeded")
    L.append("")
    if "forces" in res:
        f = np.asarray(res["forces"], dtype=float)
        L.append("Total Forces")
        for a in range(nat):
            L.append(fmt_i(a + 1, 5) + "".join(fmt_f(v, 20, 12) for v in f[a]))
        L.append("")
        L.append(format1ue("Maximal derivative component", float(np.max(np.abs(f))), "au"))
        L.append("")
        if geo.latvecs is not None:
            L.append(format1ue("Volume", abs(float(np.linalg.det(geo.latvecs))), "au^3"))
            L.append("")
    return "\n".join(L) + "\n"


# ------------------------------------------------------------------------------------------
# charge files (charges.dat / charges.bin)
# ------------------------------------------------------------------------------------------
RESTART_FORMAT = 8          # restartFormat of sccinit.F90
ELEC_TOL_MAX = 1.0e-7       # elecTolMax of accuracy.F90: the file's checksum tolerance


class ChargeFileError(ValueError):
    pass


def _fortran_logical(tok: str) -> bool:
    t = tok.strip().upper().lstrip(".")
    if t.startswith("T"):
        return True
    if t.startswith("F"):
        return False
    raise ChargeFileError(f"not a logical: {tok!r}")


def _fortran_real(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


class _ListReader:
    """List-directed READs: every READ starts a new record and takes values across records
    when one holds too few; what a record holds beyond a READ's list is skipped."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.pos = 0

    def read(self, n: int):
        vals = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise ChargeFileError("Error during reading external file of charge data")
            vals += self.lines[self.pos].replace(",", " ").split()
            self.pos += 1
        return vals[:n]


class _RecordReader:
    """gfortran unformatted sequential records: a 4-byte length before and after each."""

    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0

    def record(self) -> bytes:
        import struct
        if self.pos + 4 > len(self.data):
            raise ChargeFileError("Error during reading external file of charge data")
        (n,) = struct.unpack_from("<i", self.data, self.pos)
        rec = self.data[self.pos + 4:self.pos + 4 + n]
        self.pos += 8 + n
        return rec


def read_charges(base: str, norb: Sequence[int], ascii: bool, n_spin: int = 1):
    """The orbital populations of a charge file as initQFromFile (sccinit.F90:185-615) reads
    them for a calculation without block charges, density matrices or multipoles: base + '.dat'
    (list-directed text) or base + '.bin' (unformatted), formats 4 to 8. Returns a list of
    n_spin lists with one array of populations per atom (atoms missing from the file: zeros).
    Raises ChargeFileError with the reference's message on a format, atom-count, spin or
    checksum mismatch."""
    import struct
    nat = len(norb)
    if ascii:
        with open(base + ".dat") as fh:
            rd = _ListReader(fh.read())
        fmt = int(rd.read(1)[0])
    else:
        with open(base + ".bin", "rb") as fh:
            rr = _RecordReader(fh.read())
        (fmt,) = struct.unpack_from("<i", rr.record(), 0)
    nlog = {4: 3, 5: 3, 6: 4, 7: 5, 8: 5}.get(fmt)
    if nlog is None:
        raise ChargeFileError("Incompatible file type for external charge data")
    has_natom = fmt >= 5
    if ascii:
        # one READ: the flags, (the atom count,) the spin count and a checksum per spin of the
        # calculation
        nhead = nlog + (2 if has_natom else 1)
        head = rd.read(nhead + n_spin)
        flags = [_fortran_logical(t) for t in head[:nlog]]
        nat_file = int(head[nlog]) if has_natom else nat
        spin = int(head[nhead - 1])
        checksum = [_fortran_real(t) for t in head[nhead:]]
    else:
        rec = rr.record()
        flags = [v != 0 for v in struct.unpack_from(f"<{nlog}i", rec, 0)]
        o = 4 * nlog
        if has_natom:
            (nat_file,) = struct.unpack_from("<i", rec, o)
            o += 4
        else:
            nat_file = nat
        (spin,) = struct.unpack_from("<i", rec, o)
        o += 4
        checksum = list(struct.unpack_from(f"<{n_spin}d", rec, o))
    if any(flags):
        raise ChargeFileError("block charges, density matrices, k-point data or multipoles in "
                              "the charge file are not supported here")
    if nat_file > nat:
        raise ChargeFileError("External charge file has more atoms than are present in the "
                              "system")
    if spin != n_spin:
        raise ChargeFileError("Incorrect number of spins in restart file")
    q = [[np.zeros(norb[a]) for a in range(nat)] for _ in range(spin)]
    for s_ in range(spin):
        for a in range(nat_file):
            if ascii:
                q[s_][a] = np.array([_fortran_real(t) for t in rd.read(norb[a])])
            else:
                q[s_][a] = np.array(struct.unpack_from(f"<{norb[a]}d", rr.record(), 0))
    for s_ in range(spin):
        tot = sum(float(q[s_][a].sum()) for a in range(nat_file))
        if abs(checksum[s_] - tot) > ELEC_TOL_MAX:
            raise ChargeFileError("Error during reading external file of charge data - internal "
                                  "checksum failure, probably a damaged file")
    return q


def write_charges(base: str, populations: Sequence[np.ndarray], ascii: bool) -> str:
    """A charge file in restart format 8 as writeQToFile (sccinit.F90:619-905) writes it for a
    spin-unpolarised calculation without block charges, density matrices or multipoles:
    populations holds one array of orbital populations per atom. Text: list-directed records
    to base + '.dat'; else unformatted records to base + '.bin'. Returns the path."""
    import struct
    nat = len(populations)
    tot = float(sum(np.sum(p) for p in populations))
    if ascii:
        path = base + ".dat"
        lines = [list_directed(RESTART_FORMAT),
                 " " + " ".join(["F"] * 5) + list_directed(nat, 1, tot)[1:]]
        for p in populations:
            lines.append(list_directed(*[float(x) for x in p]))
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")
    else:
        path = base + ".bin"

        def rec(payload: bytes) -> bytes:
            n = struct.pack("<i", len(payload))
            return n + payload + n
        out = rec(struct.pack("<i", RESTART_FORMAT))
        out += rec(struct.pack("<5i", 0, 0, 0, 0, 0) + struct.pack("<2i", nat, 1) +
                   struct.pack("<d", tot))
        for p in populations:
            out += rec(struct.pack(f"<{len(p)}d", *[float(x) for x in p]))
        with open(path, "wb") as fh:
            fh.write(out)
    return path
