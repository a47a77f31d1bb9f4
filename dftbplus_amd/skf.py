"""Slater-Koster file (.skf) reading/writing and the per-species-pair integral tables.

Follows the old (s/p/d) SK file layout read by ``TOldSKData_readFromFile``
(src/dftbp/type/oldskdata.F90:100-200):

* line 1: ``dist nGrid``; the file then holds ``nGrid - 1`` integral rows, row ``i`` at
  ``r = i * dist`` (oldskdata.F90:129-131).
* homonuclear line 2: ``Ed Ep Es SPE Ud Up Us fd fp fs`` (oldskdata.F90:136-137).
* line 3: mass followed by the polynomial repulsive coefficients and cutoff (oldskdata.F90:139-140).
* integral rows: 10 Hamiltonian then 10 overlap values in the order
  dd(sigma,pi,delta) pd(sigma,pi) pp(sigma,pi) sd(sigma) sp(sigma) ss(sigma)
  (``iSKInterOld``, oldskdata.F90:77).
* a ``Spline`` block for the spline repulsive (``TOldSKData_readsplinerep``, oldskdata.F90:210-260).

The per-ordered-pair integral table used by the interpolation follows ``getFullTable``
(src/dftbp/dftbplus/parser.F90:4061-4130): for shells (l1 on species 1, l2 on species 2) the
columns come from the A-B file when l1 <= l2 and from the B-A file otherwise, one column per
m = 0..min(l1, l2) via ``skMap`` (parser.F90:4088-4094).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# Columns of the old SK format (0-based position within the 10 H or 10 S values).
OLD_COLUMNS = ["dds", "ddp", "ddd", "pds", "pdp", "pps", "ppp", "sds", "sps", "sss"]
# (lMax, lMin, m) -> column name, the skMap of parser.F90:4088-4094 restricted to l <= 2.
_SKMAP = {
    (0, 0, 0): "sss",
    (1, 0, 0): "sps",
    (2, 0, 0): "sds",
    (1, 1, 0): "pps", (1, 1, 1): "ppp",
    (2, 1, 0): "pds", (2, 1, 1): "pdp",
    (2, 2, 0): "dds", (2, 2, 1): "ddp", (2, 2, 2): "ddd",
}

AMU_TO_AU = 1822.888486209


@dataclasses.dataclass
class SplineRep:
    """Spline repulsive: exponential head, cubic segments, 5th-order last segment.

    Energy evaluation follows ``TSplineRep_getValue`` (src/dftbp/dftb/repulsive/splinerep.F90:136-160).
    """
    cutoff: float
    expcoeffs: np.ndarray          # a1, a2, a3: exp(-a1 r + a2) + a3
    xstart: np.ndarray             # (nInt,)
    spcoeffs: np.ndarray           # (nInt-1, 4) cubic coefficients
    lastcoeffs: np.ndarray         # (6,) 5th-order tail coefficients

    def value(self, r: float) -> Tuple[float, float]:
        """Energy and radial derivative at distance r."""
        if r < 1.0e-2 or r >= self.cutoff:
            return 0.0, 0.0
        if r < self.xstart[0]:
            a1, a2, a3 = self.expcoeffs
            e = math.exp(-a1 * r + a2)
            return e + a3, -a1 * e
        im = int(np.searchsorted(self.xstart, r, side="right")) - 1
        dr = r - self.xstart[im]
        if im < len(self.xstart) - 1:
            c = self.spcoeffs[im]
        else:
            c = self.lastcoeffs
        en = 0.0
        de = 0.0
        for k in range(len(c) - 1, -1, -1):
            en = en * dr + c[k]
        for k in range(len(c) - 1, 0, -1):
            de = de * dr + k * c[k]
        return en, de


@dataclasses.dataclass
class SKFile:
    """Content of one ordered-pair SK file A-B."""
    dist: float
    ham: np.ndarray                # (nRows, 10) in OLD_COLUMNS order
    over: np.ndarray               # (nRows, 10)
    rep: SplineRep
    # homonuclear only
    onsite: Optional[np.ndarray] = None     # Ed, Ep, Es
    hubbu: Optional[np.ndarray] = None      # Ud, Up, Us
    occ: Optional[np.ndarray] = None        # fd, fp, fs
    mass: float = 0.0

    @property
    def nrows(self) -> int:
        return self.ham.shape[0]


def write_skf(path: str, skf: SKFile) -> None:
    """Write an SK file in the old format (oldskdata.F90:100-200 reads it back)."""
    lines = [f"{skf.dist:.12f} {skf.nrows + 1}"]
    if skf.onsite is not None:
        vals = list(skf.onsite) + [0.0] + list(skf.hubbu) + list(skf.occ)
        lines.append(" ".join(f"{v:.17e}" for v in vals))
        lines.append(f"{skf.mass:.8f} " + " ".join(["0.0"] * 19))
    else:
        lines.append(" ".join(["0.0"] * 20))
    for i in range(skf.nrows):
        lines.append(" ".join(f"{v:.17e}" for v in np.concatenate([skf.ham[i], skf.over[i]])))
    rep = skf.rep
    nint = len(rep.xstart)
    lines.append("Spline")
    lines.append(f"{nint} {rep.cutoff:.15e}")
    lines.append(" ".join(f"{v:.17e}" for v in rep.expcoeffs))
    for j in range(nint - 1):
        x1 = rep.xstart[j + 1]
        lines.append(f"{rep.xstart[j]:.15e} {x1:.15e} " + " ".join(f"{v:.17e}" for v in rep.spcoeffs[j]))
    lines.append(f"{rep.xstart[-1]:.15e} {rep.cutoff:.15e} " + " ".join(f"{v:.17e}" for v in rep.lastcoeffs))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def _floats(line: str) -> List[float]:
    """Numbers of an SK file line; commas separate like blanks and ``n*v`` repeats v n times
    (Fortran list-directed input)."""
    out: List[float] = []
    for tok in line.replace(",", " ").split():
        if "*" in tok:
            n, v = tok.split("*", 1)
            out += [float(v)] * int(n)
        else:
            out.append(float(tok))
    return out


def read_skf(path: str, homo: bool) -> SKFile:
    """Read an old-format SK file (oldskdata.F90:100-200, spline block :210-260)."""
    with open(path) as fh:
        raw = [ln.strip() for ln in fh if ln.strip()]
    if raw[0].startswith("@"):
        raise ValueError("extended (f-orbital) SK format is not supported")
    head = raw[0].replace(",", " ").split()
    dist, ngrid = float(head[0]), int(head[1])
    nrows = ngrid - 1
    pos = 1
    onsite = hubbu = occ = None
    mass = 0.0
    if homo:
        v = _floats(raw[pos])[:10]
        onsite, hubbu, occ = np.array(v[0:3]), np.array(v[4:7]), np.array(v[7:10])
        pos += 1
        mass = _floats(raw[pos])[0]
        pos += 1
    else:
        pos += 1
    data = np.array([_floats(raw[pos + i])[:20] for i in range(nrows)])
    pos += nrows
    while raw[pos] != "Spline":
        pos += 1
    pos += 1
    nint_s, cut_s = raw[pos].split()[:2]
    nint, cutoff = int(nint_s), float(cut_s)
    pos += 1
    expc = np.array([float(x) for x in raw[pos].split()[:3]])
    pos += 1
    xstart = np.zeros(nint)
    spc = np.zeros((nint - 1, 4))
    for j in range(nint - 1):
        vals = [float(x) for x in raw[pos].split()]
        xstart[j] = vals[0]
        spc[j] = vals[2:6]
        pos += 1
    vals = [float(x) for x in raw[pos].split()]
    xstart[-1] = vals[0]
    last = np.array(vals[2:8])
    rep = SplineRep(cutoff, expc, xstart, spc, last)
    return SKFile(dist, data[:, :10].copy(), data[:, 10:].copy(), rep, onsite, hubbu, occ, mass)


def full_table(shells1: Sequence[int], shells2: Sequence[int], skf12: SKFile, skf21: SKFile
               ) -> Tuple[np.ndarray, np.ndarray]:
    """Integral table for ordered species pair (1, 2), following ``getFullTable``
    (parser.F90:4061-4130). Returns (ham, over), each (nRows, nInteg)."""
    hcols, scols = [], []
    for l1 in shells1:
        for l2 in shells2:
            if l1 <= l2:
                src, lmin, lmax = skf12, l1, l2
            else:
                src, lmin, lmax = skf21, l2, l1
            for m in range(lmin + 1):
                c = OLD_COLUMNS.index(_SKMAP[(lmax, lmin, m)])
                hcols.append(src.ham[:, c])
                scols.append(src.over[:, c])
    return np.stack(hcols, axis=1), np.stack(scols, axis=1)


def n_integrals(shells1: Sequence[int], shells2: Sequence[int]) -> int:
    """``getNSKIntegrals`` (parser.F90:4036-4055)."""
    return sum(min(l1, l2) + 1 for l1 in shells1 for l2 in shells2)
