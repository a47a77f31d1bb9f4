"""DFTB+ input ingestion: the HSD input tree, GenFormat geometries, Slater-Koster file sets,
and a single-point driver that writes DFTB+'s tagged outputs.

Covers the input subset of the SCC-DFTB path this backend implements:

* HSD syntax (src/dftbp/io/hsdparser.F90:55-142): ``Name = value``, ``Name { ... }``,
  ``Name = Method { ... }``, ``[modifier]``, quoted strings, ``#`` comments, ``;``,
  ``<<<`` (include a file as text) and ``<<+`` (include a file as HSD). Names compare without
  case; the extension prefixes ``+?*/!`` are accepted and ignored.
* GenFormat geometries (``readTGeometryGen``, src/dftbp/type/typegeometryhsd.F90): types
  ``C`` (cluster), ``S`` (supercell, Cartesian), ``F`` (supercell, fractional); lengths in
  Angstrom.
* ``Hamiltonian = DFTB`` with SCC, SCCTolerance, MaxSCCIterations, Fermi filling,
  MaxAngularMomentum (s, p), SlaterKosterFiles (explicit list or Type2FileNames),
  KPointsAndWeights (explicit list or diagonal SupercellFolding, reduced by inversion),
  Broyden mixing, Charge = 0 (parser.F90:2200-3900). Keywords for physics this backend does not
  implement raise an error instead of being ignored.
* Outputs: autotest.tag and results.tag through ``tagged`` (mainio.F90:1873-2340).
"""
from __future__ import annotations

import dataclasses
import os
import re
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import skf as skfmod
from . import outfiles, tagged
from .model import Geometry, Options, Parameters, Species, reference_charges

# src/dftbp/common/constants.F90:28-40
BOHR_AA = 0.529177249
HARTREE_EV = 27.2113845
BOLTZMANN = 0.00000316681534524639       # Hartree / Kelvin


class HsdError(ValueError):
    pass


# ------------------------------------------------------------------------------------------
# HSD tree
# ------------------------------------------------------------------------------------------
@dataclasses.dataclass
class Node:
    name: str
    modifier: str = ""
    children: List["Node"] = dataclasses.field(default_factory=list)
    data: List[str] = dataclasses.field(default_factory=list)   # raw tokens (values, lists)

    def child(self, name: str) -> Optional["Node"]:
        key = name.lower()
        for c in self.children:
            if c.name.lower() == key:
                return c
        return None

    def method(self) -> Optional["Node"]:
        """The node a ``Name = Method { ... }`` assignment holds (its only child)."""
        return self.children[0] if len(self.children) == 1 and not self.data else None

    def text(self) -> str:
        return " ".join(self.data)


_TOKEN = re.compile(r"""[ \t\r]*(?:
      (?P<nl>\n)
    | (?P<comment>\#[^\n]*)
    | (?P<inc><<[<+!])
    | (?P<str>"[^"]*"|'[^']*')
    | (?P<mod>\[[^\]]*\])
    | (?P<op>[{}=;])
    | (?P<word>[^\s{}=;\[\]"'#]+)
)""", re.VERBOSE)


def _tokens(text: str):
    pos = 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            if text[pos:].strip() == "":
                return
            raise HsdError(f"HSD: cannot read {text[pos:pos + 20]!r}")
        pos = m.end()
        kind = m.lastgroup
        if kind == "comment":
            continue
        val = m.group(kind)
        if kind == "str":
            val = val[1:-1]
        elif kind == "mod":
            val = val[1:-1].strip()
        yield kind, val


class _Parser:
    def __init__(self, text: str, base: str):
        self.toks = list(_tokens(text))
        self.i = 0
        self.base = base

    def peek(self, k: int = 0):
        j = self.i + k
        return self.toks[j] if j < len(self.toks) else ("eof", None)

    def next(self):
        t = self.peek()
        self.i += 1
        return t

    def skip_nl(self):
        while self.peek()[0] in ("nl",) or self.peek() == ("op", ";"):
            self.i += 1

    def _name(self, word: str) -> str:
        return word.lstrip("+?*/!")

    def _is_statement(self) -> bool:
        kind, _ = self.peek()
        if kind != "word":
            return False
        k = 1
        if self.peek(k)[0] == "mod":
            k += 1
        return self.peek(k) in (("op", "="), ("op", "{"))

    def _include(self, node: Node, how: str):
        kind, fname = self.next()
        if kind not in ("str", "word"):
            raise HsdError("HSD: include without a file name")
        path = os.path.join(self.base, fname)
        with open(path) as fh:
            content = fh.read()
        if how == "<<<":
            node.data += content.split()
        elif how == "<<+":
            sub = _Parser(content, os.path.dirname(path))
            node.children += sub.body(stop=None).children
        else:
            raise HsdError("HSD: XML includes are not supported")

    def body(self, stop: Optional[str]) -> Node:
        node = Node("")
        while True:
            self.skip_nl()
            kind, val = self.peek()
            if kind == "eof":
                if stop is not None:
                    raise HsdError("HSD: missing '}'")
                return node
            if (kind, val) == ("op", "}"):
                if stop is None:
                    raise HsdError("HSD: unbalanced '}'")
                self.i += 1
                return node
            if kind == "inc":
                self.i += 1
                self._include(node, val)
                continue
            if self._is_statement():
                node.children.append(self.statement())
                continue
            if kind in ("word", "str"):
                node.data.append(val)
                self.i += 1
                continue
            raise HsdError(f"HSD: unexpected {val!r}")

    def statement(self) -> Node:
        _, word = self.next()
        node = Node(self._name(word))
        if self.peek()[0] == "mod":
            node.modifier = self.next()[1]
        _, op = self.next()
        if op == "{":
            b = self.body(stop="}")
            node.children, node.data = b.children, b.data
            return node
        # assignment: a method block, an anonymous block or values up to the line end
        kind, val = self.peek()
        if (kind, val) == ("op", "{"):
            self.i += 1
            b = self.body(stop="}")
            node.children, node.data = b.children, b.data
            return node
        k = 1
        if self.peek(k)[0] == "mod":
            k += 1
        if kind == "word" and self.peek(k) == ("op", "{"):
            self.i += 1
            meth = Node(self._name(val))
            if self.peek()[0] == "mod":
                meth.modifier = self.next()[1]
            self.i += 1   # '{'
            b = self.body(stop="}")
            meth.children, meth.data = b.children, b.data
            node.children.append(meth)
            return node
        while True:
            kind, val = self.peek()
            if kind in ("nl", "eof") or (kind, val) in (("op", ";"), ("op", "}")):
                break
            if kind == "inc":
                self.i += 1
                self._include(node, val)
                continue
            node.data.append(val)
            self.i += 1
        return node


def parse_hsd(text: str, base: str = ".") -> Node:
    """The HSD tree of ``text`` (a root node without name); includes relative to ``base``."""
    return _Parser(text, base).body(stop=None)


def parse_hsd_file(path: str) -> Node:
    with open(path) as fh:
        return parse_hsd(fh.read(), os.path.dirname(os.path.abspath(path)))


# ------------------------------------------------------------------------------------------
# GenFormat
# ------------------------------------------------------------------------------------------
@dataclasses.dataclass
class GenGeometry:
    names: List[str]           # species names in the order of the gen file
    species: np.ndarray        # (nAtom,) 0-based
    coords: np.ndarray         # (nAtom, 3) bohr
    latvecs: Optional[np.ndarray] = None   # (3, 3) rows, bohr
    origin: Optional[np.ndarray] = None    # bohr


def _is_number(tok: str) -> bool:
    try:
        float(tok.replace("d", "e").replace("D", "E"))
        return True
    except ValueError:
        return False


def _f(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "E"))


def parse_gen(tokens: List[str]) -> GenGeometry:
    """A GenFormat geometry from its whitespace-separated tokens (Angstrom -> bohr)."""
    if len(tokens) < 2:
        raise HsdError("gen: too short")
    nat, kind = int(tokens[0]), tokens[1].upper()
    if kind not in ("C", "S", "F"):
        raise HsdError(f"gen: geometry type {kind!r} is not supported (C, S, F)")
    pos = 2
    names = []
    while pos < len(tokens) and not _is_number(tokens[pos]):
        names.append(tokens[pos])
        pos += 1
    if not names:
        raise HsdError("gen: no species names")
    need = 5 * nat + (12 if kind != "C" else 0)
    if len(tokens) - pos < need:
        raise HsdError("gen: too few values for the atoms and the lattice")
    rows = np.array([[_f(t) for t in tokens[pos + 5 * a: pos + 5 * a + 5]] for a in range(nat)])
    pos += 5 * nat
    species = rows[:, 1].astype(int) - 1
    if np.any(species < 0) or np.any(species >= len(names)):
        raise HsdError("gen: species index out of range")
    xyz = rows[:, 2:5]
    latvecs = origin = None
    if kind != "C":
        vals = np.array([_f(t) for t in tokens[pos: pos + 12]])
        origin = vals[:3] / BOHR_AA
        latvecs = vals[3:].reshape(3, 3) / BOHR_AA
        coords = xyz @ latvecs if kind == "F" else xyz / BOHR_AA
    else:
        coords = xyz / BOHR_AA
    return GenGeometry(names, species, coords, latvecs, origin)


def read_gen(path: str) -> GenGeometry:
    with open(path) as fh:
        return parse_gen(" ".join(ln.split("#", 1)[0] for ln in fh).split())


def _r(v) -> str:
    """Shortest text that reads back to the same double."""
    return repr(float(v))


def format_gen(names: List[str], species, coords, latvecs=None) -> str:
    """GenFormat text (Angstrom) of a geometry in bohr: type C, or S with the lattice."""
    coords = np.asarray(coords, dtype=float) * BOHR_AA
    kind = "C" if latvecs is None else "S"
    lines = [f"{len(species)} {kind}", " ".join(names)]
    for a, (s, x) in enumerate(zip(species, coords)):
        lines.append(f"{a + 1} {int(s) + 1} {_r(x[0])} {_r(x[1])} {_r(x[2])}")
    if latvecs is not None:
        lines.append("0.0 0.0 0.0")
        for v in np.asarray(latvecs, dtype=float) * BOHR_AA:
            lines.append(f"{_r(v[0])} {_r(v[1])} {_r(v[2])}")
    return "\n".join(lines) + "\n"


def export_input(directory: str, geo: Geometry, names: List[str], skfiles, options: Options,
                 forces: bool = True) -> str:
    """A DFTB+ input directory for a geometry, its SK files ({(i, j): SKFile}) and options:
    geo.gen, the SK files as A-B.skf and a dftb_in.hsd for an SCC single point that writes
    autotest.tag and results.tag. Returns the path of dftb_in.hsd."""
    os.makedirs(directory, exist_ok=True)
    for (i, j), f in skfiles.items():
        skfmod.write_skf(os.path.join(directory, f"{names[i]}-{names[j]}.skf"), f)
    with open(os.path.join(directory, "geo.gen"), "w") as fh:
        fh.write(format_gen(names, geo.species, geo.coords, geo.latvecs))
    lmax = {}
    for (i, j), f in skfiles.items():
        if i == j:
            occ = f.occ
            lmax[names[i]] = "p" if (occ is not None and (occ[1] != 0 or f.onsite[1] != 0)) else "s"
    lines = ['Geometry = GenFormat { <<< "geo.gen" }', "", "Driver = {}", "",
             "Hamiltonian = DFTB {", "  SCC = Yes", f"  SCCTolerance = {_r(options.scc_tol)}",
             f"  MaxSCCIterations = {options.max_scc_iter}",
             f"  Mixer = Broyden {{ MixingParameter = {_r(options.mix_param)} }}",
             f"  Filling = Fermi {{ Temperature [au] = {_r(options.kT)} }}",
             "  MaxAngularMomentum {"]
    lines += [f'    {n} = "{lmax[n]}"' for n in names]
    lines += ["  }", "  SlaterKosterFiles = Type2FileNames {", '    Prefix = ""',
              '    Separator = "-"', '    Suffix = ".skf"', "  }"]
    if geo.latvecs is not None:
        k = np.zeros((1, 3)) if options.kpoints is None else np.asarray(options.kpoints)
        w = np.ones(len(k)) if options.kweights is None else np.asarray(options.kweights)
        lines.append("  KPointsAndWeights {")
        lines += [f"    {_r(a[0])} {_r(a[1])} {_r(a[2])} {_r(b)}" for a, b in zip(k, w)]
        lines.append("  }")
    lines += ["}", "", "Analysis {", f"  CalculateForces = {'Yes' if forces else 'No'}", "}",
              "", "Options {", "  WriteAutotestTag = Yes", "  WriteResultsTag = Yes", "}", "",
              "ParserOptions { ParserVersion = 14 }"]
    path = os.path.join(directory, "dftb_in.hsd")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


# ------------------------------------------------------------------------------------------
# input -> model
# ------------------------------------------------------------------------------------------
@dataclasses.dataclass
class Input:
    geometry: GenGeometry
    params: Parameters
    options: Options
    forces: bool = False
    autotest_tag: bool = False
    results_tag: bool = False
    detailed_out: bool = True      # Options { WriteDetailedOut }, default Yes
    band_out: bool = True          # Options { WriteBandOut }, default Yes
    # charge files (parser.F90:3351, 4204-4216): Hamiltonian { ReadInitialCharges }, Options
    # { ReadChargesAsText, WriteCharges (default Yes), WriteChargesAsText, SkipChargeTest }
    read_charges: bool = False
    read_charges_ascii: bool = False
    write_charges: bool = True
    write_charges_ascii: bool = False
    skip_charge_test: bool = False


_UNSUPPORTED = ["SpinPolarisation", "SpinOrbit", "Dispersion", "ThirdOrder", "ThirdOrderFull",
                "HubbardDerivs", "DampXH", "HCorrection", "ElectricField", "Solvation",
                "RangeSeparated", "Hybrid", "OnSiteCorrection", "DftbPlusU", "Differentiation",
                "Excited", "ElectronDynamics", "REKS", "Transport", "ExternalPotentials"]


def _yes(node: Optional[Node], default: bool) -> bool:
    if node is None:
        return default
    t = node.text().strip().lower()
    if t in ("yes", "true", ".true."):
        return True
    if t in ("no", "false", ".false."):
        return False
    raise HsdError(f"{node.name}: expected Yes or No, got {t!r}")


def _real(node: Optional[Node], default: float) -> float:
    return default if node is None else _f(node.text())


def _temperature(node: Optional[Node]) -> float:
    """Temperature in Hartree from a [Kelvin], [eV] or atomic-unit value."""
    if node is None:
        return 0.0
    t = _f(node.text())
    unit = node.modifier.lower()
    if unit in ("kelvin", "k"):
        return t * BOLTZMANN
    if unit == "ev":
        return t / HARTREE_EV
    if unit in ("", "au", "hartree", "h"):
        return t
    raise HsdError(f"Temperature: unit {node.modifier!r} is not supported")


def _kpoints(ham: Node, periodic: bool):
    node = ham.child("KPointsAndWeights")
    if node is None:
        if periodic:
            raise HsdError("KPointsAndWeights is required for a periodic geometry")
        return None, None
    meth = node.method()
    if meth is not None and meth.name.lower() == "supercellfolding":
        v = [_f(t) for t in meth.data]
        if len(v) != 12:
            raise HsdError("SupercellFolding: 3 x 3 integers and 3 shifts expected")
        m = np.array(v[:9]).reshape(3, 3)
        if np.any(np.abs(m - np.diag(np.diag(m))) > 0):
            raise HsdError("SupercellFolding: only diagonal foldings are supported")
        n = np.diag(m).astype(int)
        shift = np.array(v[9:])
        pts = np.array([[(i + shift[0]) / n[0], (j + shift[1]) / n[1], (k + shift[2]) / n[2]]
                        for i in range(n[0]) for j in range(n[1]) for k in range(n[2])])
        # folded into [0, 1) as getSuperSampling does (periodic.F90:1236)
        pts = np.mod(pts, 1.0)
        return _reduce_by_inversion(pts, np.full(len(pts), 1.0 / len(pts)))
    v = [_f(t) for t in node.data]
    if len(v) % 4 or not v:
        raise HsdError("KPointsAndWeights: rows of k1 k2 k3 weight expected")
    a = np.array(v).reshape(-1, 4)
    w = a[:, 3] / a[:, 3].sum()
    return a[:, :3], w


def _reduce_by_inversion(k: np.ndarray, w: np.ndarray):
    """Time-reversal reduction of getSuperSampling (periodic.F90:1239-1269): for each point
    still irreducible, in order, every later point within 1e-4 (per component) of
    modulo(-k, 1) is dropped and its weight added to the earlier one."""
    nk = len(k)
    irreducible = np.ones(nk, dtype=bool)
    w = np.array(w, dtype=float)
    tol = 1e-4
    for i1 in range(nk):
        if not irreducible[i1]:
            continue
        rr = np.mod(-k[i1], 1.0)
        for i2 in range(i1 + 1, nk):
            if irreducible[i2] and np.all(np.abs(k[i2] - rr) < tol):
                irreducible[i2] = False
                w[i1] += w[i2]
    return k[irreducible].copy(), w[irreducible]


def _sk_paths(ham: Node, names: List[str], base: str) -> Dict[Tuple[int, int], str]:
    node = ham.child("SlaterKosterFiles")
    if node is None:
        raise HsdError("Hamiltonian: SlaterKosterFiles missing")
    out = {}
    meth = node.method()
    if meth is not None and meth.name.lower() == "type2filenames":
        prefix = meth.child("Prefix").text() if meth.child("Prefix") else ""
        sep = meth.child("Separator").text() if meth.child("Separator") else ""
        suffix = meth.child("Suffix").text() if meth.child("Suffix") else ""
        lower = _yes(meth.child("LowerCaseTypeName"), False)
        for i, a in enumerate(names):
            for j, b in enumerate(names):
                na, nb = (a.lower(), b.lower()) if lower else (a, b)
                out[(i, j)] = os.path.join(base, prefix + na + sep + nb + suffix)
        return out
    for c in node.children:
        pair = c.name.split("-")
        if len(pair) != 2:
            raise HsdError(f"SlaterKosterFiles: entry {c.name!r} is not A-B")
        try:
            i = [n.lower() for n in names].index(pair[0].lower())
            j = [n.lower() for n in names].index(pair[1].lower())
        except ValueError:
            continue
        if len(c.data) != 1:
            raise HsdError(f"SlaterKosterFiles: {c.name} needs exactly one file")
        out[(i, j)] = os.path.join(base, c.data[0])
    for i in range(len(names)):
        for j in range(len(names)):
            if (i, j) not in out:
                raise HsdError(f"SlaterKosterFiles: no file for {names[i]}-{names[j]}")
    return out


def load_input(path: str) -> Input:
    """Geometry, parameters and options of a dftb_in.hsd (file or directory)."""
    if os.path.isdir(path):
        path = os.path.join(path, "dftb_in.hsd")
    base = os.path.dirname(os.path.abspath(path))
    root = parse_hsd_file(path)
    gnode = root.child("Geometry")
    if gnode is None:
        raise HsdError("Geometry missing")
    gm = gnode.method()
    if gm is None or gm.name.lower() != "genformat":
        raise HsdError("Geometry: only GenFormat is supported")
    geo = parse_gen(gm.data)
    driver = root.child("Driver")
    if driver is not None and (driver.children or driver.data):
        raise HsdError("Driver: only single-point calculations (Driver = {}) are supported")
    hnode = root.child("Hamiltonian")
    if hnode is None or hnode.method() is None or hnode.method().name.lower() != "dftb":
        raise HsdError("Hamiltonian: only DFTB is supported")
    ham = hnode.method()
    for key in _UNSUPPORTED:
        c = ham.child(key)
        if c is not None and (c.children or (c.data and c.text().lower() not in ("no", "{}"))):
            raise HsdError(f"Hamiltonian: {key} is not supported by this backend")
    if not _yes(ham.child("SCC"), False):
        raise HsdError("Hamiltonian: this backend runs SCC-DFTB (SCC = Yes)")
    for key in ("ShellResolvedSCC", "OrbitalResolvedSCC"):
        if _yes(ham.child(key), False):
            raise HsdError(f"Hamiltonian: {key} = Yes is not supported")
    if abs(_real(ham.child("Charge"), 0.0)) > 0.0:
        raise HsdError("Hamiltonian: a net charge is not supported")
    kT = 0.0
    fill = ham.child("Filling")
    if fill is not None:
        fm = fill.method()
        if fm is None or fm.name.lower() != "fermi":
            raise HsdError("Filling: only Fermi filling is supported")
        kT = _temperature(fm.child("Temperature"))
    mix = 0.2
    mnode = ham.child("Mixer")
    if mnode is not None:
        mm = mnode.method()
        if mm is None or mm.name.lower() != "broyden":
            raise HsdError("Mixer: only Broyden mixing is supported")
        mix = _real(mm.child("MixingParameter"), 0.2)
    lmax_node = ham.child("MaxAngularMomentum")
    if lmax_node is None:
        raise HsdError("Hamiltonian: MaxAngularMomentum missing")
    lnames = {"s": 0, "p": 1, "d": 2, "f": 3}
    lmax = {}
    for c in lmax_node.children:
        l = lnames.get(c.text().strip().lower())
        if l is None or l > 1:
            raise HsdError(f"MaxAngularMomentum {c.name}: only s and p shells are supported")
        lmax[c.name.lower()] = l
    paths = _sk_paths(ham, geo.names, base)
    files = {}
    for (i, j), p in paths.items():
        files[(i, j)] = skfmod.read_skf(p, homo=(i == j))
    species = []
    for i, name in enumerate(geo.names):
        if name.lower() not in lmax:
            raise HsdError(f"MaxAngularMomentum: no entry for {name}")
        l = lmax[name.lower()]
        h = files[(i, i)]
        # homonuclear line: (Ed, Ep, Es), (Ud, Up, Us), (fd, fp, fs)
        onsite = [h.onsite[2], h.onsite[1]][: l + 1]
        hub = [h.hubbu[2], h.hubbu[1]][: l + 1]
        occ = [h.occ[2], h.occ[1]][: l + 1]
        species.append(Species(name, list(range(l + 1)), onsite, hub, occ, h.mass))
    params = Parameters.from_skfiles(species, files)
    kpts, kw = _kpoints(ham, geo.latvecs is not None)
    if kpts is not None and len(kpts) == 1 and np.allclose(kpts[0], 0.0):
        kpts = kw = None   # Gamma only: the real path
    opts = Options(kT=kT, scc_tol=_real(ham.child("SCCTolerance"), 1.0e-5),
                   max_scc_iter=int(_real(ham.child("MaxSCCIterations"), 100)), mix_param=mix,
                   kpoints=kpts, kweights=kw)
    analysis = root.child("Analysis")
    options = root.child("Options")
    return Input(geo, params, opts,
                 forces=_yes(analysis.child("CalculateForces") if analysis else None, False),
                 autotest_tag=_yes(options.child("WriteAutotestTag") if options else None, False),
                 results_tag=_yes(options.child("WriteResultsTag") if options else None, False),
                 detailed_out=_yes(options.child("WriteDetailedOut") if options else None, True),
                 band_out=_yes(options.child("WriteBandOut") if options else None, True),
                 read_charges=_yes(ham.child("ReadInitialCharges"), False),
                 read_charges_ascii=_yes(options.child("ReadChargesAsText") if options else None,
                                         False),
                 write_charges=_yes(options.child("WriteCharges") if options else None, True),
                 write_charges_ascii=_yes(options.child("WriteChargesAsText") if options
                                          else None, False),
                 skip_charge_test=_yes(options.child("SkipChargeTest") if options else None,
                                       False))


# ------------------------------------------------------------------------------------------
# single point and tagged outputs
# ------------------------------------------------------------------------------------------
def run(path: str, out_dir: Optional[str] = None, **backend) -> dict:
    """Single-point SCC (and forces when asked) of a DFTB+ input; writes autotest.tag and
    results.tag into ``out_dir`` (default: the input's directory) when the input asks for them.
    ``backend`` is passed on to ``api.DftbPlus`` (precision, force_mode, ...)."""
    from . import api
    inp = load_input(path)
    if out_dir is None:
        out_dir = path if os.path.isdir(path) else os.path.dirname(os.path.abspath(path))
    g = inp.geometry
    geo = Geometry(g.coords, g.species, g.latvecs)
    in_dir = path if os.path.isdir(path) else os.path.dirname(os.path.abspath(path))
    norb = [inp.params.species[s_].norb for s_ in geo.species]
    dp = api.DftbPlus(inp.params, geo, inp.options, **backend)
    try:
        if inp.read_charges:
            # initQFromFile: the orbital populations of the file, summed per atom (the
            # context's charges are atom resolved)
            q = outfiles.read_charges(os.path.join(in_dir, "charges"), norb,
                                      inp.read_charges_ascii)
            dp.set_input_charges(np.array([float(x.sum()) for x in q[0]]),
                                 check_count=not inp.skip_charge_test)
        res = _results(dp, inp, geo)
        if inp.write_charges:
            # writeQToFile: the input charges of the next iteration, each atom's population in
            # its first orbital (as the reference's files of atom-resolved charges hold them)
            qin = dp.get_input_charges()
            pops = []
            for a, n_ in enumerate(norb):
                v = np.zeros(n_)
                v[0] = qin[a]
                pops.append(v)
            outfiles.write_charges(os.path.join(out_dir, "charges"), pops,
                                   inp.write_charges_ascii)
    finally:
        dp.close()
    if inp.autotest_tag:
        tagged.write_tags(os.path.join(out_dir, "autotest.tag"), autotest_items(res, geo))
    if inp.results_tag:
        tagged.write_tags(os.path.join(out_dir, "results.tag"), results_items(res, geo))
    if inp.band_out:
        nk = res["eigenvalues"].shape[1]
        kw = np.ones(nk) if inp.options.kweights is None else np.asarray(inp.options.kweights)
        outfiles.write_band_out(os.path.join(out_dir, "band.out"), res["eigenvalues"],
                                res["filling"], kw / kw.sum())
    if inp.detailed_out:
        st = res["scc_state"]
        with open(os.path.join(out_dir, "detailed.out"), "w") as fh:
            fh.write(outfiles.detailed_out(res, geo, inp.params, st["iteration"],
                                           st["diff_elec"], st["scc_error"], st["q_input_sum"]))
    return res


def _results(dp, inp: Input, geo: Geometry) -> dict:
    e = dp.get_energy()
    en = dict(dp.energies)
    res = {"mermin_energy": e, "energies": en, "converged": dp.converged, "n_iter": dp.n_iter,
           "n_electrons": float(reference_charges(inp.params, geo.species).sum()),
           "scc_state": dp.scc_state(),
           "gross_charges": dp.get_gross_charges(), "orbital_populations":
           dp.get_orbital_populations(), "norb_atom":
           np.array([inp.params.species[s].norb for s in geo.species])}
    nk = 1 if inp.options.kpoints is None else len(inp.options.kpoints)
    ev, fl = zip(*(dp.get_eigenvalues_k(k) for k in range(nk)))
    res["eigenvalues"] = np.stack(ev, axis=1)      # (nOrb, nK)
    res["filling"] = np.stack(fl, axis=1)
    if inp.forces:
        res["forces"] = -dp.get_gradients()
    return res


def _orbital_charges(res: dict) -> np.ndarray:
    """Populations as (mOrb, nAtom), zero-padded (qOutput(:, :, 1))."""
    norb = res["norb_atom"]
    m = int(norb.max())
    out = np.zeros((m, len(norb)))
    pos = 0
    for a, n in enumerate(norb):
        out[:n, a] = res["orbital_populations"][pos: pos + n]
        pos += n
    return out


def autotest_items(res: dict, geo: Geometry) -> list:
    """The tags of writeAutotestTag (mainio.F90:1873-2000) this backend produces."""
    items = []
    if geo.latvecs is not None:
        items.append(("cell_volume", abs(np.linalg.det(geo.latvecs))))
    items.append(("orbital_charges", _orbital_charges(res)))
    if "forces" in res:
        items.append(("forces", res["forces"].T))
    items.append(("mermin_energy", res["mermin_energy"]))
    items.append(("end_coords", np.asarray(geo.coords).T))
    return items


def results_items(res: dict, geo: Geometry) -> list:
    """The tags of writeResultsTag (mainio.F90:2146-2340) this backend produces."""
    en = res["energies"]
    items = [("fermi_level", np.array([en["ef"]])),
             ("number_of_electrons", np.array([res["n_electrons"]]))]
    items.append(("mermin_energy", en["mermin"]))
    items.append(("extrapolated0_energy", en["total"] - 0.5 * en["ts"]))
    items.append(("eigenvalues", res["eigenvalues"][:, :, None]))
    items.append(("filling", res["filling"][:, :, None]))
    if "forces" in res:
        items.append(("forces", res["forces"].T))
    if geo.latvecs is not None:
        items.append(("cell_volume", abs(np.linalg.det(geo.latvecs))))
    items.append(("orbital_charges", _orbital_charges(res)[:, :, None]))
    items.append(("gross_atomic_charges", res["gross_charges"]))
    return items
