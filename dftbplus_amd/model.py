"""Input model of one SCC-DFTB calculation: species/orbital layout, Slater-Koster parameters,
geometry and the options of the SCC cycle.

The orbital layout follows ``TOrbitals`` (src/dftbp/type/orbitals.F90): per species a list of
shells with angular momenta, orbitals ordered shell by shell and within a p shell as
(p_y, p_z, p_x) (src/dftbp/dftb/sk.F90:161-186). Defaults of the options are those of the DFTB+
parser: Broyden mixing 0.2 / 0.01 / 1 / 1e5 / 1e-2 (src/dftbp/dftbplus/parser.F90:5784-5788),
finite-difference step epsilon**0.25 for the SK derivatives (parser.F90:3563-3597),
Ewald tolerance 1e-9 (parser.F90:3364), Fermi temperature floor 1e-8 Ha
(src/dftbp/common/accuracy.F90:64).
"""
from __future__ import annotations

import dataclasses
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import skf as skfmod


@dataclasses.dataclass
class Species:
    name: str
    shells: List[int]             # angular momentum of each shell
    onsite: List[float]           # on-site energy of each shell (Ha)
    hubbu: List[float]            # Hubbard U of each shell (Ha)
    occ: List[float]              # neutral-atom occupation of each shell
    mass: float = 0.0

    @property
    def norb(self) -> int:
        return sum(2 * l + 1 for l in self.shells)

    def orb_shell(self) -> List[int]:
        """Shell index of every orbital (``iShellOrb``)."""
        out = []
        for ish, l in enumerate(self.shells):
            out += [ish] * (2 * l + 1)
        return out


@dataclasses.dataclass
class Parameters:
    """Slater-Koster tables for every ordered species pair plus the repulsives."""
    species: List[Species]
    dist: float
    ham: Dict[Tuple[int, int], np.ndarray]    # (sp1, sp2) -> (nRows, nInteg)
    over: Dict[Tuple[int, int], np.ndarray]
    rep: Dict[Tuple[int, int], skfmod.SplineRep]

    @property
    def nspecies(self) -> int:
        return len(self.species)

    @property
    def nrows(self) -> int:
        return next(iter(self.ham.values())).shape[0]

    def sk_cutoff(self) -> float:
        """nGrid * dist + distFudge (src/dftbp/dftb/slakoeqgrid.F90:164-179)."""
        return self.nrows * self.dist + 1.0

    def rep_cutoff(self) -> float:
        return max(r.cutoff for r in self.rep.values())

    @staticmethod
    def from_skfiles(species: List[Species], files: Dict[Tuple[int, int], skfmod.SKFile]) -> "Parameters":
        ham, over, rep = {}, {}, {}
        dist = None
        for i, si in enumerate(species):
            for j, sj in enumerate(species):
                h, s = skfmod.full_table(si.shells, sj.shells, files[(i, j)], files[(j, i)])
                ham[(i, j)] = h
                over[(i, j)] = s
                rep[(i, j)] = files[(i, j)].rep
                dist = files[(i, j)].dist
        return Parameters(species, dist, ham, over, rep)


@dataclasses.dataclass
class Geometry:
    coords: np.ndarray                      # (nAtom, 3) bohr
    species: np.ndarray                     # (nAtom,) 0-based species index
    latvecs: Optional[np.ndarray] = None    # (3, 3) lattice vectors as rows, bohr

    @property
    def natom(self) -> int:
        return len(self.species)

    @property
    def periodic(self) -> bool:
        return self.latvecs is not None


@dataclasses.dataclass
class Options:
    kT: float = 1.0e-8                      # electron temperature (Ha); floored at minTemp
    scc_tol: float = 1.0e-8
    max_scc_iter: int = 100
    min_scc_iter: int = 1
    mix_param: float = 0.2
    omega0: float = 0.01
    min_weight: float = 1.0
    max_weight: float = 1.0e5
    weight_fac: float = 1.0e-2
    fd_delta: float = float(np.finfo(float).eps) ** 0.25
    ewald_tol: float = 1.0e-9
    # k-points of a periodic system: fractional coordinates (nK x 3) in units of the reciprocal
    # lattice vectors and weights (normalised to 1); None = Gamma point with real arithmetic
    kpoints: Optional[np.ndarray] = None
    kweights: Optional[np.ndarray] = None


def orbital_offsets(params: Parameters, species: Sequence[int]) -> np.ndarray:
    """``iAtomStart`` (0-based, length nAtom + 1) of the dense matrices."""
    norb = np.array([params.species[s].norb for s in species], dtype=np.int64)
    off = np.zeros(len(species) + 1, dtype=np.int64)
    off[1:] = np.cumsum(norb)
    return off


def reference_charges(params: Parameters, species: Sequence[int]) -> np.ndarray:
    """Neutral-atom electron count per atom (the sum of q0 over the atom's orbitals)."""
    return np.array([sum(params.species[s].occ) for s in species], dtype=np.float64)
