"""CPU oracle for the SCC-DFTB hot path: a numpy restatement of DFTB+'s algorithm.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker / timed CPU baseline; the product path
(``dftbplus_amd``) never imports it and fails loudly when its HIP library is missing.

Every function cites the reference file:line it restates (paths relative to the DFTB+ tree).
Pinning (what checks each function against something other than itself) is listed in the
tests: tests/test_oracle_pins.py. Parity status per function:

* sk_interp .............. pinned (degree-7 polynomial exactness, tail closed form/continuity)
* rotate_blocks .......... pinned (textbook SK table, rotation invariance)
* neighbour_list ......... pinned (brute-force pair count, half-list invariants)
* build_h0_s / unpack .... pinned (dense brute force, unpack->pack round trip)
* mulliken ............... pinned (sum rule Tr(PS) = nEl, dense row sums)
* expgamma ............... pinned (on-site and long-range limits, symmetry)
* invr_cluster/_ewald .... pinned (reference unit fixture coulomb.F90:31-103; NaCl Madelung,
                           alpha independence)
* fermi_filling .......... pinned (electron count, two-level closed form)
* BroydenMixer ........... pinned (reference golden sequence broyden.F90:24-58)
* density matrices ....... pinned (Tr(PS) = nEl, Tr(WS) = Eband)
* scc / forces ........... pinned (central finite differences of the total energy; momentum
                           conservation Sum F = 0 for clusters)
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Tuple

import numpy as np
import scipy.linalg
import scipy.special

DIST_FUDGE = 1.0          # accuracy.F90:44
MIN_TEMP = 1.0e-8         # accuracy.F90:64
ELEC_TOL = 1.0e-15        # accuracy.F90:55
ELEC_TOL_MAX = 1.0e-7     # accuracy.F90:60
TOL_SAME_DIST = 1.0e-5    # accuracy.F90:69
MIN_NEIGH_DIST = 1.0e-2   # accuracy.F90:78
MIN_HUB_DIFF = None       # set below from accuracy.F90 (minHubDiff)
MIN_SHORT_GAMMA = 1.0e-10  # accuracy.F90:86
TOL_SHORT_GAMMA = 1.0e-10  # accuracy.F90:94
N_INTER = 8               # slakoeqgrid.F90:73
N_RIGHT = 4               # slakoeqgrid.F90:85
DELTA_R = 1.0e-5          # slakoeqgrid.F90:89
EPS = float(np.finfo(float).eps)
MIN_HUB_DIFF = 1.0e-10    # accuracy.F90 minHubDiff


# --------------------------------------------------------------------------------------------
# Neighbour list  (src/dftbp/dftb/periodic.F90:446-696, getCellTranslations)
# --------------------------------------------------------------------------------------------

def cell_translations(latvecs: Optional[np.ndarray], cutoff: float) -> np.ndarray:
    """Lattice translations that can hold atoms within ``cutoff`` of the central cell; the zero
    translation comes first (periodic.F90:565-580 relies on rCellVec(:,1) = 0)."""
    if latvecs is None:
        return np.zeros((1, 3))
    rec = np.linalg.inv(latvecs).T            # rows: reciprocal vectors / 2pi
    nmax = [int(math.ceil(cutoff * np.linalg.norm(rec[i]))) + 1 for i in range(3)]
    out = [np.zeros(3)]
    rng = [range(-n, n + 1) for n in nmax]
    for a in rng[0]:
        for b in rng[1]:
            for c in rng[2]:
                if a == 0 and b == 0 and c == 0:
                    continue
                out.append(a * latvecs[0] + b * latvecs[1] + c * latvecs[2])
    return np.array(out)


@dataclasses.dataclass
class NeighbourList:
    coords_all: np.ndarray      # (nAll, 3): central atoms first, then images
    img2cent: np.ndarray        # (nAll,)
    neigh: List[np.ndarray]     # per atom: image indices, element 0 = the atom itself
    dist2: List[np.ndarray]     # per atom: squared distances (element 0 = 0)

    def n_within(self, iat: int, cutoff: float) -> int:
        """Number of neighbours (excluding self) within cutoff (``getNrOfNeighbours``)."""
        return int(np.searchsorted(self.dist2[iat][1:], cutoff ** 2, side="right"))


def neighbour_list(coords: np.ndarray, latvecs: Optional[np.ndarray], cutoff: float) -> NeighbourList:
    """Half neighbour list: atom a lists the images of atoms b >= a within ``cutoff``, sorted by
    distance; image atoms are appended in (cell, atom) order (periodic.F90:565-636, 672-690).
    Atoms closer than minNeighDist raise (periodic.F90:617-628)."""
    nat = len(coords)
    cells = cell_translations(latvecs, cutoff)
    cut2 = cutoff * cutoff
    coords_all = [c for c in coords]
    img = list(range(nat))
    lists = [[] for _ in range(nat)]
    d2s = [[] for _ in range(nat)]
    for ic, T in enumerate(cells):
        sh = coords + T
        for a1 in range(nat):
            rr = sh[a1]
            d2 = np.sum((coords[: a1 + 1] - rr) ** 2, axis=1)
            hits = np.nonzero(d2 <= cut2)[0]
            if len(hits) == 0:
                continue
            if ic == 0:
                idx = a1
            else:
                coords_all.append(rr)
                img.append(a1)
                idx = len(coords_all) - 1
            for a2 in hits:
                if d2[a2] < MIN_NEIGH_DIST ** 2:
                    if ic == 0 and a2 == a1:
                        continue
                    raise ValueError(f"Atoms {a1 + 1} and {a2 + 1} too close together")
                lists[a2].append(idx)
                d2s[a2].append(d2[a2])
    neigh, dist2 = [], []
    for a in range(nat):
        arr = np.array(lists[a], dtype=np.int64)
        dd = np.array(d2s[a])
        o = np.argsort(dd, kind="stable")
        neigh.append(np.concatenate([[a], arr[o]]).astype(np.int64))
        dist2.append(np.concatenate([[0.0], dd[o]]))
    return NeighbourList(np.array(coords_all), np.array(img, dtype=np.int64), neigh, dist2)


@dataclasses.dataclass
class Sparse:
    """Flattened (atom, neighbour) block list of the sparse H/S format
    (``getSparseDescriptor``, sparse2dense.F90:3617-3662)."""
    at1: np.ndarray       # central atom of the block column
    img2: np.ndarray      # image index of the neighbour
    at2: np.ndarray       # central atom of the neighbour
    start: np.ndarray     # offset of the block (column-major nOrb2 x nOrb1)
    nnz: int
    nneigh_sk: np.ndarray  # per atom, neighbours within the SK cutoff (self excluded)
    rows: Optional[np.ndarray] = None     # dense row of every sparse element
    cols: Optional[np.ndarray] = None     # dense column of every sparse element
    offsite: Optional[np.ndarray] = None  # element belongs to a block with atom1 != atom2
    prow: Optional[np.ndarray] = None     # row/column read by pack (on-site blocks: lower part)
    pcol: Optional[np.ndarray] = None

    def index_maps(self, off: np.ndarray) -> "Sparse":
        rows, cols, offs, prow, pcol = (np.zeros(self.nnz, dtype=np.int64) for _ in range(5))
        offs = np.zeros(self.nnz, dtype=bool)
        blk = np.zeros(self.nnz, dtype=np.int64)
        swap = np.zeros(self.nnz, dtype=bool)
        for k in range(len(self.at1)):
            a, b = self.at1[k], self.at2[k]
            n1, n2 = off[a + 1] - off[a], off[b + 1] - off[b]
            r = np.tile(np.arange(n2), n1)
            c = np.repeat(np.arange(n1), n2)
            sl = slice(self.start[k], self.start[k] + n1 * n2)
            rows[sl] = off[b] + r
            cols[sl] = off[a] + c
            offs[sl] = a != b
            blk[sl] = k
            if a == b:
                lo = r >= c
                swap[sl] = ~lo
                prow[sl] = np.where(lo, off[b] + r, off[b] + c)
                pcol[sl] = np.where(lo, off[a] + c, off[a] + r)
            else:
                prow[sl] = off[b] + r
                pcol[sl] = off[a] + c
        self.rows, self.cols, self.offsite, self.prow, self.pcol = rows, cols, offs, prow, pcol
        self.blk, self.swap = blk, swap
        return self


def sparse_descriptor(nl: NeighbourList, norb_atom: np.ndarray, sk_cutoff: float) -> Sparse:
    at1, img2, at2, start = [], [], [], []
    ind = 0
    nnsk = np.zeros(len(nl.neigh), dtype=np.int64)
    for a in range(len(nl.neigh)):
        nn = nl.n_within(a, sk_cutoff)
        nnsk[a] = nn
        for k in range(nn + 1):
            j = nl.neigh[a][k]
            jc = nl.img2cent[j]
            at1.append(a)
            img2.append(j)
            at2.append(jc)
            start.append(ind)
            ind += norb_atom[a] * norb_atom[jc]
    return Sparse(np.array(at1), np.array(img2), np.array(at2), np.array(start), ind, nnsk)


# --------------------------------------------------------------------------------------------
# Slater-Koster interpolation  (slakoeqgrid.F90:183-244, interpolation.F90:31-67, 182-239)
# --------------------------------------------------------------------------------------------

def _neville(xa: np.ndarray, ya: np.ndarray, xx: np.ndarray) -> np.ndarray:
    """``polyInterU2`` (interpolation.F90:182-239) vectorised over points.
    xa: (P, n) abscissae, ya: (P, nInt, n), xx: (P,) -> (P, nInt)."""
    n = xa.shape[1]
    delta = 1.0 / (xa[:, 1:] - xa[:, :1])              # (P, n-1)
    cc = ya.copy()
    dd = ya.copy()
    icl = np.ceil((xx - xa[:, 0]) * delta[:, 0]).astype(np.int64)   # 1-based
    P = xa.shape[0]
    ar = np.arange(P)
    yy = ya[ar, :, icl - 1].copy()
    icl = icl - 1
    for mm in range(1, n):
        for ii in range(0, n - mm):
            r2 = (dd[:, :, ii] - cc[:, :, ii + 1]) * delta[:, mm - 1][:, None]
            cc[:, :, ii] = (xa[:, ii] - xx)[:, None] * r2
            dd[:, :, ii] = (xa[:, ii + mm] - xx)[:, None] * r2
        take_c = 2 * icl < n - mm
        dyy = np.where(take_c[:, None], cc[ar, :, np.minimum(icl, n - 1)], dd[ar, :, np.maximum(icl - 1, 0)])
        icl = np.where(take_c, icl, icl - 1)
        yy += dyy
    return yy


def _poly5_to_zero(y0, y0p, y0pp, xx, dx, invdx):
    """``poly5ToZero`` (interpolation.F90:31-67)."""
    dx1 = y0p * dx
    dx2 = y0pp * dx * dx
    dd = 10.0 * y0 - 4.0 * dx1 + 0.5 * dx2
    ee = -15.0 * y0 + 7.0 * dx1 - 1.0 * dx2
    ff = 6.0 * y0 - 3.0 * dx1 + 0.5 * dx2
    xr = xx * invdx
    return ((ff * xr + ee) * xr + dd) * xr * xr * xr


def sk_interp(table: np.ndarray, dist: float, r: np.ndarray) -> np.ndarray:
    """``SlakoEqGrid_interNew_`` (slakoeqgrid.F90:183-244): table (nGrid, nInt) with row i at
    r = i * dist; returns (P, nInt)."""
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    leng, nint = table.shape
    rmax = leng * dist + DIST_FUDGE
    ind = np.floor(r / dist).astype(np.int64)
    out = np.zeros((len(r), nint))
    poly = (r < rmax) & (ind < leng)
    tail = (r < rmax) & (ind >= leng)
    if np.any(poly):
        rr = r[poly]
        ilast = np.maximum(np.minimum(leng, ind[poly] + N_RIGHT), N_INTER)
        rows = ilast[:, None] - N_INTER + np.arange(N_INTER)[None, :]       # 0-based rows
        xa = (rows + 1).astype(np.float64) * dist
        ya = np.transpose(table[rows], (0, 2, 1))                            # (P, nInt, 8)
        out[poly] = _neville(xa, ya, rr)
    if np.any(tail):
        rr = r[tail]
        P = len(rr)
        rows = np.arange(leng - N_INTER, leng)
        xa = np.tile((rows + 1).astype(np.float64) * dist, (P, 1))
        ya = np.tile(table[rows].T[None], (P, 1, 1))
        x8 = xa[:, -1]
        y0 = _neville(xa, ya, x8 - DELTA_R)
        y2 = _neville(xa, ya, x8 + DELTA_R)
        y1 = table[leng - 1][None, :]
        y1p = (y2 - y0) / (2.0 * DELTA_R)
        y1pp = (y2 + y0 - 2.0 * y1) / (DELTA_R * DELTA_R)
        dr = (rr - rmax)[:, None]
        out[tail] = _poly5_to_zero(y1, y1p, y1pp, dr, -DIST_FUDGE, -1.0 / DIST_FUDGE)
    return out


# --------------------------------------------------------------------------------------------
# Rotation of the SK integrals  (sk.F90:35-140 rotateH0, ss/sp/pp :144-289)
# --------------------------------------------------------------------------------------------

def rotate_blocks(sk: np.ndarray, shells1, shells2, ll, mm, nn) -> np.ndarray:
    """Diatomic blocks hh(nOrb2, nOrb1) for P pairs (s and p shells only); sk (P, nInt)."""
    P = sk.shape[0]
    n1 = sum(2 * l + 1 for l in shells1)
    n2 = sum(2 * l + 1 for l in shells2)
    hh = np.zeros((P, n2, n1))
    ind = 0
    icol = 0
    for l1 in shells1:
        o1 = 2 * l1 + 1
        irow = 0
        for l2 in shells2:
            o2 = 2 * l2 + 1
            if l1 > 1 or l2 > 1:
                raise NotImplementedError("only s and p shells")
            if l1 == 0 and l2 == 0:
                blk = sk[:, ind][:, None, None]
            elif l1 + l2 == 1:
                # sp(): hh(1:3,1) = (m, n, l) * sk  (sk.F90:178-184)
                v = np.stack([mm, nn, ll], axis=1) * sk[:, ind][:, None]
                blk = v[:, :, None] if l1 == 0 else -v[:, None, :]     # (-1)^(l1+l2) transpose
            else:
                s, p = sk[:, ind], sk[:, ind + 1]
                # pp() (sk.F90:252-289), rows/cols in (y, z, x) order
                u = np.stack([mm, nn, ll], axis=1)
                blk = (s - p)[:, None, None] * u[:, :, None] * u[:, None, :] + p[:, None, None] * np.eye(3)[None]
            hh[:, irow:irow + o2, icol:icol + o1] = blk
            ind += min(l1, l2) + 1
            irow += o2
        icol += o1
    return hh


def _pair_groups(params, geo_species, sp: Sparse, with_self=False):
    """Indices of off-site blocks grouped by ordered species pair (sp1 = atom 1, sp2 = atom 2)."""
    groups: Dict[Tuple[int, int], np.ndarray] = {}
    s1 = geo_species[sp.at1]
    s2 = geo_species[sp.at2]
    offsite = np.ones(len(sp.at1), dtype=bool) if with_self else (sp.img2 != sp.at1)
    for a in range(params.nspecies):
        for b in range(params.nspecies):
            idx = np.nonzero(offsite & (s1 == a) & (s2 == b))[0]
            if len(idx):
                groups[(a, b)] = idx
    return groups


def diatomic_blocks(params, table, sp1, sp2, vec):
    """SK blocks for bond vectors ``vec`` (P,3) from atom 1 to atom 2 (nonscc.F90:404-421)."""
    dist = np.sqrt(np.sum(vec ** 2, axis=1))
    u = vec / dist[:, None]
    ski = sk_interp(table[(sp1, sp2)], params.dist, dist)
    return rotate_blocks(ski, params.species[sp1].shells, params.species[sp2].shells, u[:, 0], u[:, 1], u[:, 2])


def build_h0_s(params, species, nl: NeighbourList, sp: Sparse):
    """``buildH0``/``buildS`` (nonscc.F90:66-181): on-site energies / unit overlap on the
    diagonal of the self block, SK blocks for every other neighbour."""
    h0 = np.zeros(sp.nnz)
    s = np.zeros(sp.nnz)
    norb = np.array([params.species[x].norb for x in species])
    for k in np.nonzero(sp.img2 == sp.at1)[0]:
        a = sp.at1[k]
        spc = params.species[species[a]]
        n = norb[a]
        eps = np.array([spc.onsite[ish] for ish in spc.orb_shell()])
        for o in range(n):
            h0[sp.start[k] + o * (n + 1)] = eps[o]
            s[sp.start[k] + o * (n + 1)] = 1.0
    for (a, b), idx in _pair_groups(params, species, sp).items():
        vec = nl.coords_all[sp.img2[idx]] - nl.coords_all[sp.at1[idx]]
        hb = diatomic_blocks(params, params.ham, a, b, vec)
        sb = diatomic_blocks(params, params.over, a, b, vec)
        n1, n2 = params.species[a].norb, params.species[b].norb
        cols = sp.start[idx][:, None] + np.arange(n1 * n2)[None, :]
        h0[cols] = hb.transpose(0, 2, 1).reshape(len(idx), -1)   # column-major flatten
        s[cols] = sb.transpose(0, 2, 1).reshape(len(idx), -1)
    return h0, s


# --------------------------------------------------------------------------------------------
# Sparse <-> dense  (sparse2dense.F90:266-321 unpackHS_real, :1024-1097 packHS_real)
# --------------------------------------------------------------------------------------------

def unpack(x: np.ndarray, params, species, sp: Sparse, off: np.ndarray) -> np.ndarray:
    """Dense lower triangle (plus full diagonal blocks), image blocks accumulated."""
    N = int(off[-1])
    out = np.zeros((N, N))
    if sp.rows is not None:
        np.add.at(out, (sp.rows, sp.cols), x)
        return out
    for k in range(len(sp.at1)):
        a, b = sp.at1[k], sp.at2[k]
        n1, n2 = off[a + 1] - off[a], off[b + 1] - off[b]
        blk = x[sp.start[k]: sp.start[k] + n1 * n2].reshape(n1, n2).T
        out[off[b]:off[b] + n2, off[a]:off[a] + n1] += blk
    return out


def block_phases(nl, sp: Sparse, coords, latvecs, kpt) -> np.ndarray:
    """Bloch phase exp(i 2 pi k.T) of every sparse block, T the fractional cell translation of
    the block's image atom (unpackHS_cmplx_kpts, sparse2dense.F90:185-262)."""
    T = nl.coords_all[sp.img2] - coords[nl.img2cent[sp.img2]]
    frac = np.rint(T @ np.linalg.inv(latvecs))
    return np.exp(2j * np.pi * (frac @ np.asarray(kpt, dtype=float)))


def unpack_k(x: np.ndarray, sp: Sparse, off: np.ndarray, phase: np.ndarray) -> np.ndarray:
    """Complex square matrix at one k-point: blocks times their phase added at (atom2 rows,
    atom1 columns), i.e. the lower triangle plus full on-site blocks (unpackHS_cmplx_kpts)."""
    N = int(off[-1])
    out = np.zeros((N, N), dtype=complex)
    np.add.at(out, (sp.rows, sp.cols), x * phase[sp.blk])
    return out


def hermitian_from_lower(m: np.ndarray) -> np.ndarray:
    """The Hermitian matrix whose lower triangle (real diagonal) is that of m ('L' storage)."""
    lo = np.tril(m, -1)
    return lo + lo.conj().T + np.diag(np.diag(m).real)


def pack_k(dense: np.ndarray, sp: Sparse, phase: np.ndarray, weight: float) -> np.ndarray:
    """kWeight * Re(exp(-i 2 pi k.T) block), on-site blocks completed from their lower triangle
    by conjugation (packHS_cmplx_kpts, sparse2dense.F90:820-918)."""
    vals = dense[sp.prow, sp.pcol]
    vals = np.where(sp.swap, np.conj(vals), vals)
    return weight * np.real(np.conj(phase[sp.blk]) * vals)


def symmetrize_lower(m: np.ndarray) -> np.ndarray:
    return np.tril(m) + np.tril(m, -1).T


def pack(dense: np.ndarray, params, species, sp: Sparse, off: np.ndarray) -> np.ndarray:
    """Gather the half-list blocks from the dense lower triangle, on-site blocks symmetrised."""
    if sp.prow is not None:
        return dense[sp.prow, sp.pcol].copy()
    out = np.zeros(sp.nnz)
    for k in range(len(sp.at1)):
        a, b = sp.at1[k], sp.at2[k]
        n1, n2 = off[a + 1] - off[a], off[b + 1] - off[b]
        blk = dense[off[b]:off[b] + n2, off[a]:off[a] + n1].copy()
        if a == b:
            blk = np.tril(blk) + np.tril(blk, -1).T
        out[sp.start[k]: sp.start[k] + n1 * n2] += blk.T.reshape(-1)
    return out


def mulliken(over: np.ndarray, rho: np.ndarray, species, sp: Sparse, off: np.ndarray) -> np.ndarray:
    """``mullikenPerOrbital`` (populations.F90:120-193): q per orbital (length N)."""
    N = int(off[-1])
    q = np.zeros(N)
    if sp.rows is not None:
        prod = over * rho
        np.add.at(q, sp.cols, prod)
        np.add.at(q, sp.rows[sp.offsite], prod[sp.offsite])
        return q
    for k in range(len(sp.at1)):
        a, b = sp.at1[k], sp.at2[k]
        n1, n2 = off[a + 1] - off[a], off[b + 1] - off[b]
        prod = (over[sp.start[k]: sp.start[k] + n1 * n2] * rho[sp.start[k]: sp.start[k] + n1 * n2]).reshape(n1, n2).T
        q[off[a]:off[a] + n1] += prod.sum(axis=0)
        if a != b:
            q[off[b]:off[b] + n2] += prod.sum(axis=1)
    return q


def add_shift(h0, over, vorb, species, sp: Sparse, off):
    """H = H0 + 1/2 S (V_mu + V_nu) for a diagonal per-orbital potential
    (``addShift_block``, shift.F90:132-229)."""
    if sp.rows is not None:
        return h0 + 0.5 * over * (vorb[sp.rows] + vorb[sp.cols])
    h = h0.copy()
    for k in range(len(sp.at1)):
        a, b = sp.at1[k], sp.at2[k]
        n1, n2 = off[a + 1] - off[a], off[b + 1] - off[b]
        v1 = vorb[off[a]:off[a] + n1]
        v2 = vorb[off[b]:off[b] + n2]
        blk = over[sp.start[k]: sp.start[k] + n1 * n2].reshape(n1, n2).T
        h[sp.start[k]: sp.start[k] + n1 * n2] += (0.5 * blk * (v2[:, None] + v1[None, :])).T.reshape(-1)
    return h


# --------------------------------------------------------------------------------------------
# SCC electrostatics  (shortgammafuncs.F90:27-335, shortgamma.F90:753-936, coulomb.F90)
# --------------------------------------------------------------------------------------------

def _gsub(r, t1, t2):
    return math.exp(-t1 * r) * ((0.5 * t2 ** 4 * t1 / (t1 ** 2 - t2 ** 2) ** 2)
                                - (t2 ** 6 - 3.0 * t2 ** 4 * t1 ** 2) / (r * (t1 ** 2 - t2 ** 2) ** 3))


def _gsubp(r, t1, t2):
    return (-t1 * math.exp(-t1 * r) * ((0.5 * t2 ** 4 * t1 / (t1 ** 2 - t2 ** 2) ** 2)
                                        - (t2 ** 6 - 3.0 * t2 ** 4 * t1 ** 2) / (r * (t1 ** 2 - t2 ** 2) ** 3))
            + math.exp(-t1 * r) * (t2 ** 6 - 3.0 * t2 ** 4 * t1 ** 2) / (r ** 2 * (t1 ** 2 - t2 ** 2) ** 3))


def expgamma(r: float, ua: float, ub: float) -> float:
    """Short-range part of gamma, exponential form (``expGamma``, shortgammafuncs.F90:97-148)."""
    ta, tb = 3.2 * ua, 3.2 * ub
    if r < TOL_SAME_DIST:
        if abs(ua - ub) < MIN_HUB_DIFF:
            return -0.5 * (ua + ub)
        return -0.5 * ((ta * tb) / (ta + tb) + (ta * tb) ** 2 / (ta + tb) ** 3)
    if abs(ua - ub) < MIN_HUB_DIFF:
        tm = 0.5 * (ta + tb)
        return math.exp(-tm * r) * (1.0 / r + 0.6875 * tm + 0.1875 * r * tm ** 2
                                    + 0.02083333333333333333 * r ** 2 * tm ** 3)
    return _gsub(r, ta, tb) + _gsub(r, tb, ta)


def expgamma_prime(r: float, ua: float, ub: float) -> float:
    """``expGammaPrime`` (shortgammafuncs.F90:152-196)."""
    if r < TOL_SAME_DIST:
        return 0.0
    if abs(ua - ub) < MIN_HUB_DIFF:
        tm = 3.2 * 0.5 * (ua + ub)
        return (-tm * math.exp(-tm * r) * (1.0 / r + 0.6875 * tm + 0.1875 * r * tm ** 2
                                          + 0.02083333333333333333 * r ** 2 * tm ** 3)
                + math.exp(-tm * r) * (-1.0 / r ** 2 + 0.1875 * tm ** 2
                                       + 2.0 * 0.02083333333333333333 * r * tm ** 3))
    ta, tb = 3.2 * ua, 3.2 * ub
    return _gsubp(r, ta, tb) + _gsubp(r, tb, ta)


def expgamma_cutoff(u2: float, u1: float, minval: float = MIN_SHORT_GAMMA) -> float:
    """``expGammaCutoff`` (shortgammafuncs.F90:27-92)."""
    rab = 1.0
    while expgamma(rab, u2, u1) > minval:
        rab *= 2.0
    mincut = rab + 0.1
    maxcut = 0.5 * rab - 0.1
    maxg = expgamma(maxcut, u2, u1)
    ming = expgamma(mincut, u2, u1)
    lowerg = expgamma(mincut, u2, u1)
    cut = maxcut + 0.1
    g = expgamma(cut, u2, u1)
    while (g - lowerg) > TOL_SHORT_GAMMA:
        maxcut = 0.5 * (cut + mincut)
        if (maxg >= ming) == (minval >= expgamma(maxcut, u2, u1)):
            mincut = maxcut
            lowerg = expgamma(mincut, u2, u1)
        else:
            cut = maxcut
            g = expgamma(cut, u2, u1)
    return mincut


def invr_cluster(coords: np.ndarray) -> np.ndarray:
    """1/R matrix with zero diagonal (``invRClusterSerial``, coulomb.F90:704-733), full."""
    d = coords[:, None, :] - coords[None, :, :]
    r = np.sqrt(np.sum(d * d, axis=2))
    np.fill_diagonal(r, np.inf)
    return 1.0 / r


def _gterm(g, alpha, vol):
    return 4.0 * math.pi * (math.exp(-0.25 * g * g / (alpha * alpha)) / (vol * g * g))


def _rterm(r, alpha):
    return math.erfc(alpha * r) / r


def _bisect_max(fn, minval):
    """Common search of ``getMaxGEwald``/``getMaxREwald`` (coulomb.F90:1911-2037)."""
    xx = 1.0e-8
    yy = fn(xx)
    while yy > minval:
        xx *= 2.0
        yy = fn(xx)
    xl, yl, xr, yr = 0.5 * xx, fn(0.5 * xx), xx, yy
    it = 1
    while yl - yr > minval and it <= 30:
        xx = 0.5 * (xl + xr)
        yy = fn(xx)
        if yy >= minval:
            xl, yl = xx, yy
        else:
            xr, yr = xx, yy
        it += 1
    return xx


def ewald_parameters(latvecs: np.ndarray, tol: float):
    """alpha, real-space cutoff and G cutoff (``getOptimalAlphaEwald`` coulomb.F90:1824-1906,
    ``getMaxGEwald``/``getMaxREwald``)."""
    vol = abs(np.linalg.det(latvecs))
    rec = 2.0 * math.pi * np.linalg.inv(latvecs).T
    ming = math.sqrt(np.min(np.sum(rec ** 2, axis=1)))
    minr = math.sqrt(np.min(np.sum(latvecs ** 2, axis=1)))

    def diff(al):
        return ((_gterm(4.0 * ming, al, vol) - _gterm(5.0 * ming, al, vol))
                - (_rterm(2.0 * minr, al) - _rterm(3.0 * minr, al)))

    alpha = 1.0e-8
    d = diff(alpha)
    while d < -tol:
        alpha *= 2.0
        d = diff(alpha)
    aleft = 0.5 * alpha
    while d < tol:
        alpha *= 2.0
        d = diff(alpha)
    aright = alpha
    alpha = (aleft + aright) / 2.0
    d = diff(alpha)
    it = 0
    while abs(d) > tol and it <= 30:
        if d < 0:
            aleft = alpha
        else:
            aright = alpha
        alpha = (aleft + aright) / 2.0
        d = diff(alpha)
        it += 1
    maxr = _bisect_max(lambda x: _rterm(x, alpha), tol)
    maxg = _bisect_max(lambda x: _gterm(x, alpha, vol), tol)
    return alpha, maxr, maxg, vol, rec


def gvectors(rec: np.ndarray, maxg: float) -> np.ndarray:
    """Reciprocal lattice points with |G| <= maxg, origin removed, one of each +-G pair
    (``getLatticePoints(..., reduceByInversion, withoutOrigin)``)."""
    nmax = [int(math.ceil(maxg / np.linalg.norm(rec[i]) * 1.0)) + 1 for i in range(3)]
    pts = []
    for a in range(-nmax[0], nmax[0] + 1):
        for b in range(-nmax[1], nmax[1] + 1):
            for c in range(-nmax[2], nmax[2] + 1):
                if (a, b, c) <= (0, 0, 0):
                    continue  # keep the half with (a,b,c) lexicographically positive
                g = a * rec[0] + b * rec[1] + c * rec[2]
                if np.dot(g, g) <= maxg * maxg:
                    pts.append(g)
    return np.array(pts).reshape(-1, 3)


def invr_ewald(coords: np.ndarray, latvecs: np.ndarray, tol: float = 1.0e-9) -> np.ndarray:
    """Ewald 1/R matrix (``invRPeriodicSerial``, coulomb.F90:968-1023): real-space erfc sum over
    images within maxR, reciprocal sum over half the G sphere (x2), -pi/(V alpha^2), and
    -2 alpha/sqrt(pi) on the diagonal. Returned full (symmetric)."""
    alpha, maxr, maxg, vol, rec = ewald_parameters(latvecs, tol)
    nat = len(coords)
    cells = cell_translations(latvecs, maxr)
    m = np.zeros((nat, nat))
    for T in cells:
        d = coords[:, None, :] - (coords[None, :, :] + T)
        r = np.sqrt(np.sum(d * d, axis=2))
        mask = (r <= maxr) & (r > 1.0e-8)
        m[mask] += scipy.special.erfc(alpha * r[mask]) / r[mask]
    G = gvectors(rec, maxg)
    g2 = np.sum(G * G, axis=1)
    w = 2.0 * 4.0 * math.pi / vol * np.exp(-g2 / (4.0 * alpha * alpha)) / g2
    ph = coords @ G.T
    c, s = np.cos(ph), np.sin(ph)
    m += (c * w) @ c.T + (s * w) @ s.T
    m -= math.pi / (vol * alpha * alpha)
    m[np.diag_indices(nat)] -= 2.0 * alpha / math.sqrt(math.pi)
    return m


def invr_ewald_deriv(coords, latvecs, dq, tol=1.0e-9):
    """Gradient of 1/2 dq^T invR dq for the Ewald sum (``addInvRPrimePeriodic``,
    coulomb.F90:1380-1494): d/dR_i."""
    alpha, maxr, maxg, vol, rec = ewald_parameters(latvecs, tol)
    nat = len(coords)
    grad = np.zeros((nat, 3))
    for T in cell_translations(latvecs, maxr):
        d = coords[:, None, :] - (coords[None, :, :] + T)
        r = np.sqrt(np.sum(d * d, axis=2))
        mask = (r <= maxr) & (r > 1.0e-8)
        rr = np.where(mask, r, 1.0)
        # d/dr [erfc(a r)/r]
        dfdr = -(2.0 * alpha / math.sqrt(math.pi) * np.exp(-(alpha * rr) ** 2) / rr + scipy.special.erfc(alpha * rr) / rr ** 2)
        coef = np.where(mask, dfdr / rr, 0.0) * dq[:, None] * dq[None, :]
        grad += np.sum(coef[:, :, None] * d, axis=1)
    G = gvectors(rec, maxg)
    g2 = np.sum(G * G, axis=1)
    w = 2.0 * 4.0 * math.pi / vol * np.exp(-g2 / (4.0 * alpha * alpha)) / g2
    ph = coords @ G.T
    c, s = np.cos(ph), np.sin(ph)
    # d/dr_i sum_j dq_i dq_j w cos(G.(ri-rj)) = -dq_i sum_G w G [sin(G ri) C_G - cos(G ri) S_G]
    C = c.T @ dq
    S = s.T @ dq
    fac = (s * C[None, :] - c * S[None, :]) * w[None, :]
    grad += -dq[:, None] * (fac @ G)
    return grad


@dataclasses.dataclass
class ShortGamma:
    """Per (atom, neighbour) short-range gamma for one unique U per species (non-shell-resolved
    SCC; first-shell U, parser.F90:3792-3794)."""
    values: List[np.ndarray]      # per atom: value per neighbour within its U-pair cutoff
    nneigh: np.ndarray


def short_gamma(params, species, nl: NeighbourList) -> ShortGamma:
    """``updateShortGammaValues_`` (shortgamma.F90:804-858) with cutoffs from ``expGammaCutoff``."""
    u = [sp.hubbu[0] for sp in params.species]
    cut = {}
    for a in range(params.nspecies):
        for b in range(params.nspecies):
            cut[(a, b)] = expgamma_cutoff(u[b], u[a])
    vals = []
    nn = np.zeros(len(species), dtype=np.int64)
    for i in range(len(species)):
        si = species[i]
        js = nl.neigh[i]
        v = []
        for k, j in enumerate(js):
            sj = species[nl.img2cent[j]]
            r = math.sqrt(nl.dist2[i][k])
            if r > cut[(sj, si)] and k > 0:
                v.append(0.0)
            else:
                v.append(expgamma(r, u[sj], u[si]))
        vals.append(np.array(v))
    return ShortGamma(vals, nn)


def gamma_matrix(params, species, nl, coords, latvecs, tol):
    """Full atomic gamma = 1/R (or Ewald) - short-range part, on-site U (the matrix whose product
    with dq gives the atom shift of ``updateShifts``, scc.F90:453-496)."""
    nat = len(species)
    g = invr_cluster(coords) if latvecs is None else invr_ewald(coords, latvecs, tol)
    sg = short_gamma(params, species, nl)
    for i in range(nat):
        for k, j in enumerate(nl.neigh[i]):
            jc = nl.img2cent[j]
            val = sg.values[i][k]
            if val == 0.0:
                continue
            g[jc, i] -= val
            if jc != i:
                g[i, jc] -= val
    return g


# --------------------------------------------------------------------------------------------
# Fermi filling  (etemp.F90:59-194, 294-455, middleGap :481)
# --------------------------------------------------------------------------------------------

def _kahan_sum(vals):
    s, c = 0.0, 0.0
    for v in vals:
        y = v - c
        t = s + y
        c = (t - s) - y
        s = t
    return s


def _kahan_dftb(vals):
    """Kahan accumulation as in DFTB+ (``kahan`` adds the compensation at the end)."""
    tot = 0.0
    comp = 0.0
    for v in vals:
        y = v - comp
        t = tot + y
        comp = (t - tot) - y
        tot = t
    return tot


def _electron_count(ef, eig, kT):
    x = (eig - ef) / kT
    with np.errstate(over="ignore"):
        occ = 1.0 / (1.0 + np.exp(x))
    return math.fsum(occ)


def _electron_fill(ef, eig, kT):
    x = (eig - ef) / kT
    with np.errstate(over="ignore"):
        f = 1.0 / (1.0 + np.exp(x))
    f = np.where(x > 700.0, 0.0, f)
    # levels stop at the first filling <= elecTol (etemp.F90 lpLevels exit)
    stop = np.nonzero(f <= ELEC_TOL)[0]
    if len(stop):
        f[stop[0] + 1:] = 0.0
        nlev = stop[0]
    else:
        nlev = len(f)
    ff = f[:nlev]
    mask = (ff > EPS) & (ff < 1.0 - EPS)
    ts = -math.fsum(np.where(mask, ff * np.log(np.where(mask, ff, 1.0)) + (1 - ff) * np.log(np.where(mask, 1 - ff, 1.0)), 0.0))
    eband = math.fsum(ff * eig[:nlev])
    ts *= kT
    return f, eband, ts


def fermi_filling_k(eig: np.ndarray, nel: float, kT: float, kw: np.ndarray):
    """Spin-unpolarised Fermi filling over k-points (``Efilling`` with kWeight, etemp.F90:59-194,
    ``electronCount``, ``electronFill``, ``middleGap`` :481-520): eig (nK, nLev), normalised
    weights kw. Returns (f (nK, nLev) with the factor 2, Ef, Eband, TS, E0), Eband and TS
    k-weighted."""
    kT = max(kT, MIN_TEMP)
    ne = 0.5 * nel
    eig = np.asarray(eig)
    nk, nlev = eig.shape

    def count(ef):
        x = (eig - ef) / kT
        with np.errstate(over="ignore"):
            occ = 1.0 / (1.0 + np.exp(x))
        return math.fsum((occ * kw[:, None]).ravel())

    def fill(ef):
        f = np.zeros_like(eig)
        eb, ts = [], []
        for k in range(nk):
            fk, ebk, tsk = _electron_fill(ef, eig[k], kT)
            f[k] = fk
            eb.append(kw[k] * ebk)
            ts.append(kw[k] * tsk)
        return f, math.fsum(eb), math.fsum(ts)

    def done(ef, scale=1.0):
        f, eb, ts = fill(ef)
        return 2 * f * scale, ef, 2 * eb, 2 * ts, 2 * (eb - 0.5 * ts)

    if ne < EPS:
        # no electrons (etemp.F90:113-122): empty levels, E_F well below the lowest level
        return (np.zeros_like(eig), eig.min() - 1000.0 * (kT + float(np.finfo(np.float32).eps)),
                0.0, 0.0, 0.0)
    if nlev <= ne:
        return done(eig.max() + 1000.0 * (kT + float(np.finfo(np.float32).eps)))
    if abs(ne - round(ne)) <= ELEC_TOL:
        flat = eig.T.ravel(order="F")           # (level, k) column-major, as reshape(eigenvals)
        order = np.argsort(flat, kind="stable")
        acc, ind = 0.0, 0
        while acc < ne:
            acc += kw[order[ind] // nlev]
            ind += 1
        ind = min(len(flat) - 1, ind)
        ef = 0.5 * (flat[order[ind]] + flat[order[ind - 1]])
        if abs(ne - count(ef)) <= ELEC_TOL_MAX:
            return done(ef)
    upper = eig.max() + 0.01
    lower = eig.min() - 0.01
    nmax, nmin = count(upper), count(lower)
    while nmin > ne:
        lower = 2.0 * (lower - upper) + lower
        nmin = count(lower)
    while nmax < ne:
        upper = 2.0 * (upper - lower) + lower
        nmax = count(upper)
    ef = 0.5 * (upper + lower)
    ncur = count(ef)
    eps2 = 2.0 * EPS
    while abs(ne - ncur) > ELEC_TOL and abs(upper - lower) >= max(abs(ef) * eps2, eps2):
        if (nmax >= nmin) == (ne >= ncur):
            lower, nmin = ef, ncur
        else:
            upper, nmax = ef, ncur
        ef = 0.5 * (upper + lower)
        ncur = count(ef)
    if abs(ne - ncur) > ELEC_TOL_MAX:
        raise RuntimeError("Fermi level search did not converge.")
    scale = ne / ncur if (ne > eps2 and abs(ne - ncur) > eps2) else 1.0
    return done(ef, scale)


def levels(f: np.ndarray) -> slice:
    """Levels entering the density matrices: up to the last filling >= machine epsilon
    (densitymatrix.F90:549-554, 688-692)."""
    big = np.nonzero(np.abs(f) >= np.finfo(float).eps)[0]
    return slice(0, int(big[-1]) + 1 if big.size else len(f))


def fermi_filling(eig: np.ndarray, nel: float, kT: float):
    """Spin-unpolarised Fermi filling of one k-point (``getFillingsAndBandEnergies``,
    main.F90:4544-4651 -> ``Efilling``): returns (f, Ef, Eband, TS, E0) with the factor 2."""
    kT = max(kT, MIN_TEMP)
    ne = 0.5 * nel
    eig = np.asarray(eig)
    n = len(eig)
    if ne < EPS:
        # no electrons (etemp.F90:113-122): empty levels, E_F well below the lowest level
        return (np.zeros_like(eig), eig.min() - 1000.0 * (kT + float(np.finfo(np.float32).eps)),
                0.0, 0.0, 0.0)
    if n <= ne:
        ef = eig.max() + 1000.0 * (kT + float(np.finfo(np.float32).eps))
        f, eb, ts = _electron_fill(ef, eig, kT)
        return 2 * f, ef, 2 * eb, 2 * ts, 2 * (eb - 0.5 * ts)
    if abs(ne - round(ne)) <= ELEC_TOL:
        # middle gap (etemp.F90:481-520): first level index where the count reaches ne
        order = np.argsort(eig, kind="stable")
        k = int(math.ceil(ne - 1e-12))
        k = min(k, n - 1)
        ef = 0.5 * (eig[order[k]] + eig[order[k - 1]])
        if abs(ne - _electron_count(ef, eig, kT)) <= ELEC_TOL_MAX:
            f, eb, ts = _electron_fill(ef, eig, kT)
            return 2 * f, ef, 2 * eb, 2 * ts, 2 * (eb - 0.5 * ts)
    upper = eig.max() + 0.01
    lower = eig.min() - 0.01
    nmax = _electron_count(upper, eig, kT)
    nmin = _electron_count(lower, eig, kT)
    while nmin > ne:
        lower = 2.0 * (lower - upper) + lower
        nmin = _electron_count(lower, eig, kT)
    while nmax < ne:
        upper = 2.0 * (upper - lower) + lower
        nmax = _electron_count(upper, eig, kT)
    ef = 0.5 * (upper + lower)
    ncur = _electron_count(ef, eig, kT)
    eps2 = 2.0 * EPS
    while abs(ne - ncur) > ELEC_TOL and abs(upper - lower) >= max(abs(ef) * eps2, eps2):
        if (nmax >= nmin) == (ne >= ncur):
            lower, nmin = ef, ncur
        else:
            upper, nmax = ef, ncur
        ef = 0.5 * (upper + lower)
        ncur = _electron_count(ef, eig, kT)
    if abs(ne - ncur) > ELEC_TOL_MAX:
        raise RuntimeError("Fermi level search did not converge.")
    f, eb, ts = _electron_fill(ef, eig, kT)
    if ne > eps2 and abs(ne - ncur) > eps2:
        f = f * ne / ncur
    return 2 * f, ef, 2 * eb, 2 * ts, 2 * (eb - 0.5 * ts)


# --------------------------------------------------------------------------------------------
# Broyden mixer  (mixer/broydenmixer.F90:180-342)
# --------------------------------------------------------------------------------------------

class BroydenMixer:
    """Modified Broyden mixing (Johnson 1988) as ``modifiedBroydenMixing``."""

    def __init__(self, n, alpha=0.2, omega0=0.01, min_weight=1.0, max_weight=1.0e5, weight_fac=1.0e-2,
                 max_iter=200):
        self.n = n
        self.alpha, self.omega0 = alpha, omega0
        self.minw, self.maxw, self.wfac = min_weight, max_weight, weight_fac
        self.it = 0
        self.ww = np.zeros(max_iter)
        self.aa = np.zeros((max_iter, max_iter))
        self.dF = np.zeros((n, max_iter))
        self.uu = np.zeros((n, max_iter))
        self.qlast = np.zeros(n)
        self.dlast = np.zeros(n)

    def mix(self, qin: np.ndarray, qdiff: np.ndarray) -> np.ndarray:
        self.it += 1
        nn = self.it
        q = qin.copy()
        if nn == 1:
            self.qlast = q.copy()
            self.dlast = qdiff.copy()
            return q + self.alpha * qdiff
        n1 = nn - 1
        w = math.sqrt(float(np.dot(qdiff, qdiff)))
        w = self.wfac / w if w > self.wfac / self.maxw else self.maxw
        w = max(w, self.minw)
        self.ww[n1 - 1] = w
        dfu = qdiff - self.dlast
        inv = max(math.sqrt(float(np.dot(dfu, dfu))), EPS)
        inv = 1.0 / inv
        dfu = inv * dfu
        cc = np.zeros(n1)
        for ii in range(nn - 2):
            self.aa[ii, n1 - 1] = np.dot(self.dF[:, ii], dfu)
            self.aa[n1 - 1, ii] = self.aa[ii, n1 - 1]
            cc[ii] = self.ww[ii] * np.dot(self.dF[:, ii], qdiff)
        self.aa[n1 - 1, n1 - 1] = 1.0
        cc[n1 - 1] = self.ww[n1 - 1] * np.dot(dfu, qdiff)
        beta = np.zeros((n1, n1))
        for ii in range(n1):
            beta[:n1, ii] = self.ww[:n1] * self.ww[ii] * self.aa[:n1, ii]
            beta[ii, ii] += self.omega0 ** 2
        cc = scipy.linalg.lu_solve(scipy.linalg.lu_factor(beta), cc, trans=1)
        self.dF[:, n1 - 1] = dfu
        dfu = self.alpha * dfu + inv * (q - self.qlast)
        self.qlast = q.copy()
        self.dlast = qdiff.copy()
        q = q + self.alpha * qdiff
        for ii in range(nn - 2):
            q = q - self.ww[ii] * cc[ii] * self.uu[:, ii]
        q = q - self.ww[n1 - 1] * cc[n1 - 1] * dfu
        self.uu[:, n1 - 1] = dfu
        return q


# --------------------------------------------------------------------------------------------
# Repulsive  (repulsive/twobodyrep.F90:142-204, splinerep.F90:136-190)
# --------------------------------------------------------------------------------------------

def repulsive(params, species, nl: NeighbourList):
    """Pair repulsive energy and gradient over the half neighbour list."""
    nat = len(species)
    e = 0.0
    grad = np.zeros((nat, 3))
    for i in range(nat):
        for k in range(1, len(nl.neigh[i])):
            j = nl.neigh[i][k]
            jc = nl.img2cent[j]
            rep = params.rep[(species[i], species[jc])]
            r = math.sqrt(nl.dist2[i][k])
            en, de = rep.value(r)
            e += en if jc != i else 0.5 * en      # twobodyrep.F90:160-166
            if jc != i:
                vec = nl.coords_all[i] - nl.coords_all[j]
                g = de * vec / r
                grad[i] += g
                grad[jc] -= g
    return e, grad


# --------------------------------------------------------------------------------------------
# SCC driver  (dftbplus/main.F90:1429-1563) and forces (main.F90:7058-7440, forces.F90:548-672)
# --------------------------------------------------------------------------------------------

@dataclasses.dataclass
class Result:
    energy_mermin: float
    energy_total: float
    e_nonscc: float
    e_scc: float
    e_rep: float
    ts: float
    eband: float
    ef: float
    charges: np.ndarray          # gross charges q0 - q per atom (dftbp_get_gross_charges)
    qout_atom: np.ndarray
    eigenvalues: np.ndarray
    fillings: np.ndarray
    n_iter: int
    converged: bool
    rho: np.ndarray              # sparse density matrix
    erho: Optional[np.ndarray] = None
    gradients: Optional[np.ndarray] = None
    vorb: Optional[np.ndarray] = None


class Calculator:
    """Serial CPU restatement of one SCC-DFTB energy + gradient evaluation."""

    def __init__(self, params, geo, opts, gamma: Optional[np.ndarray] = None):
        """gamma: a precomputed atomic gamma matrix to use instead of building it (the CPU timing
        of an SCC iteration, whose cost does not depend on its values)."""
        self.params, self.geo, self.opts = params, geo, opts
        sp = geo.species
        self.norb_atom = np.array([params.species[s].norb for s in sp])
        self.off = np.zeros(len(sp) + 1, dtype=np.int64)
        self.off[1:] = np.cumsum(self.norb_atom)
        cutoff = max(params.sk_cutoff(), params.rep_cutoff())
        # the short gamma cutoff may exceed the SK cutoff
        u = [s.hubbu[0] for s in params.species]
        gcut = max(expgamma_cutoff(a, b) for a in u for b in u)
        self.cutoff = max(cutoff, gcut)
        self.nl = neighbour_list(geo.coords, geo.latvecs, self.cutoff)
        self.sp = sparse_descriptor(self.nl, self.norb_atom, params.sk_cutoff()).index_maps(self.off)
        self.h0, self.s = build_h0_s(params, sp, self.nl, self.sp)
        self.gamma = (gamma_matrix(params, sp, self.nl, geo.coords, geo.latvecs, opts.ewald_tol)
                      if gamma is None else gamma)
        self.q0 = np.array([sum(params.species[s].occ) for s in sp])
        self.nel = float(self.q0.sum())
        self.erep, self.grad_rep = repulsive(params, sp, self.nl)
        self.kpts = None
        if getattr(opts, "kpoints", None) is not None:
            if geo.latvecs is None:
                raise ValueError("k-points need a periodic geometry")
            self.kpts = np.asarray(opts.kpoints, dtype=float).reshape(-1, 3)
            w = (np.ones(len(self.kpts)) if opts.kweights is None
                 else np.asarray(opts.kweights, dtype=float))
            self.kw = w / w.sum()
            self.phases = [block_phases(self.nl, self.sp, geo.coords, geo.latvecs, k)
                           for k in self.kpts]
            self.Sk = [hermitian_from_lower(unpack_k(self.s, self.sp, self.off, ph))
                       for ph in self.phases]

    def _density(self, h, S):
        """Eigenpairs, Fermi filling and the sparse density matrix of the Hamiltonian h (sparse):
        Gamma point in real arithmetic, else one complex Hermitian problem per k-point and the
        k-weighted density (fullDensityMatrix_cmplx + packHS_cmplx_kpts)."""
        if self.kpts is None:
            eig, C = self.solve(h, S)
            f, ef, eband, ts, e0 = fermi_filling(eig, self.nel, self.opts.kT)
            occ = levels(f)
            P = (C[:, occ] * f[occ]) @ C[:, occ].T
            return eig, C, f, ef, eband, ts, pack(P, self.params, self.geo.species, self.sp, self.off)
        eigs, Cs = [], []
        for ph, Sk in zip(self.phases, self.Sk):
            Hk = hermitian_from_lower(unpack_k(h, self.sp, self.off, ph))
            e, C = scipy.linalg.eigh(Hk, Sk, lower=True)
            eigs.append(e)
            Cs.append(C)
        eig = np.array(eigs)
        f, ef, eband, ts, e0 = fermi_filling_k(eig, self.nel, self.opts.kT, self.kw)
        rho = np.zeros(self.sp.nnz)
        for k in range(len(self.kpts)):
            occ = levels(f[k])
            P = (Cs[k][:, occ] * f[k][occ]) @ Cs[k][:, occ].conj().T
            rho += pack_k(P, self.sp, self.phases[k], self.kw[k])
        return eig, Cs, f, ef, eband, ts, rho

    def _energy_density(self, C, eig, f):
        """Sparse energy-weighted density matrix (sum_i f_i e_i c_i c_i^+, k-weighted)."""
        if self.kpts is None:
            occ = levels(f)
            W = (C[:, occ] * (f[occ] * eig[occ])) @ C[:, occ].T
            return pack(W, self.params, self.geo.species, self.sp, self.off)
        erho = np.zeros(self.sp.nnz)
        for k in range(len(self.kpts)):
            occ = levels(f[k])
            W = (C[k][:, occ] * (f[k][occ] * eig[k][occ])) @ C[k][:, occ].conj().T
            erho += pack_k(W, self.sp, self.phases[k], self.kw[k])
        return erho

    def dense(self, x):
        return symmetrize_lower(unpack(x, self.params, self.geo.species, self.sp, self.off))

    def orbital_potential(self, dq_atom):
        v_atom = self.gamma @ dq_atom
        return np.repeat(v_atom, self.norb_atom), v_atom

    def solve(self, h_sparse, S=None):
        H = self.dense(h_sparse)
        if S is None:
            S = self.dense(self.s)
        eig, C = scipy.linalg.eigh(H, S, lower=True)
        return eig, C

    def scc(self) -> Result:
        o = self.opts
        nat = len(self.geo.species)
        qin = self.q0.copy()
        mixer = BroydenMixer(nat, o.mix_param, o.omega0, o.min_weight, o.max_weight, o.weight_fac,
                             o.max_scc_iter + 1)
        S = self.dense(self.s) if self.kpts is None else None
        converged = False
        for it in range(1, o.max_scc_iter + 1):
            dqin = qin - self.q0
            vorb, _ = self.orbital_potential(dqin)
            h = add_shift(self.h0, self.s, vorb, self.geo.species, self.sp, self.off)
            eig, C, f, ef, eband, ts, rho = self._density(h, S)
            qorb = mulliken(self.s, rho, self.geo.species, self.sp, self.off)
            qout = np.add.reduceat(qorb, self.off[:-1])
            dqout = qout - self.q0
            e_nonscc = float(np.sum(mulliken(self.h0, rho, self.geo.species, self.sp, self.off)))
            vout_orb, vout_atom = self.orbital_potential(dqout)
            e_scc = 0.5 * float(np.dot(vout_atom, dqout))
            diff = qout - qin
            err = float(np.max(np.abs(diff)))
            converged = err < o.scc_tol and it >= o.min_scc_iter
            if not converged and it != o.max_scc_iter:
                qin = mixer.mix(qin, diff)
            if converged:
                break
        etot = e_nonscc + e_scc + self.erep
        res = Result(etot - ts, etot, e_nonscc, e_scc, self.erep, ts, eband, ef, self.q0 - qout, qout,
                     eig, f, it, converged, rho, vorb=vout_orb)
        res._C, res._eig, res._f, res._dqout = C, eig, f, dqout
        return res

    def gradients(self, res: Result, analytic: bool = False) -> np.ndarray:
        """Total energy gradient: 2 Sum(rho dH0 - W dS) + shift dS term (forces.F90:548-672),
        gamma derivative (shortgamma.F90:349-411, coulomb.F90:1140-1187/1380-1494), repulsive."""
        erho = self._energy_density(res._C, res._eig, res._f)
        res.erho = erho
        sp = self.sp
        spc = self.geo.species
        nat = len(spc)
        grad = np.zeros((nat, 3))
        vorb = res.vorb
        delta = self.opts.fd_delta
        for (a, b), idx in _pair_groups(self.params, spc, sp).items():
            n1, n2 = self.params.species[a].norb, self.params.species[b].norb
            vec = self.nl.coords_all[sp.img2[idx]] - self.nl.coords_all[sp.at1[idx]]
            dH = np.zeros((len(idx), 3, n2, n1))
            dS = np.zeros((len(idx), 3, n2, n1))
            if analytic:
                dH = sk_block_derivative(self.params, self.params.ham, a, b, vec)
                dS = sk_block_derivative(self.params, self.params.over, a, b, vec)
            else:
                # getFirstDerivFiniteDiff (nonscc.F90:428-484): derivative w.r.t. atom 1
                for c in range(3):
                    vp = vec.copy()
                    vm = vec.copy()
                    vp[:, c] += delta      # ii = 1: vect - (2*1-3)*delta
                    vm[:, c] -= delta      # ii = 2
                    hp = diatomic_blocks(self.params, self.params.ham, a, b, vp)
                    hm = diatomic_blocks(self.params, self.params.ham, a, b, vm)
                    spp = diatomic_blocks(self.params, self.params.over, a, b, vp)
                    smm = diatomic_blocks(self.params, self.params.over, a, b, vm)
                    dH[:, c] = (hm - hp) / (2.0 * delta)
                    dS[:, c] = (smm - spp) / (2.0 * delta)
            cols = sp.start[idx][:, None] + np.arange(n1 * n2)[None, :]
            rb = res.rho[cols].reshape(len(idx), n1, n2).transpose(0, 2, 1)
            wb = erho[cols].reshape(len(idx), n1, n2).transpose(0, 2, 1)
            at1, at2 = sp.at1[idx], sp.at2[idx]
            v1 = np.stack([vorb[self.off[x]:self.off[x] + n1] for x in at1])
            v2 = np.stack([vorb[self.off[x]:self.off[x] + n2] for x in at2])
            vsum = 0.5 * (v1[:, None, :] + v2[:, :, None])
            g = 2.0 * (np.einsum("pij,pcij->pc", rb, dH) - np.einsum("pij,pcij->pc", wb, dS)
                       + np.einsum("pij,pcij->pc", rb * vsum, dS))
            np.add.at(grad, at1, g)
            np.add.at(grad, at2, -g)
        grad += self.gamma_gradient(res._dqout)
        grad += self.grad_rep
        res.gradients = grad
        return grad

    def gamma_gradient(self, dq):
        """d/dR of 1/2 dq^T gamma dq: 1/R or Ewald part and the short-range part."""
        nat = len(dq)
        c = self.geo.coords
        if self.geo.latvecs is None:
            d = c[:, None, :] - c[None, :, :]
            r = np.sqrt(np.sum(d * d, axis=2))
            np.fill_diagonal(r, np.inf)
            grad = np.sum((-dq[:, None] * dq[None, :] / r ** 3)[:, :, None] * d, axis=1)
        else:
            grad = invr_ewald_deriv(c, self.geo.latvecs, dq, self.opts.ewald_tol)
        u = [s.hubbu[0] for s in self.params.species]
        spc = self.geo.species
        cut = {(a, b): expgamma_cutoff(u[b], u[a]) for a in range(self.params.nspecies) for b in range(self.params.nspecies)}
        for i in range(nat):
            for k in range(1, len(self.nl.neigh[i])):
                j = self.nl.neigh[i][k]
                jc = self.nl.img2cent[j]
                if jc == i:
                    continue
                r = math.sqrt(self.nl.dist2[i][k])
                if r > cut[(spc[jc], spc[i])]:
                    continue
                vec = self.nl.coords_all[i] - self.nl.coords_all[j]
                gp = -expgamma_prime(r, u[spc[jc]], u[spc[i]]) * dq[i] * dq[jc] * vec / r
                grad[i] += gp
                grad[jc] -= gp
        return grad


def sk_block_derivative(params, table, sp1, sp2, vec, h=1.0e-6):
    """Derivative of the diatomic blocks with respect to atom 1 using a high-order central
    difference of the interpolant (used to check the analytic adjoint)."""
    out = np.zeros((len(vec), 3) + diatomic_blocks(params, table, sp1, sp2, vec[:1]).shape[1:])
    for c in range(3):
        acc = 0.0
        for k, w in ((1, 8.0), (2, -1.0)):
            vp = vec.copy()
            vm = vec.copy()
            vp[:, c] += k * h
            vm[:, c] -= k * h
            acc = acc + w * (diatomic_blocks(params, table, sp1, sp2, vm) - diatomic_blocks(params, table, sp1, sp2, vp))
        out[:, c] = acc / (12.0 * h)
    return out
