"""Synthetic Slater-Koster parameters and geometries for the five benchmark workloads.

SK recipe: 500 grid rows at 0.02 bohr (last row at 10 bohr, SK cutoff 11 bohr with the 1 bohr
tail). Every nonzero integral (ss sigma, sp sigma, pp sigma, pp pi, for H and S) has the form
c (1 + a r) exp(-b r) with b drawn from [0.9, 1.4] / bohr. ss, pp sigma and pp pi are the same in
the A-B and B-A files, sp sigma differs (it couples s on the first atom to p on the second).
On-site energies: s in [-0.6, -0.2] Ha, p in [-0.35, -0.1] Ha; Hubbard U in [0.3, 0.5] Ha.
Repulsive: A exp(-B r) (1 - r / rc)^3 written as a spline with an exponential head.
"""
from __future__ import annotations

import math
import os
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import skf as skfmod
from .model import Geometry, Options, Parameters, Species

ELEMENTS = {
    # name: (shells, neutral occupations, mass in amu)
    "H": ([0], [1.0], 1.008),
    "C": ([0, 1], [2.0, 2.0], 12.011),
    "N": ([0, 1], [2.0, 3.0], 14.007),
    "O": ([0, 1], [2.0, 4.0], 15.999),
    "Si": ([0, 1], [2.0, 2.0], 28.086),
}

SK_DIST = 0.02
SK_ROWS = 500
REP_CUTOFF = 4.0


def _integral(rng, r, sign, cmin, cmax):
    c = sign * rng.uniform(cmin, cmax)
    a = rng.uniform(0.2, 0.6)
    b = rng.uniform(0.9, 1.4)
    return c * (1.0 + a * r) * np.exp(-b * r)


def _spline_rep(rng) -> skfmod.SplineRep:
    """Spline representation of f(r) = A exp(-B r) (1 - r/rc)^3 on [1, rc)."""
    A = rng.uniform(2.0, 4.0)
    B = rng.uniform(1.2, 1.8)
    rc = REP_CUTOFF

    def f(r, d=0):
        # value, first and second derivative of A e^{-Br} (1 - r/rc)^3
        e = A * math.exp(-B * r)
        u = 1.0 - r / rc
        g = [u ** 3, -3.0 * u ** 2 / rc, 6.0 * u / rc ** 2]
        ed = [e, -B * e, B * B * e]
        if d == 0:
            return ed[0] * g[0]
        if d == 1:
            return ed[1] * g[0] + ed[0] * g[1]
        return ed[2] * g[0] + 2.0 * ed[1] * g[1] + ed[0] * g[2]

    nodes = np.arange(1.0, rc - 0.5 + 1e-12, 0.25)
    xs = list(nodes)
    sp = []
    for x0, x1 in zip(xs[:-1], xs[1:]):
        h = x1 - x0
        y0, y1, d0, d1 = f(x0), f(x1), f(x0, 1), f(x1, 1)
        c2 = (3.0 * (y1 - y0) / h - 2.0 * d0 - d1) / h
        c3 = (2.0 * (y0 - y1) / h + d0 + d1) / (h * h)
        sp.append([y0, d0, c2, c3])
    # last segment [xs[-1], rc): quintic matching value/1st/2nd derivative at both ends
    x0 = xs[-1]
    h = rc - x0
    y0, d0, s0 = f(x0), f(x0, 1), f(x0, 2)
    M = np.array([[h ** 3, h ** 4, h ** 5], [3 * h ** 2, 4 * h ** 3, 5 * h ** 4], [6 * h, 12 * h ** 2, 20 * h ** 3]])
    rhs = -np.array([y0 + d0 * h + 0.5 * s0 * h * h, d0 + s0 * h, s0])
    c345 = np.linalg.solve(M, rhs)
    last = np.array([y0, d0, 0.5 * s0, *c345])
    a1 = -f(1.0, 1) / f(1.0)
    a2 = math.log(f(1.0)) + a1 * 1.0
    return skfmod.SplineRep(rc, np.array([a1, a2, 0.0]), np.array(xs), np.array(sp), last)


def make_parameters(elements: List[str], seed: int) -> Tuple[Parameters, Dict[Tuple[int, int], skfmod.SKFile]]:
    """Draw synthetic SK files for every ordered pair of ``elements``."""
    rng = np.random.default_rng(seed)
    r = SK_DIST * np.arange(1, SK_ROWS + 1)
    species = []
    for el in elements:
        shells, occ, mass = ELEMENTS[el]
        es = rng.uniform(-0.6, -0.2)
        ep = rng.uniform(-0.35, -0.1)
        u = rng.uniform(0.3, 0.5)
        onsite = [es] + ([ep] if len(shells) > 1 else [])
        species.append(Species(el, list(shells), onsite, [u] * len(shells), list(occ), mass))
    files: Dict[Tuple[int, int], skfmod.SKFile] = {}
    ns = len(elements)
    for i in range(ns):
        for j in range(i, ns):
            ham_ij = np.zeros((SK_ROWS, 10))
            over_ij = np.zeros((SK_ROWS, 10))
            col = skfmod.OLD_COLUMNS.index
            # symmetric integrals shared by A-B and B-A
            ham_ij[:, col("sss")] = _integral(rng, r, -1.0, 0.2, 0.6)
            over_ij[:, col("sss")] = _integral(rng, r, 1.0, 0.15, 0.3)
            ham_ij[:, col("pps")] = _integral(rng, r, 1.0, 0.2, 0.6)
            over_ij[:, col("pps")] = _integral(rng, r, -1.0, 0.15, 0.3)
            ham_ij[:, col("ppp")] = _integral(rng, r, -1.0, 0.1, 0.4)
            over_ij[:, col("ppp")] = _integral(rng, r, 1.0, 0.1, 0.25)
            ham_ji = ham_ij.copy()
            over_ji = over_ij.copy()
            ham_ij[:, col("sps")] = _integral(rng, r, 1.0, 0.2, 0.6)
            over_ij[:, col("sps")] = _integral(rng, r, -1.0, 0.15, 0.3)
            if i == j:
                ham_ji = ham_ij
                over_ji = over_ij
            else:
                ham_ji[:, col("sps")] = _integral(rng, r, 1.0, 0.2, 0.6)
                over_ji[:, col("sps")] = _integral(rng, r, -1.0, 0.15, 0.3)
            rep = _spline_rep(rng)
            if i == j:
                sp = species[i]
                onsite3 = np.zeros(3)
                hub3 = np.zeros(3)
                occ3 = np.zeros(3)
                for ish, l in enumerate(sp.shells):
                    onsite3[2 - l] = sp.onsite[ish]
                    hub3[2 - l] = sp.hubbu[ish]
                    occ3[2 - l] = sp.occ[ish]
                files[(i, i)] = skfmod.SKFile(SK_DIST, ham_ij, over_ij, rep, onsite3, hub3, occ3, sp.mass)
            else:
                files[(i, j)] = skfmod.SKFile(SK_DIST, ham_ij, over_ij, rep)
                files[(j, i)] = skfmod.SKFile(SK_DIST, ham_ji, over_ji, rep)
    return Parameters.from_skfiles(species, files), files


def write_skfiles(directory: str, elements: List[str], files) -> Dict[Tuple[int, int], str]:
    os.makedirs(directory, exist_ok=True)
    paths = {}
    for (i, j), f in files.items():
        p = os.path.join(directory, f"{elements[i]}-{elements[j]}.skf")
        skfmod.write_skf(p, f)
        paths[(i, j)] = p
    return paths


def _assign(rng, n, fractions):
    counts = [int(round(f * n)) for f in fractions]
    counts[-1] = n - sum(counts[:-1])
    lab = np.concatenate([np.full(c, k) for k, c in enumerate(counts)])
    return rng.permutation(lab)


def cluster_216(seed: int = 0, nside: int = 6) -> Tuple[Geometry, List[str]]:
    """216-atom C/H/O cluster: jittered 6x6x6 cubic lattice, 2.6 bohr, +-0.3 bohr noise;
    0.45 H, 0.35 C, 0.20 O (``nside`` gives smaller clusters of the same recipe)."""
    rng = np.random.default_rng(seed)
    g = np.arange(nside) * 2.6
    pos = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    pos = pos + rng.uniform(-0.3, 0.3, pos.shape)
    elements = ["H", "C", "O"]
    spec = _assign(rng, len(pos), [0.45, 0.35, 0.20])
    return Geometry(pos, spec.astype(np.int64), None), elements


def diamond_supercell(ncell: int = 8, seed: int = 1, a: float = 6.74, jitter: float = 0.1,
                      subst: float = 0.10) -> Tuple[Geometry, List[str]]:
    """Periodic carbon-rich diamond supercell (ncell^3 cubic cells, 8 atoms each), jittered, with
    a fraction ``subst`` of sites randomly replaced by H or O (half each)."""
    rng = np.random.default_rng(seed)
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                      [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    cells = np.array(np.meshgrid(*[np.arange(ncell)] * 3, indexing="ij")).reshape(3, -1).T
    frac = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    pos = frac * a + rng.uniform(-jitter, jitter, (len(frac), 3))
    n = len(pos)
    spec = np.zeros(n, dtype=np.int64)          # C
    nsub = int(round(subst * n))
    sites = rng.choice(n, nsub, replace=False)
    spec[sites[: nsub // 2]] = 1                 # H
    spec[sites[nsub // 2:]] = 2                  # O
    lat = np.eye(3) * (ncell * a)
    return Geometry(pos, spec, lat), ["C", "H", "O"]


def water_box(nmol: int = 1000, seed: int = 2) -> Tuple[Geometry, List[str]]:
    """Liquid-water-like periodic box at 1 g/cm^3: O-H 1.81 bohr, H-O-H 104.5 deg, random
    orientations, molecule centres on a jittered cubic grid."""
    rng = np.random.default_rng(seed)
    bohr_cm = 0.529177210903e-8
    vol_cm3 = nmol * 18.015 / 6.02214076e23
    L = (vol_cm3 ** (1.0 / 3.0)) / bohr_cm
    m = int(math.ceil(nmol ** (1.0 / 3.0)))
    sp = L / m
    grid = np.array(np.meshgrid(*[np.arange(m)] * 3, indexing="ij")).reshape(3, -1).T[:nmol]
    centres = (grid + 0.5) * sp + rng.uniform(-0.15 * sp, 0.15 * sp, (nmol, 3))
    th = math.radians(104.5) / 2
    local = np.array([[0, 0, 0], [1.81 * math.sin(th), 0, 1.81 * math.cos(th)],
                      [-1.81 * math.sin(th), 0, 1.81 * math.cos(th)]])
    pos = []
    spec = []
    for c in centres:
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        pos.append(c + local @ R.T)
        spec += [0, 1, 1]
    return Geometry(np.concatenate(pos), np.array(spec, dtype=np.int64), np.eye(3) * L), ["O", "H"]


def molecule_batch(nmol: int = 4096, seed: int = 3) -> Tuple[List[Geometry], List[str]]:
    """Independent non-periodic bonded-chain molecules of uneven size: target orbital counts
    log-uniform in [8, 220], 3-64 atoms, H/C/N/O, bond lengths 2.0-2.9 bohr."""
    rng = np.random.default_rng(seed)
    elements = ["H", "C", "N", "O"]
    mols = []
    for _ in range(nmol):
        target = math.exp(rng.uniform(math.log(8), math.log(220)))
        spec = []
        norb = 0
        while (norb < target and len(spec) < 64) or len(spec) < 3:
            if rng.uniform() < 0.35:
                spec.append(0)
                norb += 1
            else:
                spec.append(int(rng.integers(1, 4)))
                norb += 4
        pos = [np.zeros(3)]
        d = np.array([1.0, 0.0, 0.0])
        for k in range(1, len(spec)):
            for _try in range(100):
                nd = d + rng.normal(scale=0.8, size=3)
                nd /= np.linalg.norm(nd)
                p = pos[-1] + rng.uniform(2.0, 2.9) * nd
                if min(np.linalg.norm(p - q) for q in pos) > 1.9:
                    break
            pos.append(p)
            d = nd
        mols.append(Geometry(np.array(pos), np.array(spec, dtype=np.int64), None))
    return mols, elements


def amorphous_network(natom: int = 2000, seed: int = 4, density: float = 2.3) -> Tuple[Geometry, List[str]]:
    """Periodic amorphous C/Si/H/O network by random sequential addition at ``density`` g/cm^3."""
    rng = np.random.default_rng(seed)
    elements = ["C", "Si", "H", "O"]
    spec = _assign(rng, natom, [0.45, 0.35, 0.05, 0.15])
    masses = np.array([ELEMENTS[e][2] for e in elements])
    bohr_cm = 0.529177210903e-8
    vol = masses[spec].sum() / 6.02214076e23 / density
    L = (vol ** (1.0 / 3.0)) / bohr_cm
    pos = np.zeros((natom, 3))
    for k in range(natom):
        for _try in range(1000):
            p = rng.uniform(0, L, 3)
            if k == 0:
                break
            dv = pos[:k] - p
            dv -= L * np.round(dv / L)
            if np.min(np.einsum("ij,ij->i", dv, dv)) > 2.4 ** 2:
                break
        pos[k] = p
    return Geometry(pos, spec.astype(np.int64), np.eye(3) * L), elements


def workload(name: str):
    """(geometry, parameters, options) of a named workload."""
    if name == "cluster216":
        geo, el = cluster_216(0)
        params, _ = make_parameters(el, 100)
        return geo, params, Options(kT=1.0e-8, scc_tol=1.0e-8)
    if name == "cluster27":
        geo, el = cluster_216(0, 3)
        params, _ = make_parameters(el, 100)
        return geo, params, Options(kT=1.0e-8, scc_tol=1.0e-8)
    if name.startswith("diamond"):
        ncell = int(name[len("diamond"):] or 8)
        geo, el = diamond_supercell(ncell, 1)
        params, _ = make_parameters(el, 101)
        return geo, params, Options(kT=0.01, scc_tol=1.0e-8)
    if name.startswith("water"):
        # kT = 0.003 Ha: with this synthetic parameter set the frontier levels of the water box
        # are near-degenerate and the SCC cycle at kT = 1e-8 does not converge in 200 iterations.
        # Broyden mixing parameter 0.05: the first iteration moves about 1.8 e between the
        # atoms, and at 0.2 the 1000-molecule box oscillates without converging (300 iterations)
        geo, el = water_box(int(name[len("water"):] or 1000), 2)
        params, _ = make_parameters(el, 102)
        return geo, params, Options(kT=3.0e-3, scc_tol=1.0e-8, mix_param=0.05, max_scc_iter=300)
    if name.startswith("molecules"):
        # the first N molecules of the 4096-molecule batch recipe (seed 3), as a list of
        # geometries. kT = 0.01 Ha: with this synthetic parameter set most of these small
        # (often open-shell) molecules do not converge at kT = 1e-8 within 100 iterations
        mols, el = molecule_batch(int(name[len("molecules"):] or 4096), 3)
        params, _ = make_parameters(el, 103)
        return mols, params, Options(kT=1.0e-2, scc_tol=1.0e-8, max_scc_iter=200)
    if name.startswith("amorphous"):
        # the 2000-atom C/Si/H/O network (seed 4, 2.3 g/cm^3) or a smaller cell of the same
        # recipe (amorphousN: N atoms, same seed, density and parameter set)
        geo, el = amorphous_network(int(name[len("amorphous"):] or 2000), 4)
        params, _ = make_parameters(el, 104)
        return geo, params, Options(kT=1.0e-3, scc_tol=1.0e-8)
    raise KeyError(name)
