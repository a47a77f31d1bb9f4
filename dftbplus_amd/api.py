"""ctypes binding of libdftbhip and a ``DftbPlus`` calculator mirroring DFTB+'s Python API.

The calculator follows ``tools/pythonapi/src/dftbplus/dftbplus.py:35-399`` (``set_geometry``,
``get_energy``, ``get_gradients``, ``get_gross_charges``, ``get_nr_atoms``, ``close``); instead of
an HSD file it takes the parsed parameters (``model.Parameters``), geometry and options. The
library is mandatory: importing works without it, but every call raises ``DftbHipError`` when the
shared library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from .model import Geometry, Options, Parameters

LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdftbhip.so")
MAXSP = 8


class DftbHipError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the HIP library (raises when it is missing or cannot be loaded)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise DftbHipError(f"{LIBPATH} not found: build it with dftbplus_amd.build.build()")
        _lib = C.CDLL(LIBPATH)
        _setup(_lib)
        # the library's resident buffers are released while the HIP runtime is still up
        import atexit
        atexit.register(_lib.dftbhip_finalize)
    return _lib


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Params(C.Structure):
    _fields_ = [
        ("nspecies", C.c_int),
        ("nshell", C.c_int * MAXSP),
        ("angshell", (C.c_int * 2) * MAXSP),
        ("onsite", (C.c_double * 2) * MAXSP),
        ("hubbu", C.c_double * MAXSP),
        ("occ", (C.c_double * 2) * MAXSP),
        ("nrows", C.c_int),
        ("dist", C.c_double),
        ("ninteg", C.c_int * (MAXSP * MAXSP)),
        ("htab", _dp * (MAXSP * MAXSP)),
        ("stab", _dp * (MAXSP * MAXSP)),
        ("rep_nint", C.c_int * (MAXSP * MAXSP)),
        ("rep_cutoff", C.c_double * (MAXSP * MAXSP)),
        ("rep_expc", (C.c_double * 3) * (MAXSP * MAXSP)),
        ("rep_xs", _dp * (MAXSP * MAXSP)),
        ("rep_coef", _dp * (MAXSP * MAXSP)),
        ("rep_last", _dp * (MAXSP * MAXSP)),
    ]


class COptions(C.Structure):
    _fields_ = [
        ("kT", C.c_double),
        ("scc_tol", C.c_double),
        ("max_scc_iter", C.c_int),
        ("min_scc_iter", C.c_int),
        ("mix_param", C.c_double),
        ("omega0", C.c_double),
        ("min_weight", C.c_double),
        ("max_weight", C.c_double),
        ("weight_fac", C.c_double),
        ("fd_delta", C.c_double),
        ("ewald_tol", C.c_double),
        ("force_mode", C.c_int),
        ("precision", C.c_int),
        ("refine_below", C.c_double),
        ("refine_steps", C.c_int),
        ("tridiag", C.c_int),
        ("tail_below", C.c_double),
        ("tail_steps", C.c_int),
    ]


EXPORTS = {
    "dftbhip_get_gpus_available": [_ip],
    "dftbhip_init": [C.c_int],
    "dftbhip_finalize": [],
    "dftbhip_last_error": [C.c_char_p, C.c_int],
    "dftbhip_sygvd": [C.c_int, C.c_char, C.c_char, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp],
    "dftbhip_dgemm": [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int,
                      _dp, C.c_int, C.c_double, _dp, C.c_int],
    "dftbhip_dstedc": [C.c_int, _dp, _dp, _dp, C.c_int],
    "dftbhip_probe_mfma_f64": [_dp],
    "dftbhip_probe_mfma_f16": [_dp],
    "dftbhip_gemm_f16_tn": [C.c_int, C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_int,
                            C.c_void_p, C.c_int, C.c_void_p, C.c_int],
    "dftbhip_kernel_profile": [C.c_int],
    "dftbhip_kernel_stats": [_dp],
    "dftbhip_dm_real": [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int],
    "dftbhip_edm_real": [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, _dp, C.c_int],
    "dftbhip_dm_cmplx": [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int],
    "dftbhip_hegvd": [C.c_int, C.c_char, C.c_char, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp],
    "dftbhip_edm_cmplx": [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, _dp, C.c_int],
    "dftbhip_sk_create": [C.POINTER(C.c_void_p), C.POINTER(Params)],
    "dftbhip_sk_destroy": [C.c_void_p],
    "dftbhip_build_h0_s": [C.c_void_p, C.c_int, C.c_int, _dp, _ip, _ip, C.c_int, _ip, _ip, _ip,
                           C.c_int, _dp, _dp],
    "dftbhip_unpack_hs_real": [C.c_int, _ip, _ip, C.c_int, _ip, _ip, _ip, C.c_int, _dp, _dp,
                               C.c_int],
    "dftbhip_pack_hs_real": [C.c_int, _ip, _ip, C.c_int, _ip, _ip, _ip, C.c_int, _dp, C.c_int,
                             _dp],
    "dftbhip_mulliken": [C.c_int, _ip, _ip, C.c_int, _ip, _ip, _ip, C.c_int, _dp, _dp, _dp],
    "dftbhip_ctx_create": [C.POINTER(C.c_void_p), C.POINTER(Params), C.POINTER(COptions)],
    "dftbhip_ctx_destroy": [C.c_void_p],
    "dftbhip_ctx_set_geometry": [C.c_void_p, C.c_int, _dp, _ip, _dp],
    "dftbhip_ctx_scc": [C.c_void_p, _dp, _dp, _ip, _ip],
    "dftbhip_ctx_scc_begin": [C.c_void_p],
    "dftbhip_ctx_set_charge_restart": [C.c_void_p, C.c_int],
    "dftbhip_ctx_set_input_charges": [C.c_void_p, C.c_int, _dp, C.c_int],
    "dftbhip_ctx_get_input_charges": [C.c_void_p, C.c_int, _dp],
    "dftbhip_ctx_scc_step": [C.c_void_p, _ip, _dp],
    "dftbhip_ctx_scc_end": [C.c_void_p, _dp, _dp, _ip, _ip],
    "dftbhip_ctx_gradients": [C.c_void_p, _dp],
    "dftbhip_ctx_scc_state": [C.c_void_p, _dp, C.c_int],
    "dftbhip_ctx_eigenvalues": [C.c_void_p, _dp, _dp, C.c_int],
    "dftbhip_ctx_eigenvalues_k": [C.c_void_p, C.c_int, _dp, _dp, C.c_int],
    "dftbhip_ctx_orbital_populations": [C.c_void_p, _dp, C.c_int],
    "dftbhip_ctx_nr_orbitals": [C.c_void_p],
    "dftbhip_ctx_timings": [C.c_void_p, _dp, C.c_int],
    "dftbhip_ctx_set_kpoints": [C.c_void_p, C.c_int, _dp, _dp],
    "dftbhip_dsytrd": [C.c_char, C.c_int, _dp, C.c_int, _dp, _dp, _dp],
    "dftbhip_set_tuning": [C.c_int, C.c_int],
    "dftbhip_autotune": [C.c_int, C.c_int, _ip, C.c_int, _ip, _dp, C.c_int],
    "dftbhip_neighbour_list": [C.c_int, _dp, _dp, C.c_double, _ip, _ip, _dp, _ip, _ip, _ip, _dp],
    "dftbhip_dlacpy": [C.c_char, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int],
    "dftbhip_batch_create": [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p],
    "dftbhip_batch_destroy": [C.c_void_p],
    "dftbhip_batch_set_geometries": [C.c_void_p, C.c_int, _ip, _dp, _ip],
    "dftbhip_batch_scc": [C.c_void_p, _dp, _dp, _ip, _ip],
    "dftbhip_batch_gradients": [C.c_void_p, _dp],
    "dftbhip_batch_scc_begin": [C.c_void_p],
    "dftbhip_batch_scc_step": [C.c_void_p, _ip],
    "dftbhip_batch_scc_end": [C.c_void_p, _dp, _dp, _ip, _ip],
    "dftbhip_efilling": [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                         _dp],
    "dftbhip_mixer_create": [C.POINTER(C.c_void_p), C.c_int, C.c_double, C.c_double, C.c_double,
                             C.c_double, C.c_double],
    "dftbhip_mixer_reset": [C.c_void_p, C.c_int],
    "dftbhip_mixer_mix": [C.c_void_p, _dp, _dp],
    "dftbhip_mixer_destroy": [C.c_void_p],
}


def _setup(L):
    for name, args in EXPORTS.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int


def _check(rc: int, what: str):
    if rc != 0:
        buf = C.create_string_buffer(1024)
        lib().dftbhip_last_error(buf, 1024)
        raise DftbHipError(f"{what} failed ({rc}): {buf.value.decode(errors='replace')}")


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def make_params(params: Parameters):
    """Fill the C parameter struct; returns (struct, keep-alive list)."""
    p = Params()
    keep = []
    ns = params.nspecies
    if ns > MAXSP:
        raise DftbHipError("too many species")
    p.nspecies = ns
    for a, sp in enumerate(params.species):
        p.nshell[a] = len(sp.shells)
        for ish, l in enumerate(sp.shells):
            p.angshell[a][ish] = l
            p.onsite[a][ish] = sp.onsite[ish]
            p.occ[a][ish] = sp.occ[ish]
        p.hubbu[a] = sp.hubbu[0]
    p.nrows = params.nrows
    p.dist = params.dist
    for a in range(ns):
        for b in range(ns):
            k = a * ns + b
            h = _f64(params.ham[(a, b)])
            s = _f64(params.over[(a, b)])
            keep += [h, s]
            p.ninteg[k] = h.shape[1]
            p.htab[k] = _d(h)
            p.stab[k] = _d(s)
            rep = params.rep[(a, b)]
            xs = _f64(rep.xstart)
            cf = _f64(rep.spcoeffs.reshape(-1) if len(rep.spcoeffs) else np.zeros(4))
            last = _f64(rep.lastcoeffs)
            keep += [xs, cf, last]
            p.rep_nint[k] = len(xs)
            p.rep_cutoff[k] = rep.cutoff
            for c in range(3):
                p.rep_expc[k][c] = rep.expcoeffs[c]
            p.rep_xs[k] = _d(xs)
            p.rep_coef[k] = _d(cf)
            p.rep_last[k] = _d(last)
    return p, keep


def make_options(opts: Options, force_mode: int = 0, precision: int = 0,
                 refine_below: float = 1e-4, refine_steps: int = 2, tridiag: int = 1,
                 tail_below: float = 0.0, tail_steps: int = 1) -> COptions:
    o = COptions()
    o.kT = opts.kT
    o.scc_tol = opts.scc_tol
    o.max_scc_iter = opts.max_scc_iter
    o.min_scc_iter = opts.min_scc_iter
    o.mix_param = opts.mix_param
    o.omega0 = opts.omega0
    o.min_weight = opts.min_weight
    o.max_weight = opts.max_weight
    o.weight_fac = opts.weight_fac
    o.fd_delta = opts.fd_delta
    o.ewald_tol = opts.ewald_tol
    o.force_mode = force_mode
    o.precision = precision
    o.refine_below = refine_below
    o.refine_steps = refine_steps
    o.tridiag = tridiag
    o.tail_below = tail_below
    o.tail_steps = tail_steps
    return o


# ---------------------------------------------------------------------------------------------
# stateless drop-ins
# ---------------------------------------------------------------------------------------------

def sygvd(h: np.ndarray, s: np.ndarray, jobz: str = "V"):
    """Generalized eigenproblem (lower triangles of h and s used); returns (w, C, L)."""
    n = h.shape[0]
    a = np.asfortranarray(h, dtype=np.float64).copy(order="F")
    b = np.asfortranarray(s, dtype=np.float64).copy(order="F")
    w = np.zeros(n)
    rc = lib().dftbhip_sygvd(1, jobz.encode(), b"L", n, _d(a), n, _d(b), n, _d(w))
    if rc > 0:
        raise DftbHipError(f"sygvd info={rc}")
    _check(rc, "dftbhip_sygvd")
    return w, a, b


def hegvd(h: np.ndarray, s: np.ndarray, jobz: str = "V"):
    """Complex Hermitian generalized eigenproblem (lower triangles read): eigenvalues, the
    S-orthonormal eigenvectors (jobz 'V') and the Cholesky factor L (lower triangle)."""
    n = h.shape[0]
    a = np.asfortranarray(h, dtype=np.complex128).copy(order="F")
    b = np.asfortranarray(s, dtype=np.complex128).copy(order="F")
    w = np.zeros(n)
    rc = lib().dftbhip_hegvd(1, jobz.encode(), b"L", n, a.ctypes.data_as(_dp), n,
                             b.ctypes.data_as(_dp), n, _d(w))
    if rc > 0:
        raise DftbHipError(f"hegvd: info={rc}")
    _check(rc, "dftbhip_hegvd")
    return w, (a if jobz.upper() == "V" else None), np.tril(b)


def dsytrd(a: np.ndarray):
    """LAPACK dsytrd (lower) on the GPU: returns (reflectors in the lower part of a copy of a, d,
    e, tau)."""
    a = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    n = a.shape[0]
    d = np.zeros(n)
    e = np.zeros(max(n - 1, 1))
    tau = np.zeros(max(n - 1, 1))
    _check(lib().dftbhip_dsytrd(b"L", n, _d(a), n, _d(d), _d(e), _d(tau)), "dftbhip_dsytrd")
    return a, d, e[: max(n - 1, 0)], tau[: max(n - 1, 0)]


def dlacpy(a: np.ndarray) -> np.ndarray:
    """b = a through the device copy kernel."""
    a = np.asfortranarray(a, dtype=np.float64)
    b = np.zeros_like(a, order="F")
    m, n = a.shape
    _check(lib().dftbhip_dlacpy(b"A", m, n, _d(a), m, _d(b), m), "dftbhip_dlacpy")
    return b


def set_tuning(param: int, value: int) -> None:
    _check(lib().dftbhip_set_tuning(param, value), "dftbhip_set_tuning")


def neighbour_list(coords: np.ndarray, latvecs: Optional[np.ndarray], cutoff: float):
    """Half neighbour list built on the device (updateNeighbourList): returns (coords_all
    (nAll, 3), img2cent (nAll,), neigh, dist2) with 0-based indices; neigh[a] and dist2[a] start
    with the atom itself."""
    x = np.ascontiguousarray(coords, dtype=np.float64)
    nat = x.shape[0]
    lat = None if latvecs is None else np.ascontiguousarray(latvecs, dtype=np.float64)
    nall = np.zeros(1, dtype=np.int32)
    mnb = np.zeros(1, dtype=np.int32)
    latp = None if lat is None else _d(lat)
    _check(lib().dftbhip_neighbour_list(nat, _d(x), latp, cutoff, _i(nall), _i(mnb), None, None,
                                        None, None, None), "dftbhip_neighbour_list")
    na, mn = int(nall[0]), int(mnb[0])
    call = np.zeros((na, 3))
    i2c = np.zeros(na, dtype=np.int32)
    inb = np.zeros((nat, mn + 1), dtype=np.int32)    # row a = column a of iNeighbour
    nnb = np.zeros(nat, dtype=np.int32)
    d2 = np.zeros((nat, mn + 1))
    _check(lib().dftbhip_neighbour_list(nat, _d(x), latp, cutoff, _i(nall), _i(mnb), _d(call),
                                        _i(i2c), _i(inb), _i(nnb), _d(d2)),
           "dftbhip_neighbour_list")
    neigh = [inb[a, : nnb[a] + 1].astype(np.int64) - 1 for a in range(nat)]
    dist2 = [d2[a, : nnb[a] + 1].copy() for a in range(nat)]
    return call, i2c.astype(np.int64) - 1, neigh, dist2


def autotune(n: int, precision: int = 0):
    """Autotunes the tridiagonalisation's SYMV tile count on an n x n synthetic matrix for FP64
    (precision 0) or FP32 (1) and installs the table. Returns (rows, seconds): rows as
    [(first trailing size, tiles), ...] ascending, seconds[o, b] the summed SYMV time of option
    (2, 4, 8, 16)[o] in the 256-row bucket b of trailing sizes. A bucket keeps the option of the
    bucket below unless another is at least 2 % faster there."""
    nb = (n + 255) // 256
    table = np.zeros(2 * 256, dtype=np.int32)
    nrows = np.zeros(1, dtype=np.int32)
    detail = np.zeros(4 * nb)
    _check(lib().dftbhip_autotune(n, precision, _i(table), 256, _i(nrows), _d(detail), detail.size),
           "dftbhip_autotune")
    k = min(int(nrows[0]), 256)
    rows = [(int(table[2 * r]), int(table[2 * r + 1])) for r in range(k)]
    return rows, detail.reshape(4, nb)


def dgemm(a: np.ndarray, b: np.ndarray, transa: str = "N", transb: str = "N", alpha: float = 1.0,
          beta: float = 0.0, c: Optional[np.ndarray] = None) -> np.ndarray:
    A = np.asfortranarray(a, dtype=np.float64)
    B = np.asfortranarray(b, dtype=np.float64)
    m = A.shape[1] if transa == "T" else A.shape[0]
    k = A.shape[0] if transa == "T" else A.shape[1]
    n = B.shape[0] if transb == "T" else B.shape[1]
    Cm = np.zeros((m, n), order="F") if c is None else np.asfortranarray(c, dtype=np.float64).copy(order="F")
    _check(lib().dftbhip_dgemm(transa.encode(), transb.encode(), m, n, k, alpha, _d(A),
                               max(1, A.shape[0]), _d(B), max(1, B.shape[0]), beta, _d(Cm), max(1, m)),
           "dftbhip_dgemm")
    return Cm


def dstedc(d: np.ndarray, e: np.ndarray):
    n = len(d)
    dd = _f64(d).copy()
    ee = _f64(e).copy() if n > 1 else np.zeros(1)
    z = np.zeros((n, n), order="F")
    rc = lib().dftbhip_dstedc(n, _d(dd), _d(ee), _d(z), n)
    if rc > 0:
        raise DftbHipError(f"dstedc info={rc}")
    _check(rc, "dftbhip_dstedc")
    return dd, z


def efilling(eigenvals: np.ndarray, kweights: np.ndarray, n_electrons: float, kT: float):
    """Fermi filling of one spin channel (``Efilling``, etemp.F90:59-194) on the host:
    eigenvals (nKPoints, nLevels), ascending per k-point. Returns (filling, Ef, Ebs, TS, E0)."""
    e = np.ascontiguousarray(np.atleast_2d(eigenvals), dtype=np.float64)
    nk, nlev = e.shape
    w = _f64(kweights)
    f = np.zeros_like(e)
    out = np.zeros(4)
    r = [C.c_double() for _ in range(4)]
    _check(lib().dftbhip_efilling(nlev, nk, _d(e), _d(w), float(n_electrons), float(kT), _d(f),
                                  *[C.byref(x) for x in r]), "efilling")
    out[:] = [x.value for x in r]
    return f, out[0], out[1], out[2], out[3]


class BroydenMixer:
    """The library's modified Broyden mixer (``TBroydenMixerReal``, broydenmixer.F90:112-342)."""

    def __init__(self, n: int, mix_param: float = 0.2, omega0: float = 0.01,
                 min_weight: float = 1.0, max_weight: float = 1.0e5, weight_fac: float = 1.0e-2,
                 max_scc_iter: int = 100):
        self._h = C.c_void_p()
        _check(lib().dftbhip_mixer_create(C.byref(self._h), max_scc_iter, mix_param, omega0,
                                          min_weight, max_weight, weight_fac), "mixer_create")
        self.n = n
        _check(lib().dftbhip_mixer_reset(self._h, n), "mixer_reset")

    def mix(self, qin: np.ndarray, qdiff: np.ndarray) -> np.ndarray:
        q = np.array(qin, dtype=np.float64)
        d = _f64(qdiff)
        _check(lib().dftbhip_mixer_mix(self._h, _d(q), _d(d)), "mixer_mix")
        return q

    def close(self):
        if self._h:
            lib().dftbhip_mixer_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dm_real(evec: np.ndarray, filling: np.ndarray, nlevels: Optional[int] = None) -> np.ndarray:
    n = evec.shape[0]
    nl = evec.shape[1] if nlevels is None else nlevels
    ev = np.asfortranarray(evec, dtype=np.float64)
    f = _f64(filling)
    dm = np.zeros((n, n), order="F")
    _check(lib().dftbhip_dm_real(n, nl, _d(ev), n, _d(f), _d(dm), n), "dftbhip_dm_real")
    return dm


def edm_real(evec, filling, eigen, nlevels=None) -> np.ndarray:
    n = evec.shape[0]
    nl = evec.shape[1] if nlevels is None else nlevels
    ev = np.asfortranarray(evec, dtype=np.float64)
    f = _f64(filling)
    e = _f64(eigen)
    dm = np.zeros((n, n), order="F")
    _check(lib().dftbhip_edm_real(n, nl, _d(ev), n, _d(f), _d(e), _d(dm), n), "dftbhip_edm_real")
    return dm


def dm_cmplx(evec: np.ndarray, filling: np.ndarray, nlevels: Optional[int] = None,
             eigen: Optional[np.ndarray] = None) -> np.ndarray:
    """Lower triangle of C diag(f) C^H (with eigen: C diag(f eigen) C^H) for complex C."""
    n = evec.shape[0]
    nl = evec.shape[1] if nlevels is None else nlevels
    ev = np.asfortranarray(evec, dtype=np.complex128)
    f = _f64(filling)
    dm = np.zeros((n, n), dtype=np.complex128, order="F")
    evp = ev.ctypes.data_as(_dp)
    dmp = dm.ctypes.data_as(_dp)
    if eigen is None:
        _check(lib().dftbhip_dm_cmplx(n, nl, evp, n, _d(f), dmp, n), "dftbhip_dm_cmplx")
    else:
        e = _f64(eigen)
        _check(lib().dftbhip_edm_cmplx(n, nl, evp, n, _d(f), _d(e), dmp, n), "dftbhip_edm_cmplx")
    return dm


class SkHandle:
    def __init__(self, params: Parameters):
        self._p, self._keep = make_params(params)
        self.h = C.c_void_p()
        _check(lib().dftbhip_sk_create(C.byref(self.h), C.byref(self._p)), "dftbhip_sk_create")

    def close(self):
        if self.h:
            lib().dftbhip_sk_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RefArrays:
    """Reference-convention neighbour arrays (1-based, Fortran layout) for the sparse drop-ins."""

    def __init__(self, natom, nall, coords_all, species_all, img2cent, ineigh, nneigh_sk, ipair,
                 iatomstart, nnz):
        self.natom, self.nall, self.nnz = natom, nall, nnz
        self.coords = _f64(coords_all.reshape(-1))
        self.species = _i32(species_all)
        self.img2cent = _i32(img2cent)
        self.mneigh = ineigh.shape[0] - 1
        self.ineigh = np.asfortranarray(ineigh, dtype=np.int32)
        self.nneigh_sk = _i32(nneigh_sk)
        self.ipair = np.asfortranarray(ipair, dtype=np.int32)
        self.iatomstart = _i32(iatomstart)


def build_h0_s(sk: SkHandle, ra: RefArrays):
    h0 = np.zeros(ra.nnz)
    s = np.zeros(ra.nnz)
    _check(lib().dftbhip_build_h0_s(sk.h, ra.natom, ra.nall, _d(ra.coords), _i(ra.species),
                                    _i(ra.img2cent), ra.mneigh, _i(ra.ineigh), _i(ra.nneigh_sk),
                                    _i(ra.ipair), ra.nnz, _d(h0), _d(s)), "dftbhip_build_h0_s")
    return h0, s


def unpack_hs(ra: RefArrays, x: np.ndarray) -> np.ndarray:
    n = int(ra.iatomstart[-1]) - 1
    sq = np.zeros((n, n), order="F")
    _check(lib().dftbhip_unpack_hs_real(ra.natom, _i(ra.iatomstart), _i(ra.img2cent), ra.mneigh,
                                        _i(ra.ineigh), _i(ra.nneigh_sk), _i(ra.ipair), ra.nnz,
                                        _d(_f64(x)), _d(sq), n), "dftbhip_unpack_hs_real")
    return sq


def pack_hs(ra: RefArrays, dense: np.ndarray) -> np.ndarray:
    n = int(ra.iatomstart[-1]) - 1
    sq = np.asfortranarray(dense, dtype=np.float64)
    out = np.zeros(ra.nnz)
    _check(lib().dftbhip_pack_hs_real(ra.natom, _i(ra.iatomstart), _i(ra.img2cent), ra.mneigh,
                                      _i(ra.ineigh), _i(ra.nneigh_sk), _i(ra.ipair), ra.nnz,
                                      _d(sq), n, _d(out)), "dftbhip_pack_hs_real")
    return out


def mulliken(ra: RefArrays, over: np.ndarray, rho: np.ndarray) -> np.ndarray:
    n = int(ra.iatomstart[-1]) - 1
    q = np.zeros(n)
    _check(lib().dftbhip_mulliken(ra.natom, _i(ra.iatomstart), _i(ra.img2cent), ra.mneigh,
                                  _i(ra.ineigh), _i(ra.nneigh_sk), _i(ra.ipair), ra.nnz,
                                  _d(_f64(over)), _d(_f64(rho)), _d(q)), "dftbhip_mulliken")
    return q


# ---------------------------------------------------------------------------------------------
# calculator mirroring tools/pythonapi DftbPlus
# ---------------------------------------------------------------------------------------------
ENERGY_KEYS = ("mermin", "total", "nonscc", "scc", "repulsive", "ts", "eband", "ef")
TIMER_KEYS = ("setup", "hamiltonian", "reduce", "tridiag", "divconq", "backtransform", "density",
              "pack_mulliken", "host", "forces", "last_iteration", "refine")


class DftbPlus:
    """SCC-DFTB calculator on the GPU (energies in Ha, lengths in bohr).

    ``get_energy`` returns the Mermin free energy like ``dftbp_get_energy``
    (api/dftbplus.h:550), ``get_gradients`` dE/dR (dftbplus.h:562) and ``get_gross_charges``
    the net atomic charges q0 - q (dftbplus.h:639).
    """

    def __init__(self, params: Parameters, geometry: Geometry, options: Options = Options(),
                 force_mode: int = 0, precision: int = 0, refine_below: float = 1e-4,
                 refine_steps: int = 2, tridiag: int = 1, tail_below: float = 0.0,
                 tail_steps: int = 1):
        self.params = params
        self._p, self._keep = make_params(params)
        self._o = make_options(options, force_mode, precision, refine_below, refine_steps,
                               tridiag, tail_below, tail_steps)
        self.h = C.c_void_p()
        _check(lib().dftbhip_ctx_create(C.byref(self.h), C.byref(self._p), C.byref(self._o)),
               "dftbhip_ctx_create")
        self.species = _i32(geometry.species)
        self.natom = len(self.species)
        self._dirty = True
        kp = getattr(options, "kpoints", None)
        if kp is not None:
            k = _f64(np.asarray(kp, dtype=np.float64).reshape(-1))
            kw = getattr(options, "kweights", None)
            w = None if kw is None else _f64(np.asarray(kw, dtype=np.float64))
            _check(lib().dftbhip_ctx_set_kpoints(self.h, len(k) // 3, _d(k),
                                                 None if w is None else _d(w)),
                   "dftbhip_ctx_set_kpoints")
        self.set_geometry(geometry.coords, geometry.latvecs)

    def set_geometry(self, coords, latvecs=None):
        c = _f64(np.asarray(coords).reshape(-1))
        if len(c) != 3 * self.natom:
            raise DftbHipError("coordinate array has the wrong size")
        lv = None if latvecs is None else _f64(np.asarray(latvecs).reshape(-1))
        _check(lib().dftbhip_ctx_set_geometry(self.h, self.natom, _d(c), _i(self.species),
                                              None if lv is None else _d(lv)),
               "dftbhip_ctx_set_geometry")
        self._dirty = True

    def _ensure_scc(self):
        if not self._dirty:
            return
        e = np.zeros(8)
        q = np.zeros(self.natom)
        nit = C.c_int()
        conv = C.c_int()
        _check(lib().dftbhip_ctx_scc(self.h, _d(e), _d(q), C.byref(nit), C.byref(conv)),
               "dftbhip_ctx_scc")
        self.energies = dict(zip(ENERGY_KEYS, e))
        self.charges = q
        self.n_iter = nit.value
        self.converged = bool(conv.value)
        self._grad = None
        self._dirty = False

    def set_charge_restart(self, warm: bool):
        """Start every later SCC cycle from the input charges of the previous one (as DFTB+
        keeps qInput between geometry steps) instead of the neutral atoms."""
        _check(lib().dftbhip_ctx_set_charge_restart(self.h, int(bool(warm))),
               "dftbhip_ctx_set_charge_restart")

    # one-iteration control of the SCC cycle (used to time single iterations)
    def set_input_charges(self, populations: np.ndarray, check_count: bool = True):
        """Electron population per atom for the next SCC cycle (ReadInitialCharges): their
        sum must equal the number of electrons to 1e-3 unless check_count is False (the
        reference's SkipChargeTest)."""
        q = np.ascontiguousarray(populations, dtype=np.float64)
        _check(lib().dftbhip_ctx_set_input_charges(self.h, len(q), _d(q), int(check_count)),
               "dftbhip_ctx_set_input_charges")

    def get_input_charges(self) -> np.ndarray:
        """Electron population per atom that the next SCC iteration takes as input (the mixed
        charges after the last step; what the reference writes to its charge file)."""
        q = np.zeros(self.natom)
        _check(lib().dftbhip_ctx_get_input_charges(self.h, len(q), _d(q)),
               "dftbhip_ctx_get_input_charges")
        return q

    def scc_begin(self):
        _check(lib().dftbhip_ctx_scc_begin(self.h), "dftbhip_ctx_scc_begin")

    def scc_step(self):
        conv = C.c_int()
        err = C.c_double()
        _check(lib().dftbhip_ctx_scc_step(self.h, C.byref(conv), C.byref(err)), "dftbhip_ctx_scc_step")
        return bool(conv.value), err.value

    def scc_end(self):
        e = np.zeros(8)
        q = np.zeros(self.natom)
        nit = C.c_int()
        conv = C.c_int()
        _check(lib().dftbhip_ctx_scc_end(self.h, _d(e), _d(q), C.byref(nit), C.byref(conv)),
               "dftbhip_ctx_scc_end")
        self.energies = dict(zip(ENERGY_KEYS, e))
        self.charges = q
        self.n_iter = nit.value
        self.converged = bool(conv.value)
        self._grad = None
        self._dirty = False

    def scc_state(self) -> dict:
        """Iteration number, charge error, electronic energy of the last iteration and its
        change from the one before, input electrons of the last iteration, pairs outside the
        first-order range of the last refinement step and, for the residual form, of those
        beyond the diagonal band of X^T S X in the last refinement; whether the tail was
        switched off for the rest of the cycle."""
        out = np.zeros(8)
        _check(lib().dftbhip_ctx_scc_state(self.h, _d(out), 8), "dftbhip_ctx_scc_state")
        return {"iteration": int(out[0]), "scc_error": float(out[1]), "e_elec": float(out[2]),
                "diff_elec": float(out[2] - out[3]) if np.isfinite(out[3]) else float(out[2]),
                "q_input_sum": float(out[4]), "oa_clipped_pairs": int(out[5]),
                "oa_far_pairs": int(out[6]), "tail_off": bool(out[7])}

    def get_nr_atoms(self) -> int:
        return self.natom

    def get_nr_orbitals(self) -> int:
        return lib().dftbhip_ctx_nr_orbitals(self.h)

    def get_energy(self) -> float:
        self._ensure_scc()
        return float(self.energies["mermin"])

    def get_gradients(self) -> np.ndarray:
        self._ensure_scc()
        if self._grad is None:
            g = np.zeros(3 * self.natom)
            _check(lib().dftbhip_ctx_gradients(self.h, _d(g)), "dftbhip_ctx_gradients")
            self._grad = g.reshape(self.natom, 3)
        return self._grad.copy()

    def get_gross_charges(self) -> np.ndarray:
        self._ensure_scc()
        return self.charges.copy()

    def get_eigenvalues(self):
        self._ensure_scc()
        n = self.get_nr_orbitals()
        e = np.zeros(n)
        f = np.zeros(n)
        _check(lib().dftbhip_ctx_eigenvalues(self.h, _d(e), _d(f), n), "dftbhip_ctx_eigenvalues")
        return e, f

    def get_eigenvalues_k(self, k: int):
        """Eigenvalues and fillings (electrons per level) of k-point k (0 without k-points)."""
        self._ensure_scc()
        n = self.get_nr_orbitals()
        e = np.zeros(n)
        f = np.zeros(n)
        _check(lib().dftbhip_ctx_eigenvalues_k(self.h, k, _d(e), _d(f), n),
               "dftbhip_ctx_eigenvalues_k")
        return e, f

    def get_orbital_populations(self) -> np.ndarray:
        """Output Mulliken populations per orbital (electrons), atom by atom."""
        self._ensure_scc()
        n = self.get_nr_orbitals()
        q = np.zeros(n)
        _check(lib().dftbhip_ctx_orbital_populations(self.h, _d(q), n),
               "dftbhip_ctx_orbital_populations")
        return q

    def timings(self) -> dict:
        t = np.zeros(len(TIMER_KEYS))
        lib().dftbhip_ctx_timings(self.h, _d(t), len(t))
        return dict(zip(TIMER_KEYS, t))

    def close(self):
        if self.h:
            lib().dftbhip_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MoleculeBatch:
    """Many independent non-periodic molecules solved together (``dftbhip_batch_*``): the
    result for each molecule is that of a separate single-molecule calculation."""

    def __init__(self, params: Parameters, geometries, options: Options = Options(),
                 force_mode: int = 0):
        self._p, self._keep = make_params(params)
        self._o = make_options(options, force_mode, 0)
        self.h = C.c_void_p()
        _check(lib().dftbhip_batch_create(C.byref(self.h), C.byref(self._p), C.byref(self._o)),
               "dftbhip_batch_create")
        self.nat = np.array([g.natom for g in geometries], dtype=np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.nat)])
        coords = _f64(np.concatenate([g.coords.reshape(-1, 3) for g in geometries]).reshape(-1))
        species = _i32(np.concatenate([g.species for g in geometries]))
        self.nmol = len(geometries)
        _check(lib().dftbhip_batch_set_geometries(self.h, self.nmol, _i(self.nat), _d(coords),
                                                  _i(species)), "dftbhip_batch_set_geometries")
        self._done = False

    def scc_begin(self):
        _check(lib().dftbhip_batch_scc_begin(self.h), "dftbhip_batch_scc_begin")

    def scc_step(self) -> int:
        """One SCC iteration of every molecule not yet converged; returns how many remain."""
        na = C.c_int()
        _check(lib().dftbhip_batch_scc_step(self.h, C.byref(na)), "dftbhip_batch_scc_step")
        return na.value

    def scc_end(self):
        return self._finish("dftbhip_batch_scc_end")

    def scc(self):
        return self._finish("dftbhip_batch_scc")

    def _finish(self, fn):
        e = np.zeros((self.nmol, 8))
        q = np.zeros(int(self.offsets[-1]))
        nit = np.zeros(self.nmol, dtype=np.int32)
        conv = np.zeros(self.nmol, dtype=np.int32)
        _check(getattr(lib(), fn)(self.h, _d(e), _d(q), _i(nit), _i(conv)), fn)
        self.energies = {k: e[:, i].copy() for i, k in enumerate(ENERGY_KEYS)}
        self.charges = [q[self.offsets[m]:self.offsets[m + 1]] for m in range(self.nmol)]
        self.n_iter = nit
        self.converged = conv.astype(bool)
        self._done = True
        return self.energies["mermin"]

    def gradients(self):
        if not self._done:
            self.scc()
        g = np.zeros(3 * int(self.offsets[-1]))
        _check(lib().dftbhip_batch_gradients(self.h, _d(g)), "dftbhip_batch_gradients")
        g = g.reshape(-1, 3)
        return [g[self.offsets[m]:self.offsets[m + 1]] for m in range(self.nmol)]

    def close(self):
        if self.h:
            lib().dftbhip_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
