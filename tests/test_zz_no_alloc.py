"""The allocation-free tridiagonalisation (dftbhip_dsytrd_noalloc) against dftbhip_dsytrd and
LAPACK: device arrays allocated by the test through the HIP runtime the library itself loaded,
two calls in a row, the second with a smaller matrix (the fixed workspace keeps the first call's
contents), and the size limit of the fixed workspace reported as an argument error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime already mapped into this process by the library (same instance)."""
    from dftbplus_amd import api
    api.lib()
    path = None
    with open("/proc/self/maps") as fh:
        for line in fh:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libamdhip64 not mapped"
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevArray:
    def __init__(self, hip, host: np.ndarray):
        self.hip = hip
        self.nbytes = host.nbytes
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(self.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.p, host.ctypes.data, self.nbytes, H2D) == 0

    def get(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, self.nbytes, D2H) == 0
        return out

    def ptr(self):
        return C.cast(self.p, C.POINTER(C.c_double))

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(hip, a):
    from dftbplus_amd import api
    L = api.lib()
    n = a.shape[0]
    af = np.asfortranarray(a)
    d0, e0 = np.zeros(n), np.zeros(max(n - 1, 1))
    bufs = [DevArray(hip, af), DevArray(hip, d0), DevArray(hip, e0), DevArray(hip, e0.copy())]
    try:
        L.dftbhip_dsytrd_noalloc.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int,
                                            C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.POINTER(C.c_double)]
        rc = L.dftbhip_dsytrd_noalloc(n, bufs[0].ptr(), n, bufs[1].ptr(), bufs[2].ptr(),
                                      bufs[3].ptr())
        api._check(rc, "dftbhip_dsytrd_noalloc")
        r = bufs[0].get(af)
        d = bufs[1].get(d0)
        e = bufs[2].get(e0)[: n - 1]
        tau = bufs[3].get(e0)[: n - 1]
        return r, d, e, tau
    finally:
        for b in bufs:
            b.free()


def test_dsytrd_noalloc_matches_dsytrd():
    from dftbplus_amd import api
    api._check(api.lib().dftbhip_init(0), "dftbhip_init")
    hip = _hip()
    for n in (1500, 700):   # the second call smaller than the first
        rng = np.random.default_rng(n)
        a = rng.standard_normal((n, n))
        a = a + a.T
        r0, d0, e0, t0 = api.dsytrd(a)
        r1, d1, e1, t1 = _noalloc(hip, a)
        tol = 1e-12 * n * np.abs(a).max()
        print(f"n={n}: |dd| {np.max(np.abs(d1 - d0)):.2e} |de| {np.max(np.abs(e1 - e0)):.2e} "
              f"|dtau| {np.max(np.abs(t1 - t0)):.2e}")
        assert np.max(np.abs(d1 - d0)) < tol
        assert np.max(np.abs(e1 - e0)) < tol
        assert np.max(np.abs(t1 - t0)) < tol
        assert np.max(np.abs(np.tril(r1, -2) - np.tril(r0, -2))) < tol
        # the tridiagonal has the spectrum of a
        t = np.diag(d1) + np.diag(e1, 1) + np.diag(e1, -1)
        assert np.max(np.abs(np.linalg.eigvalsh(t) - np.linalg.eigvalsh(a))) < tol


def test_dsytrd_noalloc_rejects_sizes_beyond_workspace():
    from dftbplus_amd import api
    L = api.lib()
    api._check(L.dftbhip_init(0), "dftbhip_init")
    L.dftbhip_dsytrd_noalloc.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
    dummy = C.cast(C.c_void_p(256), C.POINTER(C.c_double))   # never dereferenced: rejected first
    rc = L.dftbhip_dsytrd_noalloc(40000, dummy, 40000, dummy, dummy, dummy)
    assert rc == -1
    buf = C.create_string_buffer(512)
    L.dftbhip_last_error(buf, 512)
    assert b"workspace" in buf.value
