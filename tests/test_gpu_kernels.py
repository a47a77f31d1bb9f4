"""GPU parity of the individual kernels against numpy/scipy and the CPU oracle, through the C ABI."""
import numpy as np
import pytest
import scipy.linalg

from conftest import reference_arrays

pytestmark = pytest.mark.gpu


def _api():
    from dftbplus_amd import api
    api.lib()
    return api


def test_mfma_layouts():
    """Accumulator element r of lane l of the 16x16x4 MFMA holds D(i, l%16) with
    i = l//16 + 4r for FP64 and i = 4(l//16) + r for FP32 (D(i, j) = i + 100 j in the probe)."""
    api = _api()
    out = np.zeros(512)
    api._check(api.lib().dftbhip_probe_mfma_f64(api._d(out)), "probe")
    f64 = out[:256].reshape(64, 4)
    f32 = out[256:].reshape(64, 4)
    e64 = np.array([[l // 16 + 4 * r + 100 * (l % 16) for r in range(4)] for l in range(64)], float)
    e32 = np.array([[4 * (l // 16) + r + 100 * (l % 16) for r in range(4)] for l in range(64)], float)
    assert np.array_equal(f64, e64)
    assert np.array_equal(f32, e32)


def test_mfma_f16_layout():
    """Accumulator element r of lane l of the FP16 16x16x32 MFMA holds D(4(l//16) + r, l%16),
    as for the FP32 16x16x4 one (D(i, j) = i + 100 j in the probe)."""
    api = _api()
    out = np.zeros(256)
    api._check(api.lib().dftbhip_probe_mfma_f16(api._d(out)), "probe")
    e = np.array([[4 * (l // 16) + r + 100 * (l % 16) for r in range(4)] for l in range(64)], float)
    assert np.array_equal(out.reshape(64, 4), e)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 53, 29), (129, 128, 64), (300, 257, 999),
                                   (128, 200, 0)])
def test_gemm_f16_tn(m, n, k):
    """c = alpha a^T b with FP16 a (k x m), b (k x n) and FP32 sums: against the float64 product
    of the same FP16 values, to the FP32 rounding of the sums (relative 1e-5 of |a|^T |b|)."""
    api = _api()
    rng = np.random.default_rng(m * n + k)
    a = np.asfortranarray(rng.standard_normal((k, m)).astype(np.float16))
    b = np.asfortranarray(rng.standard_normal((k, n)).astype(np.float16))
    c = np.zeros((m, n), dtype=np.float32, order="F")
    api._check(api.lib().dftbhip_gemm_f16_tn(m, n, k, 0.75, a.ctypes.data, max(k, 1),
                                              b.ctypes.data, max(k, 1), c.ctypes.data, m),
               "gemm_f16_tn")
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    ref = 0.75 * a64.T @ b64
    bound = 1e-5 * (np.abs(a64).T @ np.abs(b64)) + 1e-30
    assert np.all(np.abs(c - ref) <= bound)


@pytest.mark.parametrize("m,n,k,ta,tb", [(1, 1, 1, "N", "N"), (37, 53, 29, "N", "N"),
                                         (130, 257, 200, "T", "N"), (128, 128, 16, "N", "T"),
                                         (300, 129, 77, "T", "T")])
def test_dgemm(m, n, k, ta, tb):
    api = _api()
    rng = np.random.default_rng(m + n + k)
    a = rng.standard_normal((k, m) if ta == "T" else (m, k))
    b = rng.standard_normal((n, k) if tb == "T" else (k, n))
    c0 = rng.standard_normal((m, n))
    ref = 0.7 * ((a.T if ta == "T" else a) @ (b.T if tb == "T" else b)) - 0.3 * c0
    got = api.dgemm(a, b, ta, tb, alpha=0.7, beta=-0.3, c=c0)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def _tridiag_check(d, e, w, z):
    n = len(d)
    ref = scipy.linalg.eigh_tridiagonal(d, e, eigvals_only=True) if n > 1 else d
    scale = max(1.0, np.abs(ref).max())
    assert np.max(np.abs(w - ref)) < 1e-12 * scale * max(1, np.sqrt(n))
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.max(np.abs(T @ z - z * w)) < 1e-11 * scale * max(1, np.sqrt(n))
    assert np.max(np.abs(z.T @ z - np.eye(n))) < 1e-11 * max(1, np.sqrt(n))


@pytest.mark.parametrize("n", [1, 2, 5, 64, 65, 129, 300, 1000])
def test_dstedc_random(n):
    api = _api()
    rng = np.random.default_rng(n)
    d = rng.uniform(-1, 1, n)
    e = rng.uniform(-1, 1, max(n - 1, 0))
    w, z = api.dstedc(d, e)
    _tridiag_check(d, e, w, z)


def test_dstedc_clustered():
    """Glued Wilkinson matrices: many pairs of eigenvalues equal to working precision (heavy
    deflation) plus exact degeneracies from a decoupled block."""
    api = _api()
    m = 21
    dw = np.abs(np.arange(m) - m // 2).astype(float)
    d = np.concatenate([dw] * 10)
    e = np.concatenate([np.concatenate([np.ones(m - 1), [1e-8]])] * 10)[:-1]
    e[m * 5 - 1] = 0.0
    w, z = api.dstedc(d, e)
    _tridiag_check(d, e, w, z)


@pytest.mark.parametrize("n", [1, 2, 3, 10, 100, 333, 1100])
def test_sygvd_random(n):
    api = _api()
    rng = np.random.default_rng(7 * n)
    x = rng.standard_normal((n, n))
    s = x @ x.T / n + np.eye(n)
    h = rng.standard_normal((n, n))
    h = h + h.T
    w, c, l = api.sygvd(h, s)
    ref = scipy.linalg.eigh(h, s, eigvals_only=True)
    assert np.max(np.abs(w - ref)) < 1e-11 * np.abs(ref).max()
    assert np.max(np.abs(h @ c - s @ c * w)) < 1e-10 * np.abs(ref).max()
    assert np.max(np.abs(c.T @ s @ c - np.eye(n))) < 1e-11
    assert np.allclose(np.tril(l), np.linalg.cholesky(s), atol=1e-12)


def test_sygvd_single_workgroup_bulge_chasing():
    """The one-workgroup launch of the bulge chasing (the path taken when the device refuses
    the cooperative launch) runs the same sweep steps on the same data: the same bits."""
    api = _api()
    n = 333
    rng = np.random.default_rng(5)
    x = rng.standard_normal((n, n))
    s = x @ x.T / n + np.eye(n)
    h = rng.standard_normal((n, n))
    h = h + h.T
    w0, c0, _ = api.sygvd(h, s)
    api.set_tuning(13, 1)
    try:
        w1, c1, _ = api.sygvd(h, s)
    finally:
        api.set_tuning(13, 0)
    assert np.array_equal(w0, w1) and np.array_equal(c0, c1)


def test_sygvd_not_positive_definite():
    api = _api()
    s = np.eye(4)
    s[2, 2] = -1.0
    with pytest.raises(api.DftbHipError, match="info=7"):
        api.sygvd(np.eye(4), s)


def _random_hpd(rng, n):
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return x @ x.conj().T / n + np.eye(n)


def _check_hegvd(api, h, s, tol=1e-11):
    """Eigenvalues against scipy (LAPACK zhegvd), residual, S-orthonormality, the factor."""
    n = h.shape[0]
    w, c, l = api.hegvd(np.tril(h), np.tril(s))
    ref = scipy.linalg.eigh(h, s, eigvals_only=True)
    scale = max(1.0, np.abs(ref).max())
    assert np.max(np.abs(w - ref)) < tol * scale
    assert np.max(np.abs(h @ c - s @ c * w)) < 10 * tol * scale
    assert np.max(np.abs(c.conj().T @ s @ c - np.eye(n))) < tol
    assert np.max(np.abs(l @ l.conj().T - s)) < 1e-12 * np.abs(s).max()
    assert np.all(np.diag(l).real > 0) and np.max(np.abs(np.diag(l).imag)) == 0.0
    return w, c


@pytest.mark.parametrize("n", [1, 2, 65, 200, 257])
def test_hegvd_random(n):
    """magmaHegvd (dcomplex) twin: random Hermitian H, Hermitian positive definite S."""
    api = _api()
    rng = np.random.default_rng(11 * n)
    s = _random_hpd(rng, n)
    h = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    h = h + h.conj().T
    _check_hegvd(api, h, s)


def test_hegvd_degenerate_levels():
    """Levels of multiplicity 1, 2 and 3 (H = X^-H D X^-1, S = X^-H X^-1 for a random complex X,
    so H X = S X D): every degenerate cluster still gives S-orthonormal eigenvectors that span
    it."""
    api = _api()
    rng = np.random.default_rng(5)
    n = 96
    d = np.sort(np.concatenate([rng.uniform(-1, 1, n - 12), np.repeat([0.25, -0.4], 3),
                                np.repeat([0.6, 0.7, -0.1], 2)]))
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * np.eye(n)
    xi = np.linalg.inv(x)
    h = xi.conj().T @ np.diag(d) @ xi
    h = 0.5 * (h + h.conj().T)
    s = xi.conj().T @ xi
    s = 0.5 * (s + s.conj().T)
    w, c = _check_hegvd(api, h, s, tol=1e-10)
    np.testing.assert_allclose(w, d, atol=1e-10)


def test_hegvd_not_positive_definite():
    api = _api()
    s = np.eye(4, dtype=complex)
    s[2, 2] = -1.0
    with pytest.raises(api.DftbHipError, match="info=7"):
        api.hegvd(np.eye(4, dtype=complex), s)


def test_density_matrices():
    api = _api()
    rng = np.random.default_rng(3)
    n, nl = 200, 120
    c = rng.standard_normal((n, n))
    f = np.sort(rng.uniform(0, 2, n))[::-1]
    eig = rng.uniform(-1, 1, n)
    dm = api.dm_real(c, f, nl)
    edm = api.edm_real(c, f, eig, nl)
    ref = (c[:, :nl] * f[:nl]) @ c[:, :nl].T
    refe = (c[:, :nl] * (f[:nl] * eig[:nl])) @ c[:, :nl].T
    lo = np.tril_indices(n)
    assert np.allclose(dm[lo], ref[lo], rtol=1e-12, atol=1e-11)
    assert np.allclose(edm[lo], refe[lo], rtol=1e-12, atol=1e-11)


@pytest.mark.parametrize("n,nl", [(200, 120), (257, 257), (65, 1)])
def test_complex_density_matrices(n, nl):
    """makeDensityMtxCmplxGPU / makeEnergyDensityMtxCmplxGPU twins: lower triangle of
    C diag(f) C^H and C diag(f eps) C^H (negative weights included) for complex C, zeros above
    the diagonal; against numpy in complex128."""
    api = _api()
    rng = np.random.default_rng(7)
    c = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    f = np.sort(rng.uniform(0, 2, n))[::-1]
    eig = rng.uniform(-1, 1, n)
    dm = api.dm_cmplx(c, f, nl)
    edm = api.dm_cmplx(c, f, nl, eigen=eig)
    ref = (c[:, :nl] * f[:nl]) @ c[:, :nl].conj().T
    refe = (c[:, :nl] * (f[:nl] * eig[:nl])) @ c[:, :nl].conj().T
    lo = np.tril_indices(n)
    up = np.triu_indices(n, 1)
    scale = max(1.0, np.abs(ref).max())
    assert np.max(np.abs(dm[lo] - ref[lo])) < 1e-12 * scale
    assert np.max(np.abs(edm[lo] - refe[lo])) < 1e-12 * scale
    assert np.all(dm[up] == 0) and np.all(edm[up] == 0)
    # Hermitian: the diagonal is real to rounding
    assert np.max(np.abs(np.diag(dm).imag)) < 1e-12 * scale


@pytest.mark.parametrize("system", ["cluster216", "diamond2"])
def test_sparse_kernels(system, request):
    """H0/S assembly, unpack, pack and Mulliken against the oracle on the same neighbour list."""
    api = _api()
    geo, params, opts, calc, res = request.getfixturevalue(system)
    ra = reference_arrays(calc)
    sk = api.SkHandle(params)
    h0, s = api.build_h0_s(sk, ra)
    # the 5th-order tail beyond the grid uses a second difference with step 1e-5 bohr
    # (slakoeqgrid.F90:233-239), which turns last-bit differences of the Neville values into
    # ~1e-11; inside the grid the blocks agree to rounding.
    assert np.max(np.abs(h0 - calc.h0)) < 1e-10
    assert np.max(np.abs(s - calc.s)) < 1e-10
    from oracle import dftb_ref as R
    dense = api.unpack_hs(ra, calc.s)
    ref = R.unpack(calc.s, params, geo.species, calc.sp, calc.off)
    full = R.symmetrize_lower(ref)
    assert np.max(np.abs(dense - full)) < 1e-14
    rng = np.random.default_rng(1)
    m = rng.standard_normal(dense.shape)
    m = m + m.T
    packed = api.pack_hs(ra, m)
    assert np.max(np.abs(packed - R.pack(m, params, geo.species, calc.sp, calc.off))) == 0.0
    q = api.mulliken(ra, calc.s, res.rho)
    qref = R.mulliken(calc.s, res.rho, geo.species, calc.sp, calc.off)
    assert np.max(np.abs(q - qref)) < 1e-13


@pytest.mark.parametrize("tiles", [2, 4, 8, 16, 0])
@pytest.mark.parametrize("n", [1, 2, 77, 700])
def test_dsytrd_lapack_conventions(n, tiles):
    """Tridiagonalisation against LAPACK's own dsytrd ('L'): d, e, tau and the reflectors, and
    A = Q T Q^T with Q = H(0) ... H(n-2), H(i) = I - tau_i v v^T, v(i+1) = 1, v(i+2:) stored
    below the subdiagonal of column i."""
    from scipy.linalg import lapack
    api = _api()
    rng = np.random.default_rng(n + tiles)
    a = rng.standard_normal((n, n))
    a = a + a.T
    api.set_tuning(0, tiles)
    try:
        r, d, e, tau = api.dsytrd(a)
    finally:
        api.set_tuning(0, 0)
    scale = max(1.0, np.abs(a).max())
    tol = 1e-12 * max(n, 1) * scale
    ref, dr, er, taur, info = lapack.dsytrd(a, lower=1)
    assert info == 0
    assert np.max(np.abs(d - dr)) < tol
    if n > 1:
        assert np.max(np.abs(e - er)) < tol
        assert np.max(np.abs(tau - taur[: n - 1])) < tol
        assert np.max(np.abs(np.tril(r, -2) - np.tril(ref, -2))) < tol
        q = np.eye(n)
        for i in range(n - 2, -1, -1):
            v = np.zeros(n)
            v[i + 1] = 1.0
            v[i + 2:] = r[i + 2:, i]
            q -= tau[i] * np.outer(v, v @ q)
        t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert np.max(np.abs(q @ t @ q.T - a)) < tol
