// Generalized symmetric-definite eigensolver H C = S C E on the device:
//   S = L L^T (blocked Cholesky, once per geometry), Linv = L^-1;
//   A = Linv H Linv^T (two MFMA GEMMs exploiting the triangular factor);
//   A = Q T Q^T (one-stage blocked Householder tridiagonalisation);
//   T = Z E Z^T (divide and conquer);
//   C = Linv^T Q Z (blocked back-transformation of the requested leading columns).
// The contract is that of LAPACK dsygvd with itype = 1, uplo = 'L' (the routine DFTB+ calls
// through hegvd / magmaHegvd, src/dftbp/math/eigensolver.F90:561-660, 1143-1245).
#include "eigen.h"

namespace dftbhip {

int trtri_from_diag(hipStream_t s, const double* l, long ldl, double* linv, long ldv, int n,
                    double* work);
int potrf_lower_impl(hipStream_t s, double* a, long lda, int n, double* linv, long ldv,
                     int* dInfo, int* hInfo);

int EigenSolver::ensure(int n_) {
  if (n_ == n && L.p) return 0;
  n = n_;
  const size_t nn = (size_t)n * n;
  DH_TRY(L.alloc(nn));
  DH_TRY(Linv.alloc(nn));
  DH_TRY(C.alloc(nn));
  DH_TRY(Zd.alloc(nn));
  DH_TRY(T.alloc(nn));
  DH_TRY(d.alloc(n));
  DH_TRY(e.alloc(n));
  DH_TRY(tau.alloc(n));
  DH_TRY(w.alloc(n));
  DH_TRY(dinfo.alloc(1));
  workSize = std::max(stedc_work_size(n), sytrd_work_size(n));
  workSize = std::max(workSize, ormtr_work_size(n, n));
  workSize = std::max(workSize, nn / 2 + 64 * (size_t)n + 4096);
  DH_TRY(work.alloc(workSize));
  bandBytes = band_work_bytes(n, 8);
  DH_TRY(band.alloc((bandBytes + 7) / 8));
  return 0;
}

int EigenSolver::factor_overlap(hipStream_t s, const double* S, long lds, int* info) {
  DH_TRY(copy2d<double>(s, S, lds, L.p, n, n, n));
  DH_TRY(fill<double>(s, Linv.p, (size_t)n * n, 0.0));
  int hinfo = 0;
  DH_TRY(potrf_lower_impl(s, L.p, n, n, Linv.p, n, dinfo.p, &hinfo));
  *info = hinfo;
  if (hinfo != 0) return 0;
  DH_TRY(trtri_from_diag(s, L.p, n, Linv.p, n, n, work.p));
  factored = true;
  return 0;
}

int EigenSolver::reduce(hipStream_t s, const double* Li, const double* H, long ldh, double* A) {
  DH_CHECK(factored, "eigensolver: overlap not factorised");
  DH_TRY(gemm<double>(s, false, false, n, n, n, 1.0, Li, n, H, ldh, 0.0, T.p, n, false,
                      kKALower));
  DH_TRY(gemm<double>(s, false, true, n, n, n, 1.0, T.p, n, Li, n, 0.0, A, n, true, kKBUpper));
  return 0;
}

int EigenSolver::tridiag_eigen(hipStream_t s, const double* dDev, const double* eDev,
                               std::vector<double>& wHost, int* info) {
  *info = 0;
  std::vector<double> dh(n), eh(std::max(n - 1, 1));
  DH_HIP(hipMemcpyAsync(dh.data(), dDev, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (n > 1)
    DH_HIP(hipMemcpyAsync(eh.data(), eDev, (n - 1) * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  DH_TRY(stedc(s, n, dh.data(), eh.data(), w.p, Zd.p, n, work.p, workSize, info, dc_ws()));
  if (*info != 0) return 0;
  wHost.resize(n);
  DH_HIP(hipMemcpyAsync(wHost.data(), w.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

int EigenSolver::vectors_from(hipStream_t s, const double* Li, const double* A,
                              const double* tauDev, int nvec, double* X, long ldx) {
  if (nvec <= 0) return 0;
  DH_TRY(ormtr_lower(s, A, n, n, tauDev, Zd.p, n, nvec, work.p, workSize));
  DH_TRY(gemm<double>(s, true, false, n, nvec, n, 1.0, Li, n, Zd.p, n, 0.0, X, ldx, false,
                      kKAUpper));
  return 0;
}

int EigenSolver::solve(hipStream_t s, const double* H, long ldh, std::vector<double>& wHost,
                       int* info, double* times) {
  DH_CHECK(factored, "eigensolver: overlap not factorised");
  *info = 0;
  Events<5> ev;
  DH_TRY(ev.create());
  DH_HIP(hipEventRecord(ev[0], s));
  DH_TRY(reduce(s, Linv.p, H, ldh, C.p));
  DH_HIP(hipEventRecord(ev[1], s));
  if (twoStage) {
    DH_TRY(symmetrize_lower(s, C.p, n, n));
    DH_TRY(sytrd_2stage(s, C.p, n, n, d.p, e.p, band.p, bandBytes));
  } else {
    DH_TRY(sytrd_lower(s, C.p, n, n, d.p, e.p, tau.p, work.p, workSize));
  }
  DH_HIP(hipEventRecord(ev[2], s));
  DH_TRY(tridiag_eigen(s, d.p, e.p, wHost, info));
  if (*info != 0) return 0;
  DH_HIP(hipEventRecord(ev[3], s));
  DH_HIP(hipEventSynchronize(ev[3]));
  if (times) {
    float ms;
    DH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    times[0] += ms * 1e-3;
    DH_HIP(hipEventElapsedTime(&ms, ev[1], ev[2]));
    times[1] += ms * 1e-3;
    DH_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
    times[2] += ms * 1e-3;
  }
  return 0;
}

int EigenSolver::vectors(hipStream_t s, int nvec, double* X, long ldx) {
  if (nvec <= 0) return 0;
  if (twoStage) {
    DH_TRY(ormtr_2stage(s, C.p, n, n, Zd.p, n, nvec, band.p, bandBytes, work.p,
                        workSize * sizeof(double)));
    DH_TRY(gemm<double>(s, true, false, n, nvec, n, 1.0, Linv.p, n, Zd.p, n, 0.0, X, ldx, false,
                        kKAUpper));
    return 0;
  }
  return vectors_from(s, Linv.p, C.p, tau.p, nvec, X, ldx);
}

// Ogita-Aishima correction E (n x nv) from G = X^T S X and M = X^T H X (both full):
//   lam_j = M_jj / G_jj,  R = I - G,
//   E_jj = R_jj / 2,
//   E_ij = (M_ij + lam_j R_ij) / (lam_j - lam_i)   where that is small (|E_ij| < kOaMax),
//   E_ij = R_ij / 2                                 otherwise (near-degenerate pairs: the pair
//                                                   is only re-orthonormalised).
constexpr double kOaMax = 0.25;

__global__ void oa_lambda_kernel(const double* M, const double* G, int n, double* lam) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lam[i] = M[i + (long)i * n] / G[i + (long)i * n];
}

__global__ void oa_correction_kernel(const double* M, const double* G, const double* lam, int n,
                                     int nv, double* E) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n || j >= nv) return;
  const long ij = i + (long)j * n;
  double e;
  if (i == j) {
    e = 0.5 * (1.0 - G[ij]);
  } else {
    const double r = -G[ij];
    const double num = M[ij] + lam[j] * r, den = lam[j] - lam[i];
    e = (std::fabs(num) < kOaMax * std::fabs(den)) ? num / den : 0.5 * r;
  }
  E[ij] = e;
}

// The same correction written transposed, Et(j, i) = E(i, j), from the symmetric M and G (both
// triangles stored): thread (row j, column i) reads M and G at (j, i) = (i, j), so reads and
// writes run along columns. lam is formed first by oa_lambda_kernel.
// jacobi = 1: a pair outside the first-order range (near-degenerate against its coupling) is
// rotated by the angle of the exact 2x2 problem, theta = atan(2 M_ij / (lam_j - lam_i)) / 2
// (|theta| <= pi/4, equal to the first-order E_ij for a weak coupling), instead of only being
// re-orthonormalised
__global__ void oa_correction_t_kernel(const double* M, const double* G, const double* lam, int n,
                                       double* Et, int jacobi, int* nClipped) {
  const int jr = blockIdx.x * blockDim.x + threadIdx.x;   // row of Et: the vector index j of E
  const int ic = blockIdx.y;                               // column of Et: the index i of E
  const bool in = jr < n && ic < n;
  const long q = jr + (long)ic * n;
  double e = 0.0;
  bool clipped = false;
  if (in) {
    if (jr == ic) {
      e = 0.5 * (1.0 - G[q]);
    } else {
      const double r = -G[q];
      const double num = M[q] + lam[jr] * r, den = lam[jr] - lam[ic];
      if (std::fabs(num) < kOaMax * std::fabs(den)) {
        e = num / den;
      } else {
        clipped = jr > ic;   // each pair counted once
        if (jacobi) {
          const double th =
              den != 0.0 ? 0.5 * atan(2.0 * M[q] / den) : copysign(0.25 * M_PI, M[q]);
          e = 0.5 * r + sin(th);
        } else {
          e = 0.5 * r;
        }
      }
    }
    Et[q] = e;
  }
  // one atomic per wave for the count of pairs outside the first-order range
  if (nClipped) {
    const unsigned long long m = __ballot(clipped);
    if (m != 0ull && (threadIdx.x & 63) == (unsigned)(__ffsll((long long)m) - 1))
      atomicAdd(nClipped, (int)__popcll(m));
  }
}

static int g_oaJacobi = 0;
int set_oa_mode(int mode) {
  DH_CHECK(mode == 0 || mode == 1, "oa mode: 0 or 1");
  g_oaJacobi = mode;
  return 0;
}

int oa_correction_t(hipStream_t s, const double* M, const double* G, int n, double* lam,
                    double* Et, int* nClipped) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(oa_lambda_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, M, G, n, lam);
  DH_LAUNCH_CHECK();
  hipLaunchKernelGGL(oa_correction_t_kernel, dim3(cdiv(n, 256), n), dim3(256), 0, s, M, G, lam, n,
                     Et, g_oaJacobi, nClipped);
  DH_LAUNCH_CHECK();
  return 0;
}

// ---- residual form (FP16 products) ----------------------------------------------------------
constexpr int kOaChunks = 64;   // k-chunks of the row reductions (partial results per chunk)

// part[c nv + j] = sum over the chunk c of k of a(j, k) b(j, k) (mode 0), or the maximum of
// |a(j, k) - lam_j b(j, k)| (mode 1); threads run along j, so every load is coalesced
__global__ void oa_rowreduce_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                    long ld, int nv, int n, const double* __restrict__ lam,
                                    int mode, double* __restrict__ part) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  const int chunk = (n + kOaChunks - 1) / kOaChunks;
  const int k0 = blockIdx.y * chunk, k1 = min(n, k0 + chunk);
  double acc = 0.0;
  if (mode == 0) {
    for (int k = k0; k < k1; ++k) acc = fma(a[j + (long)k * ld], b[j + (long)k * ld], acc);
  } else {
    const double l = lam[j];
    for (int k = k0; k < k1; ++k)
      acc = fmax(acc, fabs(a[j + (long)k * ld] - l * b[j + (long)k * ld]));
  }
  part[(long)blockIdx.y * nv + j] = acc;
}

__global__ void oa_rayleigh_kernel(const double* __restrict__ part, int nv,
                                   const double* __restrict__ G, long ldg, double* __restrict__ lam) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  double m = 0.0;
  for (int c = 0; c < kOaChunks; ++c) m += part[(long)c * nv + j];
  lam[j] = m / G[j + (long)j * ldg];
}

__global__ void oa_scale_kernel(const double* __restrict__ part, int nv, double* __restrict__ rsc,
                                float* __restrict__ rinv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  double mx = 0.0;
  for (int c = 0; c < kOaChunks; ++c) mx = fmax(mx, part[(long)c * nv + j]);
  // max |r_j| 2^p in [2^14, 2^15); p bounded so that 2^-p stays a normal FP32 number
  const int p = mx > 0.0 ? min(100, max(-100, 14 - ilogb(mx))) : 0;
  rsc[j] = ldexp(1.0, p);
  rinv[j] = ldexpf(1.0f, -p);
}

int oa_rayleigh(hipStream_t s, const double* Xt, const double* YH, long ld, int nv, int n,
                const double* G, long ldg, double* part, double* lam) {
  if (nv <= 0) return 0;
  hipLaunchKernelGGL(oa_rowreduce_kernel, dim3(cdiv(nv, 256), kOaChunks), dim3(256), 0, s, Xt, YH,
                     ld, nv, n, (const double*)nullptr, 0, part);
  DH_LAUNCH_CHECK();
  hipLaunchKernelGGL(oa_rayleigh_kernel, dim3(cdiv(nv, 256)), dim3(256), 0, s, part, nv, G, ldg,
                     lam);
  DH_LAUNCH_CHECK();
  return 0;
}

int oa_residual_scale(hipStream_t s, const double* YH, const double* YS, long ld, int nv, int n,
                      const double* lam, double* part, double* rsc, float* rinv) {
  if (nv <= 0) return 0;
  hipLaunchKernelGGL(oa_rowreduce_kernel, dim3(cdiv(nv, 256), kOaChunks), dim3(256), 0, s, YH, YS,
                     ld, nv, n, lam, 1, part);
  DH_LAUNCH_CHECK();
  hipLaunchKernelGGL(oa_scale_kernel, dim3(cdiv(nv, 256)), dim3(256), 0, s, part, nv, rsc, rinv);
  DH_LAUNCH_CHECK();
  return 0;
}

// 64 x 64 tiles of (j, i): Ft and G read along j, E16 written along i through LDS
// In-band couplings Ct[j + i ldc] = x_i^T r_j = M_ij - lam_j G_ij (0 for i = j) for the pairs
// whose 128 x 128 tiles are at most `band` apart, from the band tiles of M = X^T H X and
// G = X^T S X (lower triangle, leading dimension ldb); threads run along j.
__global__ void oa_band_coupling_kernel(const double* __restrict__ Mb, const double* __restrict__ Gb,
                                        long ldb, const double* __restrict__ lam, int n, int band,
                                        double* __restrict__ Ct, long ldc) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int i = (j / kGemmTile - band) * kGemmTile + (int)blockIdx.y;
  if (i < 0 || i >= n) return;
  const long q = max(i, j) + (long)min(i, j) * ldb;
  Ct[j + (long)i * ldc] = i == j ? 0.0 : Mb[q] - lam[j] * Gb[q];
}

int oa_band_coupling(hipStream_t s, const double* Mb, const double* Gb, long ldb, int n, int band,
                     double* lam, double* Ct, long ldc) {
  if (n <= 0) return 0;
  DH_CHECK(band >= 0, "oa_band_coupling: negative band");
  hipLaunchKernelGGL(oa_lambda_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, Mb, Gb, n, lam);
  DH_LAUNCH_CHECK();
  hipLaunchKernelGGL(oa_band_coupling_kernel, dim3(cdiv(n, 256), kGemmTile * (2 * band + 1)),
                     dim3(256), 0, s, Mb, Gb, ldb, lam, n, band, Ct, ldc);
  DH_LAUNCH_CHECK();
  return 0;
}

// G is read from its lower triangle, (max(i, j), min(i, j)), and only where it is used: on the
// diagonal and for the pairs outside the first-order range. gband >= 0: G holds only the
// 128 x 128 diagonal tiles and `gband` tile rows below them; a pair outside the first-order
// range whose element lies outside that band is counted in nFar and left uncorrected (the
// caller then forms the whole lower triangle and repeats the correction). Ct (may be null):
// the FP64 couplings of the pairs whose tiles are at most `band` apart, used instead of Ft.
__device__ __forceinline__ double oa_f16_element(const float* __restrict__ Ft, long ldf,
                                                 const double* __restrict__ G, long ldg,
                                                 const double* __restrict__ Ct, long ldc,
                                                 const double* __restrict__ lam, double lj,
                                                 int i, int j, int jacobi, int band, int gband,
                                                 int& nclip, int& nfar) {
  const int hi = max(i, j), lo = min(i, j);
  const long qg = hi + (long)lo * ldg;
  if (i == j) return 0.5 * (1.0 - G[qg]);
  const int dt = hi / kGemmTile - lo / kGemmTile;
  const double num = (Ct && dt <= band) ? Ct[j + (long)i * ldc] : (double)Ft[j + (long)i * ldf];
  const double den = lj - lam[i];
  if (fabs(num) < kOaMax * fabs(den)) return num / den;
  if (gband >= 0 && dt > gband) {
    nfar += j > i;
    return 0.0;
  }
  const double r = -G[qg];
  nclip += j > i;   // each pair counted once
  if (!jacobi) return 0.5 * r;
  const double mij = num - lj * r;   // x_i^T H x_j
  const double th = den != 0.0 ? 0.5 * atan(2.0 * mij / den) : copysign(0.25 * M_PI, mij);
  return 0.5 * r + sin(th);
}

// pass 0: the largest |E(i, j)| of each column j (non-negative floats ordered as their bit
// patterns: one atomicMax per thread and column) and the pair counters; pass 1: E16(i, j) =
// E(i, j) 2^q_j with q_j from that maximum (largest scaled value in [2^14, 2^15)), written
// along i through LDS, and rinv[j] = 2^-q_j. 64 x 64 tiles of (j, i): Ft and G read along j.
__global__ __launch_bounds__(256) void oa_correction_f16_kernel(
    const float* __restrict__ Ft, long ldf, const double* __restrict__ G, long ldg,
    const double* __restrict__ Ct, long ldc,
    const double* __restrict__ lam, int n, _Float16* __restrict__ E16, long lde, int jacobi,
    int band, int gband, int pass, unsigned* __restrict__ colMax, float* __restrict__ rinv,
    int* __restrict__ nClipped, int* __restrict__ nFar) {
  __shared__ float t[64][65];
  const int j0 = blockIdx.x * 64, i0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int j = j0 + tx;
  const double lj = j < n ? lam[j] : 0.0;
  int nclip = 0, nfar = 0;
  if (pass == 0) {
    float mx = 0.0f;
    if (j < n)
      for (int ii = ty; ii < 64; ii += 4) {
        const int i = i0 + ii;
        if (i < n)
          mx = fmaxf(mx, (float)fabs(oa_f16_element(Ft, ldf, G, ldg, Ct, ldc, lam, lj, i, j,
                                                    jacobi, band, gband, nclip, nfar)));
      }
    if (j < n && mx > 0.0f) atomicMax(colMax + j, __float_as_uint(mx));
    // one atomic per wave and counter
    int w = nclip, f = nfar;
    for (int o = 32; o > 0; o >>= 1) {
      w += __shfl_xor(w, o);
      f += __shfl_xor(f, o);
    }
    if ((threadIdx.x & 63) == 0) {
      if (nClipped && w != 0) atomicAdd(nClipped, w);
      if (nFar && f != 0) atomicAdd(nFar, f);
    }
    return;
  }
  double sc = 1.0;
  if (j < n) {
    const float mx = __uint_as_float(colMax[j]);
    const int q = mx > 0.0f ? min(100, max(-100, 14 - ilogbf(mx))) : 0;
    sc = ldexp(1.0, q);
    if (blockIdx.y == 0 && ty == 0) rinv[j] = ldexpf(1.0f, -q);
  }
  for (int ii = ty; ii < 64; ii += 4) {
    const int i = i0 + ii;
    double e = 0.0;
    if (j < n && i < n)
      e = oa_f16_element(Ft, ldf, G, ldg, Ct, ldc, lam, lj, i, j, jacobi, band, gband, nclip,
                         nfar);
    t[ii][tx] = (float)(e * sc);
  }
  __syncthreads();
  for (int jj = ty; jj < 64; jj += 4) {
    const int i = i0 + tx, jo = j0 + jj;
    if (i < n && jo < n) E16[i + (long)jo * lde] = (_Float16)t[tx][jj];
  }
}

int oa_correction_f16(hipStream_t s, const float* Ft, long ldf, const double* G, long ldg,
                      const double* Ct, long ldc, const double* lam, int n, int band, int gband,
                      unsigned* colMax, int* nClipped, int* nFar) {
  if (n <= 0) return 0;
  DH_CHECK(gband < 0 || nFar, "oa_correction_f16: a banded G needs the far-pair counter");
  DH_TRY(fill<unsigned>(s, colMax, n, 0u));
  hipLaunchKernelGGL(oa_correction_f16_kernel, dim3(cdiv(n, 64), cdiv(n, 64)), dim3(256), 0, s,
                     Ft, ldf, G, ldg, Ct, ldc, lam, n, (_Float16*)nullptr, 0L, g_oaJacobi, band,
                     gband, 0, colMax, (float*)nullptr, nClipped, nFar);
  DH_LAUNCH_CHECK();
  return 0;
}

int oa_correction_f16_write(hipStream_t s, const float* Ft, long ldf, const double* G, long ldg,
                            const double* Ct, long ldc, const double* lam, int n, int band,
                            int gband, const unsigned* colMax, _Float16* E16, long lde,
                            float* rinv) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(oa_correction_f16_kernel, dim3(cdiv(n, 64), cdiv(n, 64)), dim3(256), 0, s,
                     Ft, ldf, G, ldg, Ct, ldc, lam, n, E16, lde, g_oaJacobi, band, gband, 1,
                     const_cast<unsigned*>(colMax), rinv, (int*)nullptr, (int*)nullptr);
  DH_LAUNCH_CHECK();
  return 0;
}

int EigenSolver::refine(hipStream_t s, const double* H, const double* S, const float* X0,
                        int steps, int nvec, double* X, long ldx, std::vector<double>& lamHost) {
  DH_CHECK(workSize >= 2 * (size_t)n * n, "refine: workspace too small");
  DH_TRY((convert2d<float, double>(s, X0, n, Zd.p, n, n, n)));
  return refine_from_basis(s, H, S, steps, nvec, X, ldx, lamHost);
}

int EigenSolver::refine(hipStream_t s, const double* H, const double* S, const double* X0,
                        int steps, int nvec, double* X, long ldx, std::vector<double>& lamHost) {
  DH_CHECK(workSize >= 2 * (size_t)n * n, "refine: workspace too small");
  DH_TRY(copy2d<double>(s, X0, n, Zd.p, n, n, n));
  return refine_from_basis(s, H, S, steps, nvec, X, ldx, lamHost);
}

// the steps on the basis already in Zd
int EigenSolver::refine_from_basis(hipStream_t s, const double* H, const double* S, int steps,
                                   int nvec, double* X, long ldx,
                                   std::vector<double>& lamHost) {
  steps = std::max(steps, 1);
  double* Y = C.p;                  // H X or S X
  double* M = work.p;               // X^T H X
  double* G = T.p;                  // X^T S X
  double* E = work.p + (size_t)n * n;
  double* Xa = Zd.p;                // basis of the current step
  for (int st = 0; st < steps; ++st) {
    const bool last = st == steps - 1;
    const int nv = last ? nvec : n;
    DH_TRY(gemm<double>(s, false, false, n, n, n, 1.0, H, n, Xa, n, 0.0, Y, n));
    DH_TRY(gemm<double>(s, true, false, n, n, n, 1.0, Xa, n, Y, n, 0.0, M, n, true));
    DH_TRY(symmetrize_lower(s, M, n, n));
    DH_TRY(gemm<double>(s, false, false, n, n, n, 1.0, S, n, Xa, n, 0.0, Y, n));
    DH_TRY(gemm<double>(s, true, false, n, n, n, 1.0, Xa, n, Y, n, 0.0, G, n, true));
    DH_TRY(symmetrize_lower(s, G, n, n));
    hipLaunchKernelGGL(oa_lambda_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, M, G, n, w.p);
    DH_LAUNCH_CHECK();
    if (nv > 0) {
      hipLaunchKernelGGL(oa_correction_kernel, dim3(cdiv(n, 256), nv), dim3(256), 0, s, M, G,
                         w.p, n, nv, E);
      DH_LAUNCH_CHECK();
    }
    // X' = X (I + E): into X (last step) or into Y (the next basis, then swapped with Xa)
    double* out = last ? X : Y;
    const long ldo = last ? ldx : n;
    if (nv > 0) {
      DH_TRY(copy2d<double>(s, Xa, n, out, ldo, n, nv));
      DH_TRY(gemm<double>(s, false, false, n, nv, n, 1.0, Xa, n, E, n, 1.0, out, ldo));
    }
    if (!last) DH_TRY(copy2d<double>(s, Y, n, Xa, n, n, n));
  }
  lamHost.resize(n);
  DH_HIP(hipMemcpyAsync(lamHost.data(), w.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

int EigenSolver::prepare_mixed(hipStream_t s) {
  DH_CHECK(factored, "eigensolver: overlap not factorised");
  DH_TRY(Linv32.alloc((size_t)n * n));
  DH_TRY(d32.alloc(n));
  DH_TRY(e32.alloc(std::max(n, 2)));
  DH_TRY(tau32.alloc(std::max(n, 1)));
  DH_TRY((convert2d<double, float>(s, Linv.p, n, Linv32.p, n, n, n)));
  return 0;
}

// FP32 work arrays share the FP64 buffers: the reduced matrix and the Householder vectors live
// in C, the GEMM temporary and the FP32 copy of the tridiagonal eigenvectors in T.
int EigenSolver::solve_mixed(hipStream_t s, const float* H, long ldh, std::vector<double>& wHost,
                             int* info, double* times) {
  DH_CHECK(factored && Linv32.p, "eigensolver: mixed path not prepared");
  *info = 0;
  float* A32 = reinterpret_cast<float*>(C.p);
  float* T32 = reinterpret_cast<float*>(T.p);
  Events<4> ev;
  DH_TRY(ev.create());
  DH_HIP(hipEventRecord(ev[0], s));
  DH_TRY(gemm<float>(s, false, false, n, n, n, 1.0f, Linv32.p, n, H, ldh, 0.0f, T32, n, false,
                     kKALower));
  DH_TRY(gemm<float>(s, false, true, n, n, n, 1.0f, T32, n, Linv32.p, n, 0.0f, A32, n, true,
                     kKBUpper));
  DH_HIP(hipEventRecord(ev[1], s));
  if (twoStage) {
    DH_TRY(symmetrize_lower(s, A32, n, n));
    DH_TRY(sytrd_2stage_f32(s, A32, n, n, d32.p, e32.p, band.p, bandBytes));
  } else {
    DH_TRY(sytrd_lower_f32(s, A32, n, n, d32.p, e32.p, tau32.p, work.p, workSize * sizeof(double)));
  }
  DH_HIP(hipEventRecord(ev[2], s));
  std::vector<float> df(n), ef(std::max(n - 1, 1));
  DH_HIP(hipMemcpyAsync(df.data(), d32.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
  if (n > 1)
    DH_HIP(hipMemcpyAsync(ef.data(), e32.p, (n - 1) * sizeof(float), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  std::vector<double> dh(df.begin(), df.end()), eh(ef.begin(), ef.end());
  DH_TRY(stedc(s, n, dh.data(), eh.data(), w.p, Zd.p, n, work.p, workSize, info, dc_ws()));
  DH_HIP(hipEventRecord(ev[3], s));
  wHost.resize(n);
  DH_HIP(hipMemcpyAsync(wHost.data(), w.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  if (times) {
    float ms;
    for (int k = 0; k < 3; ++k) {
      DH_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      times[k] += ms * 1e-3;
    }
  }
  return 0;
}

int EigenSolver::vectors_mixed(hipStream_t s, int nvec, float* X, long ldx) {
  if (nvec <= 0) return 0;
  float* A32 = reinterpret_cast<float*>(C.p);
  float* Z32 = reinterpret_cast<float*>(T.p);
  DH_TRY((convert2d<double, float>(s, Zd.p, n, Z32, n, n, nvec)));
  if (twoStage)
    DH_TRY(ormtr_2stage_f32(s, A32, n, n, Z32, n, nvec, band.p, bandBytes, work.p,
                            workSize * sizeof(double)));
  else
    DH_TRY(ormtr_lower_f32(s, A32, n, n, tau32.p, Z32, n, nvec, work.p,
                           workSize * sizeof(double)));
  DH_TRY(gemm<float>(s, true, false, n, nvec, n, 1.0f, Linv32.p, n, Z32, n, 0.0f, X, ldx, false,
                     kKAUpper));
  return 0;
}

}  // namespace dftbhip
