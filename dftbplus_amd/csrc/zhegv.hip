// Complex Hermitian generalized eigensolver H C = S C E (itype 1, lower storage), the contract of
// LAPACK zhegvd that DFTB+ reaches through hegvd / magmaHegvd (dcomplex)
// (src/dftbp/math/eigensolver.F90:113-117, 1143-1245).
//
// Complex matrices live as two real planes (re, im), so every product is a set of real MFMA
// GEMMs (gemm.hip):
//   S = L L^H         blocked complex Cholesky (64-wide LDS diagonal blocks, plane GEMMs);
//   Linv = L^-1       recursive doubling over the diagonal-block inverses;
//   A = Linv H Linv^H two complex GEMMs with triangular k-limits;
//   A = Q T Q^T       on the real embedding [[Re A, -Im A], [Im A, Re A]] (2n, symmetric), whose
//                     spectrum is that of A with every level twice; one-stage tridiagonalisation,
//                     divide and conquer and back-transformation of the real solver;
//   Y                 one complex eigenvector per level: an embedded eigenvector [y_r; y_i] is
//                     the complex y_r + i y_i, and its partner in the pair is i times the same
//                     vector; levels within 1e-8 (relative) form a cluster whose 2m embedded
//                     vectors are reduced to m orthonormal complex ones by Gram-Schmidt in
//                     eigenvalue order;
//   C = Linv^H Y      one complex GEMM.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "zhegv.h"

namespace dftbhip {

// ---------------------------------------------------------------------------------------------
// complex GEMM on planes: C = alpha op(A) op(B) + beta C, op = 'N' or 'C' (conjugate transpose),
// real alpha and beta. With op(A) = Ar' + i sA Ai' (sA = -1 for 'C', ' = transposed for 'C'):
//   Re C = alpha (Ar' Br' - sA sB Ai' Bi') + beta Re C
//   Im C = alpha (sB Ar' Bi' + sA Ai' Br') + beta Im C
// C must not overlap A or B.
// ---------------------------------------------------------------------------------------------
int zgemm(hipStream_t s, char opA, char opB, int M, int N, int K, double alpha, ZcPlanes A,
          ZcPlanes B, double beta, ZPlanes C, bool lowerOut, GemmKLimit klim) {
  const bool ta = opA == 'C', tb = opB == 'C';
  const double sA = ta ? -1.0 : 1.0, sB = tb ? -1.0 : 1.0;
  DH_TRY(gemm<double>(s, ta, tb, M, N, K, alpha, A.r, A.ld, B.r, B.ld, beta, C.r, C.ld, lowerOut,
                      klim));
  DH_TRY(gemm<double>(s, ta, tb, M, N, K, -alpha * sA * sB, A.i, A.ld, B.i, B.ld, 1.0, C.r, C.ld,
                      lowerOut, klim));
  DH_TRY(gemm<double>(s, ta, tb, M, N, K, alpha * sB, A.r, A.ld, B.i, B.ld, beta, C.i, C.ld,
                      lowerOut, klim));
  DH_TRY(gemm<double>(s, ta, tb, M, N, K, alpha * sA, A.i, A.ld, B.r, B.ld, 1.0, C.i, C.ld,
                      lowerOut, klim));
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Complex Cholesky
// ---------------------------------------------------------------------------------------------
constexpr int kZNB = 64;

// Factor the jb x jb Hermitian diagonal block (lower part read) in place, write its inverse to
// linv. dInfo receives offset + k + 1 for the first pivot whose real part is not positive.
__global__ __launch_bounds__(256) void zpotrf_diag_kernel(double* ar, double* ai, long lda,
                                                          int jb, double* vr, double* vi,
                                                          long ldv, int offset, int* dInfo) {
  __shared__ double Lr[kZNB][kZNB + 1], Li[kZNB][kZNB + 1];
  __shared__ double Xr[kZNB][kZNB + 1], Xi[kZNB][kZNB + 1];
  __shared__ int failed;
  const int tid = threadIdx.x;
  if (tid == 0) failed = (*dInfo != 0);
  for (int e = tid; e < jb * jb; e += blockDim.x) {
    const int i = e % jb, c = e / jb;
    const long o = i + (long)c * lda;
    Lr[i][c] = (i >= c) ? ar[o] : 0.0;
    Li[i][c] = (i > c) ? ai[o] : 0.0;   // the diagonal of a Hermitian matrix is real
  }
  __syncthreads();
  if (failed) return;
  for (int k = 0; k < jb; ++k) {
    if (tid == 0) {
      const double akk = Lr[k][k];
      if (!(akk > 0.0)) {
        failed = 1;
        if (*dInfo == 0) *dInfo = offset + k + 1;
      } else {
        Lr[k][k] = sqrt(akk);
      }
    }
    __syncthreads();
    if (failed) return;
    const double dkk = Lr[k][k];
    for (int i = k + 1 + tid; i < jb; i += blockDim.x) {
      Lr[i][k] /= dkk;
      Li[i][k] /= dkk;
    }
    __syncthreads();
    // a(i, c) -= l(i, k) conj(l(c, k)) for i >= c > k
    const int m = jb - k - 1;
    for (int e = tid; e < m * m; e += blockDim.x) {
      const int i = k + 1 + e % m, c = k + 1 + e / m;
      if (i >= c) {
        const double pr = Lr[i][k], pi = Li[i][k], qr = Lr[c][k], qi = Li[c][k];
        Lr[i][c] -= pr * qr + pi * qi;
        if (i > c) Li[i][c] -= pi * qr - pr * qi;
      }
    }
    __syncthreads();
  }
  // inverse by forward substitution, one column per thread
  if (tid < jb) {
    const int c = tid;
    for (int i = 0; i < c; ++i) {
      Xr[i][c] = 0.0;
      Xi[i][c] = 0.0;
    }
    Xr[c][c] = 1.0 / Lr[c][c];
    Xi[c][c] = 0.0;
    for (int i = c + 1; i < jb; ++i) {
      double sr = 0.0, si = 0.0;
      for (int k = c; k < i; ++k) {
        sr += Lr[i][k] * Xr[k][c] - Li[i][k] * Xi[k][c];
        si += Lr[i][k] * Xi[k][c] + Li[i][k] * Xr[k][c];
      }
      const double d = Lr[i][i];   // real diagonal
      Xr[i][c] = -sr / d;
      Xi[i][c] = -si / d;
    }
  }
  __syncthreads();
  for (int e = tid; e < jb * jb; e += blockDim.x) {
    const int i = e % jb, c = e / jb;
    const long o = i + (long)c * lda;
    if (i >= c) {
      ar[o] = Lr[i][c];
      ai[o] = Li[i][c];
    }
    const long ov = i + (long)c * ldv;
    vr[ov] = Xr[i][c];
    vi[ov] = Xi[i][c];
  }
}

// S = L L^H in place (lower), the diagonal-block inverses into linv (zeroed beforehand);
// tmp: two planes of n x 64. *hInfo: LAPACK-style minor of the first non-positive pivot.
static int zpotrf_lower(hipStream_t s, ZPlanes a, int n, ZPlanes linv, ZPlanes tmp, int* dInfo,
                        int* hInfo) {
  DH_TRY(fill<int>(s, dInfo, 1, 0));
  for (int j = 0; j < n; j += kZNB) {
    const int jb = std::min(kZNB, n - j);
    const long dj = j + (long)j * a.ld, dv = j + (long)j * linv.ld;
    hipLaunchKernelGGL(zpotrf_diag_kernel, dim3(1), dim3(256), 0, s, a.r + dj, a.i + dj, a.ld, jb,
                       linv.r + dv, linv.i + dv, linv.ld, j, dInfo);
    DH_LAUNCH_CHECK();
    const int m = n - j - jb;
    if (m > 0) {
      const long o21 = (j + jb) + (long)j * a.ld, o22 = (j + jb) + (long)(j + jb) * a.ld;
      ZPlanes a21{a.r + o21, a.i + o21, a.ld}, a22{a.r + o22, a.i + o22, a.ld};
      ZcPlanes x11{linv.r + dv, linv.i + dv, linv.ld};
      // L21 = A21 L11^-H through tmp, then back into place
      DH_TRY(zgemm(s, 'N', 'C', m, jb, jb, 1.0, a21.c(), x11, 0.0, tmp, false, kKFull));
      DH_TRY(copy2d<double>(s, tmp.r, tmp.ld, a21.r, a21.ld, m, jb));
      DH_TRY(copy2d<double>(s, tmp.i, tmp.ld, a21.i, a21.ld, m, jb));
      // A22 -= L21 L21^H (lower triangle)
      DH_TRY(zgemm(s, 'N', 'C', m, m, jb, -1.0, a21.c(), a21.c(), 1.0, a22, true, kKFull));
    }
  }
  DH_HIP(hipMemcpyAsync(hInfo, dInfo, sizeof(int), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

// Full L^-1 from the factor and the diagonal-block inverses already in linv (strict upper part
// zero): blocks of 64, 128, ... are joined pairwise, X21 = -X22 (L21 X11). tmp: two planes of
// n x n (leading dimension n).
static int ztrtri_from_diag(hipStream_t s, ZcPlanes l, int n, ZPlanes linv, ZPlanes tmp) {
  const int nleaf = (n + kZNB - 1) / kZNB;
  for (int p = 1; p < nleaf; p *= 2) {
    for (int s0 = 0; s0 + p < nleaf; s0 += 2 * p) {
      const int r1 = s0 * kZNB, r2 = (s0 + p) * kZNB;
      const int n1 = r2 - r1;
      const int n2 = std::min((s0 + 2 * p) * kZNB, n) - r2;
      const long o21 = r2 + (long)r1 * l.ld;
      const long v11 = r1 + (long)r1 * linv.ld, v22 = r2 + (long)r2 * linv.ld,
                 v21 = r2 + (long)r1 * linv.ld;
      ZcPlanes l21{l.r + o21, l.i + o21, l.ld};
      ZcPlanes x11{linv.r + v11, linv.i + v11, linv.ld}, x22{linv.r + v22, linv.i + v22, linv.ld};
      ZPlanes x21{linv.r + v21, linv.i + v21, linv.ld};
      DH_TRY(zgemm(s, 'N', 'N', n2, n1, n1, 1.0, l21, x11, 0.0, tmp, false, kKFull));
      DH_TRY(zgemm(s, 'N', 'N', n2, n1, n2, -1.0, x22, tmp.c(), 0.0, x21, false, kKALower));
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Embedding, selection of one complex vector per level, Hermitian completion
// ---------------------------------------------------------------------------------------------
// Hermitian matrix from its lower triangle: re(j, i) = re(i, j), im(j, i) = -im(i, j) for i > j,
// im(i, i) = 0
__global__ void hermitian_from_lower_kernel(double* re, double* im, long ld, int n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int i = (int)(e % n), j = (int)(e / n);
  if (i < j) {
    re[i + (long)j * ld] = re[j + (long)i * ld];
    im[i + (long)j * ld] = -im[j + (long)i * ld];
  } else if (i == j) {
    im[i + (long)j * ld] = 0.0;
  }
}

// lower triangle of the real embedding E = [[Ar, -Ai], [Ai, Ar]] (2n x 2n) from the lower
// triangles of Ar and Ai (A Hermitian: Ai(p, q) = -Ai(q, p))
__global__ void embed_lower_kernel(const double* ar, const double* ai, long ld, int n, double* E,
                                   long lde) {
  const long n2 = 2L * n;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n2 * n2) return;
  const int i = (int)(e % n2), j = (int)(e / n2);
  if (i < j) return;
  double v;
  if (j < n) {
    if (i < n) {
      v = ar[i + (long)j * ld];
    } else {
      const int p = i - n, q = j;
      v = (p > q) ? ai[p + (long)q * ld] : (p < q ? -ai[q + (long)p * ld] : 0.0);
    }
  } else {
    v = ar[(i - n) + (long)(j - n) * ld];
  }
  E[i + (long)j * lde] = v;
}

// column j of Y from embedded column 2 j: y = z(0:n) + i z(n:2n)
__global__ void pair_gather_kernel(const double* z, long ldz, int n, double* yr, double* yi,
                                   long ldy) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int i = (int)(e % n), j = (int)(e / n);
  const double* col = z + (long)(2 * j) * ldz;
  yr[i + (long)j * ldy] = col[i];
  yi[i + (long)j * ldy] = col[n + i];
}

__device__ inline double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
  return t;
}

// One cluster of m levels (pairs j0 .. j0+m-1) per workgroup: its 2m embedded eigenvectors
// (columns 2 j0 ...) as complex vectors, Gram-Schmidt (twice) against the accepted ones in
// eigenvalue order, accepted when the remaining norm exceeds 1/2 (the candidates have norm 1),
// until m are accepted. The candidate is staged in work (2n doubles per cluster).
__global__ __launch_bounds__(256) void cluster_gs_kernel(const double* z, long ldz, int n,
                                                         const int* clusters, double* yr,
                                                         double* yi, long ldy, double* work,
                                                         int* failed) {
  __shared__ double red[4];
  const int j0 = clusters[2 * blockIdx.x], m = clusters[2 * blockIdx.x + 1];
  double* vr = work + (long)blockIdx.x * 2 * n;
  double* vi = vr + n;
  int accepted = 0;
  for (int t = 0; t < 2 * m && accepted < m; ++t) {
    const double* col = z + (long)(2 * j0 + t) * ldz;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      vr[i] = col[i];
      vi[i] = col[n + i];
    }
    __syncthreads();
    for (int rep = 0; rep < 2; ++rep) {
      for (int a = 0; a < accepted; ++a) {
        const double* ar = yr + (long)(j0 + a) * ldy;
        const double* ai = yi + (long)(j0 + a) * ldy;
        double hr = 0.0, hi = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
          hr += ar[i] * vr[i] + ai[i] * vi[i];
          hi += ar[i] * vi[i] - ai[i] * vr[i];
        }
        hr = block_sum(hr, red);
        hi = block_sum(hi, red);
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
          vr[i] -= hr * ar[i] - hi * ai[i];
          vi[i] -= hr * ai[i] + hi * ar[i];
        }
        __syncthreads();
      }
    }
    double nr = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) nr += vr[i] * vr[i] + vi[i] * vi[i];
    nr = sqrt(block_sum(nr, red));
    if (nr > 0.5) {
      double* orr = yr + (long)(j0 + accepted) * ldy;
      double* oi = yi + (long)(j0 + accepted) * ldy;
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        orr[i] = vr[i] / nr;
        oi[i] = vi[i] / nr;
      }
      ++accepted;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && accepted < m) atomicAdd(failed, 1);
}

// ---------------------------------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------------------------------
int ZhegvSolver::ensure(int n_) {
  if (n_ == n && Hr.p) return 0;
  n = n_;
  const size_t nn = (size_t)n * n, n2 = 2 * (size_t)n;
  for (DevBuf<double>* b : {&Hr, &Hi, &Lr, &Li, &Vr, &Vi, &Ar, &Ai}) DH_TRY(b->alloc(nn));
  DH_TRY(E2.alloc(n2 * n2));
  DH_TRY(Z2.alloc(n2 * n2));
  DH_TRY(d2.alloc(n2));
  DH_TRY(e2.alloc(n2));
  DH_TRY(tau2.alloc(n2));
  DH_TRY(w2.alloc(n2));
  DH_TRY(dinfo.alloc(2));
  // stedc's 4 (2n)^2 also covers the Cholesky panel planes (2 n x 64), the inverse's products
  // (2 n^2) and the Gram-Schmidt staging (at most n clusters x 2n)
  workSize = std::max({stedc_work_size((int)n2), sytrd_work_size((int)n2),
                       ormtr_work_size((int)n2, (int)n2), 2 * nn + 64});
  DH_TRY(work.alloc(workSize));
  return 0;
}

void ZhegvSolver::release() {
  for (DevBuf<double>* b : {&Hr, &Hi, &Lr, &Li, &Vr, &Vi, &Ar, &Ai, &E2, &Z2, &d2, &e2, &tau2,
                            &w2, &work})
    b->release();
  dinfo.release();
  clusterDev.release();
  sws.reset();
  n = 0;
}

int ZhegvSolver::solve(hipStream_t s, const double* aDev, long lda, const double* bDev, long ldb,
                       bool vectors, std::vector<double>& wHost, int* info) {
  *info = 0;
  const long ld = n;
  const size_t nn = (size_t)n * n;
  // planes of H (full Hermitian) and S (the factorisation reads the lower triangle)
  DH_TRY(split_complex(s, aDev, lda, Hr.p, Hi.p, ld, n, n));
  DH_TRY(hermitian_from_lower(s, Hr.p, Hi.p, ld, n));
  DH_TRY(split_complex(s, bDev, ldb, Lr.p, Li.p, ld, n, n));
  DH_TRY(fill<double>(s, Vr.p, nn, 0.0));
  DH_TRY(fill<double>(s, Vi.p, nn, 0.0));
  ZPlanes L{Lr.p, Li.p, ld}, V{Vr.p, Vi.p, ld};
  int hinfo = 0;
  {
    // the Cholesky's panel temporary: two n x 64 planes
    ZPlanes ptmp{work.p, work.p + (size_t)n * kZNB, ld};
    DH_TRY(zpotrf_lower(s, L, n, V, ptmp, dinfo.p, &hinfo));
  }
  if (hinfo != 0) {
    *info = n + hinfo;
    return 0;
  }
  // the joined blocks' products: up to n x n per plane at leading dimension n
  DH_TRY(ztrtri_from_diag(s, L.c(), n, V, ZPlanes{work.p, work.p + nn, ld}));
  // A = Linv H Linv^H (lower triangle)
  ZPlanes T{Ar.p, Ai.p, ld};
  ZPlanes H{Hr.p, Hi.p, ld};
  DH_TRY(zgemm(s, 'N', 'N', n, n, n, 1.0, V.c(), H.c(), 0.0, T, false, kKALower));
  // the product T Linv^H goes to H's planes (free now)
  DH_TRY(zgemm(s, 'N', 'C', n, n, n, 1.0, T.c(), V.c(), 0.0, H, true, kKBUpper));
  // real embedding (2n), tridiagonalisation and divide and conquer
  const int n2 = 2 * n;
  hipLaunchKernelGGL(embed_lower_kernel, dim3(cdiv((long)n2 * n2, 256)), dim3(256), 0, s, Hr.p,
                     Hi.p, ld, n, E2.p, (long)n2);
  DH_LAUNCH_CHECK();
  DH_TRY(sytrd_lower(s, E2.p, n2, n2, d2.p, e2.p, tau2.p, work.p, workSize));
  std::vector<double> dh(n2), eh(n2);
  DH_HIP(hipMemcpyAsync(dh.data(), d2.p, n2 * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(eh.data(), e2.p, (n2 - 1) * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  if (!sws) sws.reset(stedc_ws_new(), stedc_ws_delete);
  int dcInfo = 0;
  DH_TRY(stedc(s, n2, dh.data(), eh.data(), w2.p, Z2.p, n2, work.p, workSize, &dcInfo,
               sws.get()));
  if (dcInfo != 0) {
    *info = dcInfo;
    return 0;
  }
  std::vector<double> w2h(n2);
  DH_HIP(hipMemcpyAsync(w2h.data(), w2.p, n2 * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  wHost.resize(n);
  for (int j = 0; j < n; ++j) wHost[j] = w2h[2 * j];
  if (!vectors) return 0;
  // back-transformation of all embedded vectors, one complex vector per level
  DH_TRY(ormtr_lower(s, E2.p, n2, n2, tau2.p, Z2.p, n2, n2, work.p, workSize));
  ZPlanes Y{Ar.p, Ai.p, ld};   // T's planes are free once A is formed
  hipLaunchKernelGGL(pair_gather_kernel, dim3(cdiv((long)n * n, 256)), dim3(256), 0, s, Z2.p,
                     (long)n2, n, Y.r, Y.i, ld);
  DH_LAUNCH_CHECK();
  std::vector<int> cl;
  for (int j = 0; j < n;) {
    int k = j + 1;
    while (k < n && std::fabs(wHost[k] - wHost[k - 1]) <=
                        1e-8 * std::max(1.0, std::fabs(wHost[k])))
      ++k;
    if (k - j > 1) {
      cl.push_back(j);
      cl.push_back(k - j);
    }
    j = k;
  }
  nClusters = (int)cl.size() / 2;
  if (nClusters > 0) {
    DH_TRY(clusterDev.alloc(cl.size()));
    DH_HIP(hipMemcpyAsync(clusterDev.p, cl.data(), cl.size() * sizeof(int),
                          hipMemcpyHostToDevice, s));
    DH_TRY(fill<int>(s, dinfo.p + 1, 1, 0));
    DH_CHECK((size_t)nClusters * 2 * n <= workSize, "hegvd: cluster workspace too small");
    hipLaunchKernelGGL(cluster_gs_kernel, dim3(nClusters), dim3(256), 0, s, Z2.p, (long)n2, n,
                       clusterDev.p, Y.r, Y.i, ld, work.p, dinfo.p + 1);
    DH_LAUNCH_CHECK();
    int failed = 0;
    DH_HIP(hipMemcpyAsync(&failed, dinfo.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    DH_HIP(hipStreamSynchronize(s));
    DH_CHECK(failed == 0, "hegvd: a degenerate cluster did not yield independent vectors");
  }
  // C = Linv^H Y into H's planes
  DH_TRY(zgemm(s, 'C', 'N', n, n, n, 1.0, V.c(), Y.c(), 0.0, H, false, kKAUpper));
  return 0;
}

int hermitian_from_lower(hipStream_t s, double* re, double* im, long ld, int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(hermitian_from_lower_kernel, dim3(cdiv((long)n * n, 256)), dim3(256), 0, s,
                     re, im, ld, n);
  DH_LAUNCH_CHECK();
  return 0;
}

}  // namespace dftbhip
