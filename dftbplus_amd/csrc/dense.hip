// Dense helpers: fills/copies, blocked Cholesky (potrf, lower) and the inverse of the Cholesky
// factor (trtri, lower) built on the MFMA GEMM core.
//
// Cholesky: right-looking, 64-wide panels. The 64x64 diagonal block is factorised inside one
// workgroup in LDS, which also forms its inverse; the panel below is L21 = A21 * inv(L11)^T and
// the trailing update A22 -= L21 L21^T is a lower-output GEMM. The inverse of the whole factor is
// assembled by recursive doubling: inv([L11 0; L21 L22]) = [X11 0; -X22 L21 X11, X22], one
// batched GEMM pair per level.
#include "common.h"

namespace dftbhip {

template <typename T>
__global__ void fill_kernel(T* p, size_t n, T v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename T>
__global__ void fill2d_kernel(T* p, long ld, int m, int n, T v) {
  const int j = blockIdx.y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
    p[i + (long)j * ld] = v;
}

template <typename Tin, typename Tout>
__global__ void convert2d_kernel(const Tin* src, long lds, Tout* dst, long ldd, int m, int n) {
  for (int j = blockIdx.y; j < n; j += gridDim.y)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
      dst[i + (long)j * ldd] = (Tout)src[i + (long)j * lds];
}

template <typename T>
int fill(hipStream_t s, T* p, size_t n, T v) {
  if (n == 0) return 0;
  unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(fill_kernel<T>, dim3(g), dim3(256), 0, s, p, n, v);
  DH_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int fill2d(hipStream_t s, T* p, long ld, int m, int n, T v) {
  if (m <= 0 || n <= 0) return 0;
  if (ld == m) return fill<T>(s, p, (size_t)m * n, v);
  hipLaunchKernelGGL(fill2d_kernel<T>, dim3(cdiv(m, 256), n), dim3(256), 0, s, p, ld, m, n, v);
  DH_LAUNCH_CHECK();
  return 0;
}

template <typename Tin, typename Tout>
int convert2d(hipStream_t s, const Tin* src, long lds, Tout* dst, long ldd, int m, int n) {
  if (m <= 0 || n <= 0) return 0;
  dim3 grid(std::min(cdiv(m, 256), 64u), std::min(n, 65535));
  hipLaunchKernelGGL((convert2d_kernel<Tin, Tout>), grid, dim3(256), 0, s, src, lds, dst, ldd, m,
                     n);
  DH_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int copy2d(hipStream_t s, const T* src, long lds, T* dst, long ldd, int m, int n) {
  return convert2d<T, T>(s, src, lds, dst, ldd, m, n);
}

template <typename T, bool kSqrt>
__global__ void scale_cols_kernel(const T* src, long lds, const T* w, T* dst, long ldd, int m,
                                  int n) {
  for (int j = blockIdx.y; j < n; j += gridDim.y) {
    const T f = kSqrt ? sqrt(w[j]) : w[j];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
      dst[i + (long)j * ldd] = src[i + (long)j * lds] * f;
  }
}

// dst(:, j) = src(:, j) * sqrt(w(j)), w >= 0
template <typename T>
int scale_cols_sqrt(hipStream_t s, const T* src, long lds, const T* w, T* dst, long ldd, int m,
                    int n) {
  if (m <= 0 || n <= 0) return 0;
  dim3 grid(std::min(cdiv(m, 256), 64u), std::min(n, 65535));
  hipLaunchKernelGGL((scale_cols_kernel<T, true>), grid, dim3(256), 0, s, src, lds, w, dst, ldd,
                     m, n);
  DH_LAUNCH_CHECK();
  return 0;
}

// dst(:, j) = src(:, j) * w(j)
template <typename T>
int scale_cols(hipStream_t s, const T* src, long lds, const T* w, T* dst, long ldd, int m, int n) {
  if (m <= 0 || n <= 0) return 0;
  dim3 grid(std::min(cdiv(m, 256), 64u), std::min(n, 65535));
  hipLaunchKernelGGL((scale_cols_kernel<T, false>), grid, dim3(256), 0, s, src, lds, w, dst, ldd,
                     m, n);
  DH_LAUNCH_CHECK();
  return 0;
}
template int scale_cols<double>(hipStream_t, const double*, long, const double*, double*, long,
                                int, int);
template int scale_cols_sqrt<double>(hipStream_t, const double*, long, const double*, double*,
                                     long, int, int);
template int scale_cols_sqrt<float>(hipStream_t, const float*, long, const float*, float*, long,
                                    int, int);

// a(j, i) = a(i, j) for i > j, 32x32 tiles through LDS.
template <typename T>
__global__ void symmetrize_lower_kernel(T* a, long lda, int n) {
  __shared__ T t[32][33];
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bj > bi) return;
  const int tx = threadIdx.x, ty = threadIdx.y;   // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int i = bi * 32 + tx, j = bj * 32 + k;
    t[k][tx] = (i < n && j < n) ? a[i + (long)j * lda] : T(0);
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int i = bj * 32 + tx, j = bi * 32 + k;   // target (i, j) = source (j, i), i < j
    if (i < n && j < n && i < j) a[i + (long)j * lda] = t[tx][k];
  }
}

__global__ void zero_upper_kernel(double* a, long lda, int n) {
  const int j = blockIdx.y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < j && i < n; i += gridDim.x * blockDim.x)
    a[i + (long)j * lda] = 0.0;
}

template <typename T>
static int symmetrize_lower_t(hipStream_t s, T* a, long lda, int n) {
  if (n <= 1) return 0;
  dim3 grid(cdiv(n, 32), cdiv(n, 32));
  hipLaunchKernelGGL(symmetrize_lower_kernel<T>, grid, dim3(32, 8), 0, s, a, lda, n);
  DH_LAUNCH_CHECK();
  return 0;
}
int symmetrize_lower(hipStream_t s, double* a, long lda, int n) {
  return symmetrize_lower_t<double>(s, a, lda, n);
}
int symmetrize_lower(hipStream_t s, float* a, long lda, int n) {
  return symmetrize_lower_t<float>(s, a, lda, n);
}

// dst(j, i) = src(i, j) for the m x n matrix src, 32x32 tiles through LDS (grid-stride over
// tile columns so that any n fits the grid)
template <typename Tin, typename Tout>
__global__ void transpose2d_kernel(const Tin* src, long lds, Tout* dst, long ldd, int m, int n) {
  __shared__ Tout t[32][33];
  const int tx = threadIdx.x, ty = threadIdx.y;   // 32 x 8
  const int bi = blockIdx.x;
  for (int bj = blockIdx.y; bj * 32 < n; bj += gridDim.y) {
    for (int k = ty; k < 32; k += 8) {
      const int i = bi * 32 + tx, j = bj * 32 + k;
      t[k][tx] = (i < m && j < n) ? (Tout)src[i + (long)j * lds] : Tout(0);
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
      const int j = bj * 32 + tx, i = bi * 32 + k;   // dst(j, i) = src(i, j)
      if (i < m && j < n) dst[j + (long)i * ldd] = t[tx][k];
    }
    __syncthreads();
  }
}

template <typename Tin, typename Tout>
int transpose2d(hipStream_t s, const Tin* src, long lds, Tout* dst, long ldd, int m, int n) {
  if (m <= 0 || n <= 0) return 0;
  dim3 grid(cdiv(m, 32), std::min(cdiv(n, 32), 65535u));
  hipLaunchKernelGGL((transpose2d_kernel<Tin, Tout>), grid, dim3(32, 8), 0, s, src, lds, dst, ldd,
                     m, n);
  DH_LAUNCH_CHECK();
  return 0;
}
// dst(i, j) = a(i, j) + b(i, j), b in FP32 (an FP32 product added to an FP64 matrix)
__global__ void add_f32_kernel(const double* a, long lda, const float* b, long ldb, double* dst,
                               long ldd, int m, int n) {
  for (int j = blockIdx.y; j < n; j += gridDim.y)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
      dst[i + (long)j * ldd] = a[i + (long)j * lda] + (double)b[i + (long)j * ldb];
}

int add_f32(hipStream_t s, const double* a, long lda, const float* b, long ldb, double* dst,
            long ldd, int m, int n) {
  if (m <= 0 || n <= 0) return 0;
  dim3 grid(std::min(cdiv(m, 256), 64u), std::min(n, 65535));
  hipLaunchKernelGGL(add_f32_kernel, grid, dim3(256), 0, s, a, lda, b, ldb, dst, ldd, m, n);
  DH_LAUNCH_CHECK();
  return 0;
}

// FP16 operands of the refinement's products, from the m x n FP64 matrix x (or the residual
// x - diag(lam) y when y is given), each row i scaled by rsc[i] (when given): same(i, j) and
// tr(j, i). 64 x 64 tiles through LDS, so that reads and both writes run along columns.
__global__ void f16_convert_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                   const double* __restrict__ lam, const double* __restrict__ rsc,
                                   long ldx, int m, int n, _Float16* __restrict__ same, long ld1,
                                   _Float16* __restrict__ tr, long ld2) {
  __shared__ float t[64][65];
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int i = i0 + tx;
  const double li = (y && i < m) ? lam[i] : 0.0;
  const double si = (rsc && i < m) ? rsc[i] : 1.0;
  for (int jj = ty; jj < 64; jj += 4) {
    const int j = j0 + jj;
    float v = 0.0f;
    if (i < m && j < n) {
      double d = x[i + (long)j * ldx];
      if (y) d -= li * y[i + (long)j * ldx];
      v = (float)(d * si);
      if (same) same[i + (long)j * ld1] = (_Float16)v;
    }
    t[jj][tx] = v;
  }
  if (!tr) return;
  __syncthreads();
  for (int ii = ty; ii < 64; ii += 4) {
    const int j = j0 + tx, r = i0 + ii;
    if (r < m && j < n) tr[j + (long)r * ld2] = (_Float16)t[tx][ii];
  }
}

int f16_convert(hipStream_t s, const double* x, const double* y, const double* lam,
                const double* rsc, long ldx, int m, int n, _Float16* same, long ld1, _Float16* tr,
                long ld2) {
  if (m <= 0 || n <= 0) return 0;
  DH_CHECK(!y || lam, "f16_convert: the residual needs lam");
  hipLaunchKernelGGL(f16_convert_kernel, dim3(cdiv(m, 64), cdiv(n, 64)), dim3(256), 0, s, x, y,
                     lam, rsc, ldx, m, n, same, ld1, tr, ld2);
  DH_LAUNCH_CHECK();
  return 0;
}

template int transpose2d<double, double>(hipStream_t, const double*, long, double*, long, int, int);
template int transpose2d<float, double>(hipStream_t, const float*, long, double*, long, int, int);

int zero_strict_upper(hipStream_t s, double* a, long lda, int n) {
  if (n <= 1) return 0;
  hipLaunchKernelGGL(zero_upper_kernel, dim3(cdiv(n, 256), n), dim3(256), 0, s, a, lda, n);
  DH_LAUNCH_CHECK();
  return 0;
}

constexpr int kPotrfNB = 64;

// Factor the jb x jb diagonal block at a (lda) in place and write its inverse to linv (ldv).
// dInfo receives (offset + k + 1) for the first non-positive pivot.
__global__ __launch_bounds__(256) void potrf_diag_kernel(double* a, long lda, int jb,
                                                         double* linv, long ldv, int offset,
                                                         int* dInfo) {
  __shared__ double L[kPotrfNB][kPotrfNB + 1];   // L[i][c]
  __shared__ double X[kPotrfNB][kPotrfNB + 1];
  __shared__ int failed;
  const int tid = threadIdx.x;
  if (tid == 0) failed = (*dInfo != 0);
  for (int e = tid; e < jb * jb; e += blockDim.x) {
    const int i = e % jb, c = e / jb;
    L[i][c] = (i >= c) ? a[i + (long)c * lda] : 0.0;
  }
  __syncthreads();
  if (failed) return;
  for (int k = 0; k < jb; ++k) {
    if (tid == 0) {
      const double akk = L[k][k];
      if (!(akk > 0.0)) {
        failed = 1;
        if (*dInfo == 0) *dInfo = offset + k + 1;
      } else {
        L[k][k] = sqrt(akk);
      }
    }
    __syncthreads();
    if (failed) return;
    const double dkk = L[k][k];
    for (int i = k + 1 + tid; i < jb; i += blockDim.x) L[i][k] /= dkk;
    __syncthreads();
    const int m = jb - k - 1;
    for (int e = tid; e < m * m; e += blockDim.x) {
      const int i = k + 1 + e % m, c = k + 1 + e / m;
      if (i >= c) L[i][c] -= L[i][k] * L[c][k];
    }
    __syncthreads();
  }
  // inverse by forward substitution, one column per thread
  if (tid < jb) {
    const int c = tid;
    for (int i = 0; i < c; ++i) X[i][c] = 0.0;
    X[c][c] = 1.0 / L[c][c];
    for (int i = c + 1; i < jb; ++i) {
      double acc = 0.0;
      for (int k = c; k < i; ++k) acc += L[i][k] * X[k][c];
      X[i][c] = -acc / L[i][i];
    }
  }
  __syncthreads();
  for (int e = tid; e < jb * jb; e += blockDim.x) {
    const int i = e % jb, c = e / jb;
    if (i >= c) a[i + (long)c * lda] = L[i][c];
    if (linv) linv[i + (long)c * ldv] = X[i][c];
  }
}

int potrf_lower_impl(hipStream_t s, double* a, long lda, int n, double* linv, long ldv,
                     int* dInfo, int* hInfo) {
  DH_TRY(fill<int>(s, dInfo, 1, 0));
  for (int j = 0; j < n; j += kPotrfNB) {
    const int jb = std::min(kPotrfNB, n - j);
    double* ljj = linv + j + (long)j * ldv;
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(256), 0, s, a + j + (long)j * lda, lda, jb,
                       ljj, ldv, j, dInfo);
    DH_LAUNCH_CHECK();
    const int m = n - j - jb;
    if (m > 0) {
      double* a21 = a + (j + jb) + (long)j * lda;
      double* a22 = a + (j + jb) + (long)(j + jb) * lda;
      DH_TRY(gemm<double>(s, false, true, m, jb, jb, 1.0, a21, lda, ljj, ldv, 0.0, a21, lda));
      DH_TRY(gemm<double>(s, false, true, m, m, jb, -1.0, a21, lda, a21, lda, 1.0, a22, lda,
                          true));
    }
  }
  DH_HIP(hipMemcpyAsync(hInfo, dInfo, sizeof(int), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

// Builds the full inverse in linv whose 64x64 diagonal blocks already hold the block inverses
// (written by the Cholesky); the strict upper part must be zero on entry. work: n*n/2 + 4096.
int trtri_from_diag(hipStream_t s, const double* l, long ldl, double* linv, long ldv, int n,
                    double* work) {
  const int nleaf = (n + kPotrfNB - 1) / kPotrfNB;
  std::vector<GemmDesc<double>> d1, d2;
  DevBuf<GemmDesc<double>> dd;
  for (int p = 1; p < nleaf; p *= 2) {
    d1.clear();
    d2.clear();
    int maxb = 0;
    size_t woff = 0;
    for (int s0 = 0; s0 + p < nleaf; s0 += 2 * p) {
      const int r1 = s0 * kPotrfNB;                       // first block rows/cols
      const int r2 = (s0 + p) * kPotrfNB;                 // second block
      const int n1 = r2 - r1;
      const int n2 = std::min((s0 + 2 * p) * kPotrfNB, n) - r2;
      double* t = work + woff;
      woff += (size_t)n2 * n1;
      // T = L21 * X11   (n2 x n1)
      d1.push_back({n2, n1, n1, l + r2 + (long)r1 * ldl, ldl, linv + r1 + (long)r1 * ldv, ldv, t,
                    (long)n2});
      // X21 = -X22 * T  (n2 x n1)
      d2.push_back({n2, n1, n2, linv + r2 + (long)r2 * ldv, ldv, t, (long)n2,
                    linv + r2 + (long)r1 * ldv, ldv});
      maxb = std::max(maxb, std::max(n1, n2));
    }
    if (d1.empty()) continue;
    DH_TRY(dd.alloc(2 * d1.size()));
    DH_HIP(hipMemcpyAsync(dd.p, d1.data(), d1.size() * sizeof(GemmDesc<double>),
                          hipMemcpyHostToDevice, s));
    DH_HIP(hipMemcpyAsync(dd.p + d1.size(), d2.data(), d2.size() * sizeof(GemmDesc<double>),
                          hipMemcpyHostToDevice, s));
    DH_TRY(gemm_batched<double>(s, false, false, dd.p, (int)d1.size(), maxb, maxb, 1.0, 0.0));
    DH_TRY(gemm_batched<double>(s, false, false, dd.p + d1.size(), (int)d2.size(), maxb, maxb,
                                -1.0, 0.0));
    // descriptors must outlive the launches
    DH_HIP(hipStreamSynchronize(s));
  }
  return 0;
}

// ------------------------------------------------------------------------------------------
// Interleaved complex (re, im pairs, Fortran complex(dp)) <-> split planes
// ------------------------------------------------------------------------------------------
__global__ void split_complex_kernel(const double* z, long ldz, double* re, double* im, long ld,
                                     int m, int n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * n) return;
  const int i = (int)(e % m), j = (int)(e / m);
  const double* src = z + 2 * (i + (long)j * ldz);
  re[i + (long)j * ld] = src[0];
  im[i + (long)j * ld] = src[1];
}

// mode 0: every element; 1: the lower triangle, zeros above it; 2: the lower triangle only
__global__ void merge_complex_kernel(const double* re, const double* im, long ld, double* z,
                                     long ldz, int m, int n, int mode) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * n) return;
  const int i = (int)(e % m), j = (int)(e / m);
  double* dst = z + 2 * (i + (long)j * ldz);
  const bool keep = mode == 0 || i >= j;
  if (!keep && mode == 2) return;
  dst[0] = keep ? re[i + (long)j * ld] : 0.0;
  dst[1] = keep ? im[i + (long)j * ld] : 0.0;
}

int split_complex(hipStream_t s, const double* z, long ldz, double* re, double* im, long ld,
                  int m, int n) {
  if (m <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(split_complex_kernel, dim3(cdiv((long)m * n, 256)), dim3(256), 0, s, z, ldz,
                     re, im, ld, m, n);
  DH_LAUNCH_CHECK();
  return 0;
}

int merge_complex(hipStream_t s, const double* re, const double* im, long ld, double* z, long ldz,
                  int m, int n, int mode) {
  if (m <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(merge_complex_kernel, dim3(cdiv((long)m * n, 256)), dim3(256), 0, s, re, im,
                     ld, z, ldz, m, n, mode);
  DH_LAUNCH_CHECK();
  return 0;
}

template int fill<double>(hipStream_t, double*, size_t, double);
template int fill<float>(hipStream_t, float*, size_t, float);
template int fill<int>(hipStream_t, int*, size_t, int);
template int fill<unsigned>(hipStream_t, unsigned*, size_t, unsigned);
template int fill<long>(hipStream_t, long*, size_t, long);
template int fill<_Float16>(hipStream_t, _Float16*, size_t, _Float16);
template int fill2d<double>(hipStream_t, double*, long, int, int, double);
template int fill2d<float>(hipStream_t, float*, long, int, int, float);
template int copy2d<double>(hipStream_t, const double*, long, double*, long, int, int);
template int copy2d<float>(hipStream_t, const float*, long, float*, long, int, int);
template int convert2d<double, float>(hipStream_t, const double*, long, float*, long, int, int);
template int convert2d<float, double>(hipStream_t, const float*, long, double*, long, int, int);

}  // namespace dftbhip
