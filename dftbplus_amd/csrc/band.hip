// Two-stage reduction of a symmetric matrix to tridiagonal form, and its back-transformation.
//
// Stage 1 (dense -> band of width kBW): for each panel of kBW columns a Householder QR of the
// part below the band (column kernels), then the two-sided update of the trailing matrix with
// MFMA GEMMs (A22 -= V W^T + W V^T, W = A22 V T - 1/2 V T^T V^T A22 V T). The matrix is kept
// in full symmetric storage so that every product is a plain GEMM.
//
// Stage 2 (band -> tridiagonal): bulge chasing. Sweep s annihilates column s of the band with a
// reflector on rows s+1..s+kBW, then chases the bulge it creates down the band, one kBW x kBW
// block step at a time. Sweeps run concurrently on the workgroups of one cooperative launch:
// step k of sweep s starts once sweep s-1 has finished step k+2 (the first step that does not
// touch rows or columns of step k of sweep s), published through per-sweep progress counters.
// Every wait is bounded; a wait that runs out sets an abort flag that ends all workgroups.
//
// Back-transformation: the stage-2 reflectors of kBW consecutive sweeps at the same step index
// form one block reflector (2 kBW - 1 rows); sweep groups are applied from the last to the first
// and, inside a group, step blocks in ascending order (reflectors of later sweeps and earlier
// steps commute with the ones they are moved past). Stage-1 reflectors are applied with the
// blocked routine of the one-stage path (unit element kBW rows below the column).
//
// LAPACK's dsytrd_sy2sb / dsytrd_sb2st / dormtr pair plays this role in the reference's
// two-stage solvers (src/dftbp/math/eigensolver.F90 hegvr/hegvd routes through LAPACK).
#include <algorithm>
#include <type_traits>

#include "common.h"
#include "mfma.h"

namespace dftbhip {

constexpr int kBW = 32;            // bandwidth of the intermediate band matrix
constexpr int kLdab = 2 * kBW;     // rows of the band storage: band plus bulge
constexpr int kQ2Cols = 32;        // columns of Z per workgroup in the stage-2 back-transform

template <typename T>
int ormqr_shift(hipStream_t s, const T* a, long lda, int n, const T* tau, T* z, long ldz,
                int ncols, void* work, size_t workBytes, int shift, int nref);

// Householder reflector (I - tau v v^T) x = beta e1, v = [1, x(2:) * scal] (LAPACK dlarfg
// without the rescaling of tiny norms)
__device__ __forceinline__ void hh_reflector(double alpha, double xnrm2, double& beta, double& tau,
                                             double& scal) {
  if (xnrm2 == 0.0) {
    beta = alpha;
    tau = 0.0;
    scal = 0.0;
    return;
  }
  beta = -copysign(sqrt(alpha * alpha + xnrm2), alpha);
  tau = (beta - alpha) / beta;
  scal = 1.0 / (alpha - beta);
}

__device__ __forceinline__ double wsum64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__device__ __forceinline__ double block_sum256(double v, double* sh) {
  v = wsum64(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) t = sh[0] + sh[1] + sh[2] + sh[3];
  return t;
}

// number of bulge-chasing steps of sweep s (steps k with s + 1 + k kBW <= n - 2)
__host__ __device__ __forceinline__ int chase_steps(int n, int s) {
  return (s + 1 <= n - 2) ? (n - 3 - s) / kBW + 1 : 0;
}

// ------------------------------------------------------------------ stage 1: panel QR kernels
// The panel P (m x kBW) starts at row j0 + kBW, column j0 of the matrix.

template <typename T>
__global__ __launch_bounds__(256) void qr_norm0_kernel(const T* P, long lda, int m, double* nrm2) {
  __shared__ double sh[4];
  const int r = 1 + blockIdx.x * 256 + threadIdx.x;
  double x = 0.0;
  if (r < m) {
    x = (double)P[r];
    x *= x;
  }
  const double t = block_sum256(x, sh);
  if (threadIdx.x == 0 && t != 0.0) atomicAdd(&nrm2[0], t);
}

// reflector of column i from alpha = P(i,i) and nrm2[i] = ||P(i+1:m, i)||^2, and the products
// dots(i, j) = v^T P(i:m, j) for the columns j > i of the panel (partial sums per block)
template <typename T>
__global__ __launch_bounds__(256) void qr_dots_kernel(const T* P, long lda, int m, int i,
                                                      const double* nrm2, double* dots,
                                                      double* refl, T* tauOut) {
  __shared__ double red[4][kBW];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double beta, tau, scal;
  hh_reflector((double)P[i + (long)i * lda], nrm2[i], beta, tau, scal);
  if (blockIdx.x == 0 && tid == 0) {
    refl[3 * i] = beta;
    refl[3 * i + 1] = tau;
    refl[3 * i + 2] = scal;
    tauOut[i] = (T)tau;
  }
  const int r = i + blockIdx.x * 256 + tid;
  double v = 0.0;
  if (r < m) v = (r == i) ? 1.0 : (double)P[r + (long)i * lda] * scal;
  for (int j = i + 1; j < kBW; ++j) {
    double x = (r < m) ? v * (double)P[r + (long)j * lda] : 0.0;
    x = wsum64(x);
    if (lane == 0) red[wv][j] = x;
  }
  __syncthreads();
  if (tid > i && tid < kBW) {
    const double t = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
    if (t != 0.0) atomicAdd(&dots[i * kBW + tid], t);
  }
}

// P(i:m, j) -= tau v dots(i, j) for j > i; v and beta into column i; ||P(i+2:m, i+1)||^2
template <typename T>
__global__ __launch_bounds__(256) void qr_update_kernel(T* P, long lda, int m, int i,
                                                        const double* refl, const double* dots,
                                                        double* nrm2) {
  __shared__ double sh[4];
  const double beta = refl[3 * i], tau = refl[3 * i + 1], scal = refl[3 * i + 2];
  const int r = i + blockIdx.x * 256 + threadIdx.x;
  double sq = 0.0;
  if (r < m) {
    const double v = (r == i) ? 1.0 : (double)P[r + (long)i * lda] * scal;
    for (int j = i + 1; j < kBW; ++j) {
      const T x = (T)((double)P[r + (long)j * lda] - tau * v * dots[i * kBW + j]);
      P[r + (long)j * lda] = x;
      if (j == i + 1 && r >= i + 2) sq = (double)x * (double)x;
    }
    P[r + (long)i * lda] = (T)((r == i) ? beta : v);
  }
  const double t = block_sum256(sq, sh);
  if (threadIdx.x == 0 && t != 0.0 && i + 1 < kBW) atomicAdd(&nrm2[i + 1], t);
}

// explicit V (m x kr, unit lower trapezoidal) into V1 and V2 (two copies, leading dim m)
template <typename T>
__global__ void qr_form_v_kernel(const T* P, long lda, int m, int kr, T* V1, T* V2) {
  const int c = blockIdx.y;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m || c >= kr) return;
  T v;
  if (r < c) v = T(0);
  else if (r == c) v = T(1);
  else v = P[r + (long)c * lda];
  V1[r + (long)c * m] = v;
  V2[r + (long)c * m] = v;
}

// T(i,i) = tau_i, T(0:i,i) = -tau_i T(0:i,0:i) G(0:i,i) with G = V^T V (ld kBW)
template <typename T>
__global__ __launch_bounds__(64) void larft32_kernel(const T* G, const T* tau, int kr, T* Tout) {
  __shared__ double Ts[kBW][kBW + 1];
  const int t = threadIdx.x;
  for (int e = t; e < kBW * kBW; e += 64) Ts[e % kBW][e / kBW] = 0.0;
  __syncthreads();
  for (int i = 0; i < kr; ++i) {
    const double ti = (double)tau[i];
    if (t < i) {
      double acc = 0.0;
      for (int k = t; k < i; ++k) acc += Ts[t][k] * (double)G[k + i * kBW];
      Ts[t][i] = -ti * acc;
    }
    if (t == 0) Ts[i][i] = ti;
    __syncthreads();
  }
  for (int e = t; e < kBW * kBW; e += 64) Tout[e] = (T)Ts[e % kBW][e / kBW];
}

// lower band (width kBW) of the dense matrix into the band storage, bulge rows zeroed
template <typename T>
__global__ void band_extract_kernel(const T* A, long lda, int n, T* AB) {
  const int j = blockIdx.x;
  const int d = threadIdx.x;   // 0 .. kLdab-1
  if (j >= n || d >= kLdab) return;
  const int i = j + d;
  T v = T(0);
  if (d <= kBW && i < n) v = A[i + (long)j * lda];
  AB[d + (long)j * kLdab] = v;
}

template <typename T>
__global__ void band_diag_kernel(const T* AB, int n, T* d, T* e) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  d[j] = AB[(long)j * kLdab];
  if (j < n - 1) e[j] = AB[1 + (long)j * kLdab];
}

// ------------------------------------------------------------------ stage 2: bulge chasing

template <typename T>
__device__ __forceinline__ T& band_at(T* AB, int i, int j) {
  return AB[(i - j) + (long)j * kLdab];
}

template <typename T>
__global__ __launch_bounds__(256) void sb2st_kernel(T* AB, int n, int nsweep, int kmax, T* V2,
                                                    T* tau2, int* prog, int* abortFlag,
                                                    long spinLimit) {
  constexpr int b = kBW;
  __shared__ double O[b][b + 1];    // rows q0..q1-1, columns col..col+ncO-1
  __shared__ double D[b][b + 1];    // rows and columns q0..q1-1 (full symmetric)
  __shared__ double Bk[b][b + 1];   // rows q1..r1-1, columns q0..q1-1
  __shared__ double vs[b], ws[b], ys[b];
  __shared__ double sc[4];
  __shared__ int stop;
  const int tid = threadIdx.x;
  for (int s = blockIdx.x; s < nsweep; s += gridDim.x) {
    const int ns = chase_steps(n, s);
    for (int k = 0; k < ns; ++k) {
      if (s > 0) {
        if (tid == 0) {
          const int need = min(k + 3, chase_steps(n, s - 1));
          long spins = 0;
          int ab = 0;
          // relaxed polls (an acquire load would invalidate the L2 on every poll); one
          // acquire fence after the wait
          while (__hip_atomic_load(&prog[s - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                 need) {
            ab = __hip_atomic_load(abortFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (ab != 0) break;
            if (++spins > spinLimit) {
              ab = 1;
              __hip_atomic_store(abortFlag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              break;
            }
            __builtin_amdgcn_s_sleep(2);
          }
          stop = ab;
        }
        __syncthreads();
        if (stop != 0) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      const int q0 = s + 1 + k * b;
      const int q1 = min(q0 + b, n);
      const int L = q1 - q0;
      const int col = (k == 0) ? s : q0 - b;
      const int ncO = (k == 0) ? 1 : b;
      const int r1 = min(q1 + b, n);
      const int LB = r1 - q1;
      for (int e = tid; e < L * ncO; e += 256) {
        const int i = e % L, j = e / L;
        O[i][j] = (double)band_at(AB, q0 + i, col + j);
      }
      for (int e = tid; e < L * L; e += 256) {
        const int i = e % L, j = e / L;
        if (i >= j) {
          const double x = (double)band_at(AB, q0 + i, q0 + j);
          D[i][j] = x;
          D[j][i] = x;
        }
      }
      for (int e = tid; e < LB * L; e += 256) {
        const int i = e % LB, j = e / LB;
        Bk[i][j] = (double)band_at(AB, q1 + i, q0 + j);
      }
      __syncthreads();
      // reflector from the first column of O
      if (tid < 64) {
        const double x = (tid >= 1 && tid < L) ? O[tid][0] : 0.0;
        const double xn2 = wsum64(x * x);
        if (tid == 0) {
          double beta, tau, scal;
          hh_reflector(O[0][0], xn2, beta, tau, scal);
          sc[0] = beta;
          sc[1] = tau;
          sc[2] = scal;
        }
      }
      __syncthreads();
      const double beta = sc[0], tau = sc[1], scal = sc[2];
      if (tid < L) vs[tid] = (tid == 0) ? 1.0 : O[tid][0] * scal;
      __syncthreads();
      const long vo = ((long)s * kmax + k) * b;
      if (tid < b) V2[vo + tid] = (T)(tid < L ? vs[tid] : 0.0);
      if (tid == 0) tau2[(long)s * kmax + k] = (T)tau;
      // left application to the remaining columns of O; its first column becomes beta e1
      if (tid >= 1 && tid < ncO) {
        double y = 0.0;
        for (int i = 0; i < L; ++i) y += vs[i] * O[i][tid];
        ys[tid] = tau * y;
      }
      // w = tau D v
      if (tid >= 64 && tid < 64 + L) {
        const int i = tid - 64;
        double w = 0.0;
        for (int j = 0; j < L; ++j) w += D[i][j] * vs[j];
        ws[i] = tau * w;
      }
      __syncthreads();
      for (int e = tid; e < L * ncO; e += 256) {
        const int i = e % L, j = e / L;
        if (j == 0) O[i][0] = (i == 0) ? beta : 0.0;
        else O[i][j] -= vs[i] * ys[j];
      }
      if (tid < 64) {
        const double p = (tid < L) ? vs[tid] * ws[tid] : 0.0;
        const double dot = wsum64(p);
        if (tid == 0) sc[3] = dot;
      }
      __syncthreads();
      const double half = 0.5 * tau * sc[3];
      if (tid < L) ws[tid] -= half * vs[tid];
      // y = Bk v (right application below the diagonal block)
      if (tid >= 64 && tid < 64 + LB) {
        const int i = tid - 64;
        double y = 0.0;
        for (int j = 0; j < L; ++j) y += Bk[i][j] * vs[j];
        ys[i] = tau * y;
      }
      __syncthreads();
      for (int e = tid; e < L * L; e += 256) {
        const int i = e % L, j = e / L;
        D[i][j] -= vs[i] * ws[j] + ws[i] * vs[j];
      }
      for (int e = tid; e < LB * L; e += 256) {
        const int i = e % LB, j = e / LB;
        Bk[i][j] -= ys[i] * vs[j];
      }
      __syncthreads();
      for (int e = tid; e < L * ncO; e += 256) {
        const int i = e % L, j = e / L;
        band_at(AB, q0 + i, col + j) = (T)O[i][j];
      }
      for (int e = tid; e < L * L; e += 256) {
        const int i = e % L, j = e / L;
        if (i >= j) band_at(AB, q0 + i, q0 + j) = (T)D[i][j];
      }
      for (int e = tid; e < LB * L; e += 256) {
        const int i = e % LB, j = e / LB;
        band_at(AB, q1 + i, q0 + j) = (T)Bk[i][j];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __syncthreads();
      if (tid == 0) __hip_atomic_store(&prog[s], k + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ------------------------------------------------------------------ stage-2 back-transformation
// Block (J, k): sweeps s0 = J kBW .. s1-1, step k; rows r0 = s0 + 1 + k kBW .. r1-1.

__host__ __device__ __forceinline__ void q2_rows(int n, int nsweep, int J, int k, int& s0, int& s1,
                                                 int& r0, int& r1) {
  s0 = J * kBW;
  s1 = min(s0 + kBW, nsweep);
  r0 = s0 + 1 + k * kBW;
  r1 = min(s1 + k * kBW + kBW, n);
}

// explicit V of block (J, k) into LDS: column t holds the reflector of sweep s0+t at rows t..
template <typename T>
__device__ void q2_load_v(const T* V2, const T* tau2, int n, int nsweep, int kmax, int J, int k,
                          double (*Vs)[kBW + 1], double* taus, int R) {
  int s0, s1, r0, r1;
  q2_rows(n, nsweep, J, k, s0, s1, r0, r1);
  for (int e = threadIdx.x; e < R * kBW; e += blockDim.x) {
    const int i = e % R, t = e / R;
    const int s = s0 + t;
    double v = 0.0;
    if (s < s1 && k < chase_steps(n, s)) {
      const int q0 = s + 1 + k * kBW;
      const int L = min(q0 + kBW, n) - q0;
      const int ii = i - t;
      if (ii >= 0 && ii < L) v = (double)V2[((long)s * kmax + k) * kBW + ii];
    }
    Vs[i][t] = v;
  }
  for (int t = threadIdx.x; t < kBW; t += blockDim.x) {
    const int s = s0 + t;
    taus[t] = (s < s1 && k < chase_steps(n, s)) ? (double)tau2[(long)s * kmax + k] : 0.0;
  }
}

// T of block (J, k) (forward, columnwise, LAPACK dlarft) and its explicit V, zero padded to
// 2 kBW rows, into Vx (column-major 2 kBW x kBW per block)
template <typename T>
__global__ __launch_bounds__(256) void q2_larft_kernel(const T* V2, const T* tau2, int n,
                                                       int nsweep, int kmax, T* T2, T* Vx) {
  __shared__ double Vs[2 * kBW][kBW + 1];
  __shared__ double G[kBW][kBW + 1];
  __shared__ double Ts[kBW][kBW + 1];
  __shared__ double taus[kBW];
  const int J = blockIdx.x, k = blockIdx.y;
  int s0, s1, r0, r1;
  q2_rows(n, nsweep, J, k, s0, s1, r0, r1);
  if (k >= chase_steps(n, s0)) return;
  const int R = r1 - r0;
  q2_load_v(V2, tau2, n, nsweep, kmax, J, k, Vs, taus, R);
  for (int e = threadIdx.x; e < (2 * kBW - R) * kBW; e += 256) {
    const int i = R + e % (2 * kBW - R), t = e / (2 * kBW - R);
    Vs[i][t] = 0.0;
  }
  __syncthreads();
  T* vout = Vx + ((long)J * kmax + k) * 2 * kBW * kBW;
  for (int e = threadIdx.x; e < 2 * kBW * kBW; e += 256) vout[e] = (T)Vs[e % (2 * kBW)][e / (2 * kBW)];
  for (int e = threadIdx.x; e < kBW * kBW; e += 256) {
    const int a = e % kBW, c = e / kBW;
    double g = 0.0;
    for (int i = 0; i < R; ++i) g += Vs[i][a] * Vs[i][c];
    G[a][c] = g;
    Ts[a][c] = 0.0;
  }
  __syncthreads();
  const int t = threadIdx.x;
  for (int i = 0; i < kBW; ++i) {
    if (t < i) {
      double acc = 0.0;
      for (int q = t; q < i; ++q) acc += Ts[t][q] * G[q][i];
      Ts[t][i] = -taus[i] * acc;
    }
    if (t == 0) Ts[i][i] = taus[i];
    __syncthreads();
  }
  T* out = T2 + ((long)J * kmax + k) * kBW * kBW;
  for (int e = threadIdx.x; e < kBW * kBW; e += 256) out[e] = (T)Ts[e % kBW][e / kBW];
}

// Z <- Q2 Z for the columns c0 .. c0 + kQ2Cols - 1 of one workgroup (4 wave64s, MFMA 16x16x4).
// For each block: W = V^T Z (32 x 32), W2 = T W, Z -= V W2 (64 x 32). Inside a sweep group the
// row windows of consecutive step blocks overlap by kBW rows: Z rows live in a 2 kBW-row ring in
// LDS, the half that leaves the window is stored and the half that enters it is loaded.
template <typename T>
__global__ __launch_bounds__(256) void q2_apply_kernel(const T* Vx, const T* T2, int n,
                                                       int nsweep, int kmax, T* Z, long ldz,
                                                       int ncols) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  constexpr int R2 = 2 * kBW;        // 64 rows of a block window
  constexpr int NC = kQ2Cols;        // 32 columns
  __shared__ T Vs[R2][kBW + 1];
  __shared__ T Ts[kBW][kBW + 1];
  __shared__ T Zs[R2][NC + 1];       // ring: logical row i at physical (i + ofs) % R2
  __shared__ T Ws[kBW][NC + 1];
  __shared__ T W2s[kBW][NC + 1];
  const int c0 = blockIdx.x * NC;
  const int nc = min(NC, ncols - c0);
  if (nc <= 0) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nJ = (nsweep + kBW - 1) / kBW;
  auto zload = [&](int grow0, int lrow0, int nrows, int ofs) {
    for (int e = tid; e < nrows * NC; e += 256) {
      const int i = e % nrows, c = e / nrows;
      const int gr = grow0 + i;
      Zs[(lrow0 + i + ofs) % R2][c] = (gr < n && c < nc) ? Z[gr + (long)(c0 + c) * ldz] : T(0);
    }
  };
  auto zstore = [&](int grow0, int lrow0, int nrows, int ofs) {
    for (int e = tid; e < nrows * NC; e += 256) {
      const int i = e % nrows, c = e / nrows;
      const int gr = grow0 + i;
      if (gr < n && c < nc) Z[gr + (long)(c0 + c) * ldz] = Zs[(lrow0 + i + ofs) % R2][c];
    }
  };
  for (int J = nJ - 1; J >= 0; --J) {
    const int kn = chase_steps(n, J * kBW);
    for (int k = 0; k < kn; ++k) {
      int s0, s1, r0, r1;
      q2_rows(n, nsweep, J, k, s0, s1, r0, r1);
      const int ofs = (k & 1) * kBW;
      if (k == 0) zload(r0, 0, R2, ofs);
      else zload(r0 + kBW, kBW, kBW, ofs);
      const T* vb = Vx + ((long)J * kmax + k) * R2 * kBW;
      for (int e = tid; e < R2 * kBW; e += 256) Vs[e % R2][e / R2] = vb[e];
      const T* tb = T2 + ((long)J * kmax + k) * kBW * kBW;
      for (int e = tid; e < kBW * kBW; e += 256) Ts[e % kBW][e / kBW] = tb[e];
      __syncthreads();
      // W = V^T Z: wave wv owns the 16 x 16 tile (ti, tj)
      {
        const int ti = wv & 1, tj = wv >> 1;
        acc_t acc = {0, 0, 0, 0};
        for (int k0 = 0; k0 < R2; k0 += 4) {
          const T a = Vs[k0 + lk][ti * 16 + lr];
          const T b = Zs[(k0 + lk + ofs) % R2][tj * 16 + lr];
          acc = M::op(a, b, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) Ws[ti * 16 + M::row(lane, r)][tj * 16 + lr] = acc[r];
      }
      __syncthreads();
      // W2 = T W
      {
        const int ti = wv & 1, tj = wv >> 1;
        acc_t acc = {0, 0, 0, 0};
        for (int p0 = 0; p0 < kBW; p0 += 4) {
          const T a = Ts[ti * 16 + lr][p0 + lk];
          const T b = Ws[p0 + lk][tj * 16 + lr];
          acc = M::op(a, b, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) W2s[ti * 16 + M::row(lane, r)][tj * 16 + lr] = acc[r];
      }
      __syncthreads();
      // Z -= V W2: wave wv owns rows wv*16 .. wv*16+15, both column tiles
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        acc_t acc = {0, 0, 0, 0};
        for (int a0 = 0; a0 < kBW; a0 += 4) {
          const T a = Vs[wv * 16 + lr][a0 + lk];
          const T b = W2s[a0 + lk][ct * 16 + lr];
          acc = M::op(a, b, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = (wv * 16 + M::row(lane, r) + ofs) % R2;
          Zs[row][ct * 16 + lr] -= acc[r];
        }
      }
      __syncthreads();
      // the first kBW rows leave the window (all of it after the last step of the group)
      if (k + 1 < kn) zstore(r0, 0, kBW, ofs);
      else zstore(r0, 0, R2, ofs);
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ host side

template <typename T>
struct BandWs {
  T* AB;          // kLdab x n
  T* V2;          // nsweep x kmax x kBW
  T* tau2;        // nsweep x kmax
  T* T2;          // nJ x kmax x kBW x kBW
  T* Vx;          // nJ x kmax x 2kBW x kBW  explicit block reflectors
  T* tau1;        // n
  T* P1;          // n x 2kBW   [V W]
  T* P2;          // n x 2kBW   [W V]
  T* X;           // n x kBW
  T* G;           // kBW x kBW
  T* Tm;          // kBW x kBW
  T* Zs;          // kBW x kBW
  T* U;           // kBW x kBW
  double* dots;   // kBW x kBW
  double* nrm2;   // kBW
  double* refl;   // 3 kBW
  int* prog;      // nsweep
  int* abortFlag; // 1
};

// Stage 2 on one workgroup of an ordinary launch instead of the cooperative grid: the sweeps
// then run in order, so every progress wait is already met when it is polled. Taken when the
// device refuses the cooperative launch (while another process holds the device's cooperative
// launch resource, hipLaunchCooperativeKernel returns hipErrorIllegalState) or when forced.
static bool g_sb2stSingle = false;
int set_sb2st_single(bool on) {
  g_sb2stSingle = on;
  return 0;
}

template <typename T>
static size_t band_layout(int n, BandWs<T>* ws, char* base) {
  size_t off = 0;
  auto take = [&](auto** p, size_t bytes) {
    if (ws) *p = reinterpret_cast<std::remove_reference_t<decltype(**p)>*>(base + off);
    off += (bytes + 255) / 256 * 256;
  };
  BandWs<T> dummy;
  BandWs<T>* w = ws ? ws : &dummy;
  const size_t nsweep = std::max(n - 2, 1);
  const size_t kmax = std::max(chase_steps(n, 0), 1);
  const size_t nJ = (nsweep + kBW - 1) / kBW;
  take(&w->AB, sizeof(T) * kLdab * std::max(n, 1));
  take(&w->V2, sizeof(T) * nsweep * kmax * kBW);
  take(&w->tau2, sizeof(T) * nsweep * kmax);
  take(&w->T2, sizeof(T) * nJ * kmax * kBW * kBW);
  take(&w->Vx, sizeof(T) * nJ * kmax * 2 * kBW * kBW);
  take(&w->tau1, sizeof(T) * std::max(n, 1));
  take(&w->P1, sizeof(T) * std::max(n, 1) * 2 * kBW);
  take(&w->P2, sizeof(T) * std::max(n, 1) * 2 * kBW);
  take(&w->X, sizeof(T) * std::max(n, 1) * kBW);
  take(&w->G, sizeof(T) * kBW * kBW);
  take(&w->Tm, sizeof(T) * kBW * kBW);
  take(&w->Zs, sizeof(T) * kBW * kBW);
  take(&w->U, sizeof(T) * kBW * kBW);
  take(&w->dots, sizeof(double) * kBW * kBW);
  take(&w->nrm2, sizeof(double) * kBW);
  take(&w->refl, sizeof(double) * 3 * kBW);
  take(&w->prog, sizeof(int) * nsweep);
  take(&w->abortFlag, sizeof(int));
  return off;
}

size_t band_work_bytes(int n, int elemSize) {
  return elemSize == 8 ? band_layout<double>(n, nullptr, nullptr)
                       : band_layout<float>(n, nullptr, nullptr);
}

// Dense symmetric a (n x n, both triangles) -> tridiagonal (d, e); the reflectors stay in a
// (stage 1, below the band) and in the workspace (stage 2) for the back-transformation.
template <typename T>
int sytrd_2stage_t(hipStream_t s, T* a, long lda, int n, T* d, T* e, void* work,
                   size_t workBytes) {
  DH_CHECK(workBytes >= band_layout<T>(n, nullptr, nullptr), "sytrd_2stage: workspace too small");
  BandWs<T> ws;
  band_layout<T>(n, &ws, static_cast<char*>(work));
  DH_TRY(fill<T>(s, ws.tau1, std::max(n, 1), T(0)));
  // stage 1
  for (int j0 = 0;; j0 += kBW) {
    const int m = n - j0 - kBW;
    if (m < 2) break;
    const int kr = std::min(kBW, m - 1);
    T* P = a + (j0 + kBW) + (long)j0 * lda;
    DH_TRY(fill<double>(s, ws.dots, kBW * kBW, 0.0));
    DH_TRY(fill<double>(s, ws.nrm2, kBW, 0.0));
    hipLaunchKernelGGL(qr_norm0_kernel<T>, dim3(cdiv(m, 256)), dim3(256), 0, s, P, lda, m,
                       ws.nrm2);
    DH_LAUNCH_CHECK();
    for (int i = 0; i < kr; ++i) {
      const dim3 g(cdiv(m - i, 256));
      hipLaunchKernelGGL(qr_dots_kernel<T>, g, dim3(256), 0, s, P, lda, m, i, ws.nrm2, ws.dots,
                         ws.refl, ws.tau1 + j0);
      DH_LAUNCH_CHECK();
      hipLaunchKernelGGL(qr_update_kernel<T>, g, dim3(256), 0, s, P, lda, m, i, ws.refl, ws.dots,
                         ws.nrm2);
      DH_LAUNCH_CHECK();
    }
    // V into P1[:, 0:kr] and P2[:, kr:2kr]
    T* V = ws.P1;
    T* W = ws.P1 + (size_t)m * kr;
    hipLaunchKernelGGL(qr_form_v_kernel<T>, dim3(cdiv(m, 256), kr), dim3(256), 0, s, P, lda, m,
                       kr, V, ws.P2 + (size_t)m * kr);
    DH_LAUNCH_CHECK();
    const int sk = std::max(1, std::min(32, m / 512));
    DH_TRY(gemm<T>(s, true, false, kr, kr, m, T(1), V, m, V, m, T(0), ws.G, kBW, false, kKFull,
                   nullptr, sk));
    hipLaunchKernelGGL(larft32_kernel<T>, dim3(1), dim3(64), 0, s, ws.G, ws.tau1 + j0, kr, ws.Tm);
    DH_LAUNCH_CHECK();
    T* a22 = a + (j0 + kBW) + (long)(j0 + kBW) * lda;
    DH_TRY(gemm<T>(s, false, false, m, kr, m, T(1), a22, lda, V, m, T(0), ws.X, m, false, kKFull,
                   nullptr, std::max(1, std::min(16, m / 1024))));
    DH_TRY(gemm<T>(s, false, false, m, kr, kr, T(1), ws.X, m, ws.Tm, kBW, T(0), W, m));
    DH_TRY(gemm<T>(s, true, false, kr, kr, m, T(1), V, m, W, m, T(0), ws.Zs, kBW, false, kKFull,
                   nullptr, sk));
    DH_TRY(gemm<T>(s, true, false, kr, kr, kr, T(1), ws.Tm, kBW, ws.Zs, kBW, T(0), ws.U, kBW));
    DH_TRY(gemm<T>(s, false, false, m, kr, kr, T(-0.5), V, m, ws.U, kBW, T(1), W, m));
    DH_TRY(copy2d<T>(s, W, m, ws.P2, m, m, kr));
    DH_TRY(gemm<T>(s, false, true, m, m, 2 * kr, T(-1), ws.P1, m, ws.P2, m, T(1), a22, lda));
  }
  // stage 2
  hipLaunchKernelGGL(band_extract_kernel<T>, dim3(std::max(n, 1)), dim3(kLdab), 0, s, a, lda, n,
                     ws.AB);
  DH_LAUNCH_CHECK();
  const int nsweep = std::max(n - 2, 0);
  const int kmax = std::max(chase_steps(n, 0), 1);
  if (nsweep > 0) {
    DH_TRY(fill<int>(s, ws.prog, nsweep, 0));
    DH_TRY(fill<int>(s, ws.abortFlag, 1, 0));
    int dev = 0, ncu = 0, perCu = 0;
    DH_HIP(hipGetDevice(&dev));
    DH_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    DH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, sb2st_kernel<T>, 256, 0));
    DH_CHECK(perCu >= 1, "sb2st: kernel does not fit on a compute unit");
    const int grid = std::min(nsweep, std::min(ncu * perCu, 2 * ncu));
    T* AB = ws.AB;
    T* V2 = ws.V2;
    T* tau2 = ws.tau2;
    int* prog = ws.prog;
    int* abortFlag = ws.abortFlag;
    int nn = n, nsw = nsweep, km = kmax;
    long spinLimit = 1L << 24;
    void* args[] = {&AB, &nn, &nsw, &km, &V2, &tau2, &prog, &abortFlag, &spinLimit};
    hipError_t err = hipErrorIllegalState;
    if (!g_sb2stSingle)
      err = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(sb2st_kernel<T>),
                                       dim3(grid), dim3(256), args, 0, s);
    if (err == hipErrorIllegalState || err == hipErrorCooperativeLaunchTooLarge ||
        err == hipErrorNotSupported) {
      (void)hipGetLastError();   // the refused launch ran nothing; prog is still zero
      hipLaunchKernelGGL(sb2st_kernel<T>, dim3(1), dim3(256), 0, s, AB, nn, nsw, km, V2, tau2,
                         prog, abortFlag, spinLimit);
      DH_LAUNCH_CHECK();
    } else {
      DH_HIP(err);
    }
    int hAbort = 0;
    DH_HIP(hipMemcpyAsync(&hAbort, ws.abortFlag, sizeof(int), hipMemcpyDeviceToHost, s));
    DH_HIP(hipStreamSynchronize(s));
    DH_CHECK(hAbort == 0, "sb2st: a progress wait ran out (bulge chasing aborted)");
    const int nJ = (nsweep + kBW - 1) / kBW;
    hipLaunchKernelGGL(q2_larft_kernel<T>, dim3(nJ, kmax), dim3(256), 0, s, ws.V2, ws.tau2, n,
                       nsweep, kmax, ws.T2, ws.Vx);
    DH_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(band_diag_kernel<T>, dim3(cdiv(std::max(n, 1), 256)), dim3(256), 0, s, ws.AB,
                     n, d, e);
  DH_LAUNCH_CHECK();
  return 0;
}

// Z (n x ncols) <- Q Z with Q = Q1 Q2 of sytrd_2stage_t
template <typename T>
int ormtr_2stage_t(hipStream_t s, const T* a, long lda, int n, T* z, long ldz, int ncols,
                   void* work, size_t workBytes, void* work2, size_t work2Bytes) {
  DH_CHECK(workBytes >= band_layout<T>(n, nullptr, nullptr), "ormtr_2stage: workspace too small");
  if (ncols <= 0 || n <= 2) return 0;
  BandWs<T> ws;
  band_layout<T>(n, &ws, static_cast<char*>(work));
  const int nsweep = n - 2;
  const int kmax = std::max(chase_steps(n, 0), 1);
  hipLaunchKernelGGL(q2_apply_kernel<T>, dim3(cdiv(ncols, kQ2Cols)), dim3(256), 0, s, ws.Vx,
                     ws.T2, n, nsweep, kmax, z, ldz, ncols);
  DH_LAUNCH_CHECK();
  const int nref = n - kBW - 1;
  if (nref > 0)
    DH_TRY(ormqr_shift<T>(s, a, lda, n, ws.tau1, z, ldz, ncols, work2, work2Bytes, kBW, nref));
  return 0;
}

int sytrd_2stage(hipStream_t s, double* a, long lda, int n, double* d, double* e, void* work,
                 size_t workBytes) {
  return sytrd_2stage_t<double>(s, a, lda, n, d, e, work, workBytes);
}
int sytrd_2stage_f32(hipStream_t s, float* a, long lda, int n, float* d, float* e, void* work,
                     size_t workBytes) {
  return sytrd_2stage_t<float>(s, a, lda, n, d, e, work, workBytes);
}
int ormtr_2stage(hipStream_t s, const double* a, long lda, int n, double* z, long ldz, int ncols,
                 void* work, size_t workBytes, void* work2, size_t work2Bytes) {
  return ormtr_2stage_t<double>(s, a, lda, n, z, ldz, ncols, work, workBytes, work2, work2Bytes);
}
int ormtr_2stage_f32(hipStream_t s, const float* a, long lda, int n, float* z, long ldz,
                     int ncols, void* work, size_t workBytes, void* work2, size_t work2Bytes) {
  return ormtr_2stage_t<float>(s, a, lda, n, z, ldz, ncols, work, workBytes, work2, work2Bytes);
}

}  // namespace dftbhip
