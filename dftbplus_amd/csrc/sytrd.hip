// Householder tridiagonalisation of a real symmetric matrix (lower storage) and the
// back-transformation of eigenvectors, one-stage blocked form:
//
//   for each panel of kNB columns:
//     for each column c: update column c with the panel's V W^T + W V^T;
//                        form the reflector (alpha, ||x||) -> (beta, tau, 1/(alpha-beta));
//                        w = tau * (A22 v - V (W^T v) - W (V^T v)),  w += -tau/2 (w.v) v;
//     trailing update A22 -= [V W] [W V]^T on the MFMA GEMM (lower tiles).
//
// The symmetric product A22 v reads only the lower triangle: each workgroup owns a 64-row block
// and a chunk of <= 8 column blocks; row-block partial sums and transposed-tile partial sums go
// to workspaces that the next kernel reduces (no atomics on the vector). This product is the
// HBM-bound part of the eigensolver (it streams the trailing matrix once per column).
#include "common.h"

namespace dftbhip {

constexpr int kNB = 32;      // panel width

// Optional per-launch timing of the symmetric product kernel (HIP events on the launch stream):
// launches, summed duration and the algorithmic bytes (lower triangle of the trailing matrix).
struct SymvProfile {
  bool on = false;
  std::vector<hipEvent_t> ev;
  std::vector<double> bytes;
  size_t used = 0;
  double launches = 0, seconds = 0, algBytes = 0;
};
static SymvProfile g_symvProf;

void symv_profiling(bool on) {
  g_symvProf.on = on;
  g_symvProf.launches = g_symvProf.seconds = g_symvProf.algBytes = 0;
}

void symv_profile_stats(double* out) {
  out[0] = g_symvProf.launches;
  out[1] = g_symvProf.seconds;
  out[2] = g_symvProf.algBytes;
}

static int symv_prof_collect(hipStream_t s) {
  SymvProfile& p = g_symvProf;
  if (!p.on || p.used == 0) return 0;
  DH_HIP(hipStreamSynchronize(s));
  for (size_t i = 0; i < p.used; ++i) {
    float ms = 0.f;
    DH_HIP(hipEventElapsedTime(&ms, p.ev[2 * i], p.ev[2 * i + 1]));
    p.seconds += ms * 1e-3;
    p.algBytes += p.bytes[i];
    p.launches += 1;
  }
  p.used = 0;
  return 0;
}
constexpr int kTB = 64;      // SYMV tile
constexpr int kCH = 8;       // column tiles per SYMV workgroup

struct SytrdWs {
  double* W;       // n x kNB
  double* Y;       // maxChunk x n   row-block partial sums
  double* Z;       // nRB x n        transposed partial sums
  double* vbuf;    // n
  double* nrm2;    // n   ||x||^2 per column
  double* dot;     // n   w.v per column
  double* t12;     // n x 2kNB
  double* scal;    // n
  double* P;       // n x 2kNB
  double* Q;       // n x 2kNB
};

static size_t sytrd_layout(int n, SytrdWs* ws, double* base) {
  const size_t nRB = (n + kTB - 1) / kTB;
  const size_t nCh = (nRB + kCH - 1) / kCH;
  size_t off = 0;
  auto take = [&](double** p, size_t cnt) {
    if (ws) *p = base + off;
    off += (cnt + 63) / 64 * 64;
  };
  SytrdWs dummy;
  SytrdWs* w = ws ? ws : &dummy;
  take(&w->W, (size_t)n * kNB);
  take(&w->Y, nCh * n);
  take(&w->Z, nRB * n);
  take(&w->vbuf, n);
  take(&w->nrm2, n);
  take(&w->dot, n);
  take(&w->t12, (size_t)n * 2 * kNB);
  take(&w->scal, n);
  take(&w->P, (size_t)n * 2 * kNB);
  take(&w->Q, (size_t)n * 2 * kNB);
  return off;
}

size_t sytrd_work_size(int n) { return sytrd_layout(n, nullptr, nullptr); }

__device__ __forceinline__ double block_reduce_sum(double v, double* sh) {
  // 256 threads
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return r;
}

// Reflector constants of column c from alpha = A(c+1,c) and ||x||^2 (LAPACK dlarfg without the
// rescaling loop): beta = -sign(alpha) sqrt(alpha^2 + ||x||^2), tau = (beta - alpha)/beta,
// x scale = 1/(alpha - beta).
__device__ __forceinline__ void reflector(double alpha, double xn2, double& beta, double& tau,
                                          double& scal) {
  if (xn2 == 0.0) {
    beta = alpha;
    tau = 0.0;
    scal = 0.0;
    return;
  }
  const double nrm = sqrt(alpha * alpha + xn2);
  beta = alpha >= 0.0 ? -nrm : nrm;
  tau = (beta - alpha) / beta;
  scal = 1.0 / (alpha - beta);
}

// A(r,c) -= sum_p V(r,p) W(c,p) + W(r,p) V(c,p), rows r >= c; accumulate ||A(c+2:n,c)||^2.
__global__ __launch_bounds__(256) void col_update_kernel(double* a, long lda, int n, int c, int j0,
                                                         const double* W, long ldw,
                                                         double* nrm2) {
  __shared__ double vc[kNB], wc[kNB], sh[4];
  const int i = c - j0;
  if (threadIdx.x < i) {
    vc[threadIdx.x] = a[c + (long)(j0 + threadIdx.x) * lda];
    wc[threadIdx.x] = W[c + (long)threadIdx.x * ldw];
  }
  __syncthreads();
  const int r = c + blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (r < n) {
    double v = a[r + (long)c * lda];
    for (int p = 0; p < i; ++p)
      v -= a[r + (long)(j0 + p) * lda] * wc[p] + W[r + (long)p * ldw] * vc[p];
    a[r + (long)c * lda] = v;
    if (r >= c + 2) sq = v * v;
  }
  const double tot = block_reduce_sum(sq, sh);
  if (threadIdx.x == 0 && tot != 0.0) atomicAdd(&nrm2[c], tot);
}

// Partial sums of y = A22 v (A22 = a(c+1:n, c+1:n), lower) and of t1 = W^T v, t2 = V^T v.
__global__ __launch_bounds__(256) void symv_kernel(const double* a, long lda, int n, int c, int j0,
                                                   const double* W, long ldw, const double* nrm2,
                                                   double* Y, double* Z, double* vbuf,
                                                   double* t12, double* scalOut, double* eOut,
                                                   double* tauOut) {
  __shared__ double T[kTB][kTB + 1];    // T[jj][ii]
  __shared__ double vI[kTB], vJ[kTB];
  __shared__ double part[4][kTB];
  const int m = n - c - 1;
  const int I = blockIdx.x, ch = blockIdx.y;
  const int jb0 = ch * kCH;
  if (jb0 > I) return;
  const int jb1 = min(jb0 + kCH, I + 1);
  const int tid = threadIdx.x;
  const double alpha = a[(c + 1) + (long)c * lda];
  double beta, tau, scal;
  reflector(alpha, nrm2[c], beta, tau, scal);
  if (I == 0 && ch == 0 && tid == 0) {
    scalOut[c] = scal;
    eOut[c] = beta;
    tauOut[c] = tau;
  }
  auto vload = [&](int blk, double* dst) {
    if (tid < kTB) {
      const int k = blk * kTB + tid;
      double v = 0.0;
      if (k < m) v = (k == 0) ? 1.0 : a[(c + 1 + k) + (long)c * lda] * scal;
      dst[tid] = v;
    }
  };
  vload(I, vI);
  const int rows0 = c + 1 + I * kTB;
  // the ch == 0 workgroup of each row block publishes v and the panel products t1, t2
  if (ch == 0) {
    __syncthreads();
    if (tid < kTB && I * kTB + tid < m) vbuf[rows0 + tid] = vI[tid];
    const int i = c - j0;
    // t1[p] = sum_r W(r,p) v(r), t2[p] = sum_r V(r,p) v(r) over this row block
    for (int p = tid / kTB; p < 2 * i; p += 4) {
      const int rr = tid % kTB;
      double prod = 0.0;
      if (I * kTB + rr < m) {
        const int r = rows0 + rr;
        const double x = (p < i) ? W[r + (long)p * ldw] : a[r + (long)(j0 + p - i) * lda];
        prod = x * vI[rr];
      }
      for (int o = 32; o > 0; o >>= 1) prod += __shfl_down(prod, o, 64);
      if ((tid & 63) == 0 && prod != 0.0) atomicAdd(&t12[(long)c * 2 * kNB + p], prod);
    }
  }
  double yacc = 0.0;   // row ii = tid & 63, quarter q = tid >> 6
  const int ii = tid & 63, q = tid >> 6;
  for (int J = jb0; J < jb1; ++J) {
    __syncthreads();
    vload(J, vJ);
    const int cols0 = c + 1 + J * kTB;
#pragma unroll 4
    for (int e = 0; e < 16; ++e) {
      const int jj = q * 16 + e;
      const int gr = I * kTB + ii, gc = J * kTB + jj;
      double v = 0.0;
      if (gr < m && gc < m && (I > J || ii >= jj)) v = a[(rows0 + ii) + (long)(cols0 + jj) * lda];
      T[jj][ii] = v;
    }
    __syncthreads();
    // row sums over this thread's quarter of columns
#pragma unroll
    for (int e = 0; e < 16; ++e) yacc += T[q * 16 + e][ii] * vJ[q * 16 + e];
    // column sums (transposed product) over this thread's quarter of rows
    const int jj = tid & 63;
    double zc = 0.0;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int r = q * 16 + e;
      if (I > J || r > jj) zc += T[jj][r] * vI[r];
    }
    part[q][jj] = zc;
    __syncthreads();
    if (tid < kTB && J * kTB + tid < m)
      Z[(long)I * n + cols0 + tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
  }
  __syncthreads();
  part[q][ii] = yacc;
  __syncthreads();
  if (tid < kTB && I * kTB + tid < m)
    Y[(long)ch * n + rows0 + tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
}

// w = tau * (y - V t1 - W t2) with y reduced from the partial sums; store v into column c.
__global__ __launch_bounds__(256) void finalize_w_kernel(double* a, long lda, int n, int c, int j0,
                                                         double* W, long ldw, const double* Y,
                                                         const double* Z, const double* vbuf,
                                                         const double* t12, const double* tauA,
                                                         double* dot) {
  __shared__ double t1[kNB], t2[kNB], sh[4];
  const int i = c - j0;
  if (threadIdx.x < i) {
    t1[threadIdx.x] = t12[(long)c * 2 * kNB + threadIdx.x];
    t2[threadIdx.x] = t12[(long)c * 2 * kNB + i + threadIdx.x];
  }
  __syncthreads();
  const int m = n - c - 1;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double prod = 0.0;
  if (k < m) {
    const int r = c + 1 + k;
    const int Jr = k / kTB;
    const int nRB = (m + kTB - 1) / kTB;
    double y = 0.0;
    for (int ch = 0; ch * kCH <= nRB - 1; ++ch)
      if (ch * kCH <= Jr || true) {
        // chunk ch contributes to row block I only if ch*kCH <= I; rows of block Jr
        if (ch * kCH <= Jr) y += Y[(long)ch * n + r];
      }
    for (int I = Jr; I < nRB; ++I) y += Z[(long)I * n + r];
    for (int p = 0; p < i; ++p)
      y -= a[r + (long)(j0 + p) * lda] * t1[p] + W[r + (long)p * ldw] * t2[p];
    const double tau = tauA[c];
    const double w = tau * y;
    W[r + (long)i * ldw] = w;
    const double v = vbuf[r];
    a[r + (long)c * lda] = v;
    prod = w * v;
  }
  const double tot = block_reduce_sum(prod, sh);
  if (threadIdx.x == 0 && tot != 0.0) atomicAdd(&dot[c], tot);
}

__global__ void finish_w_kernel(int n, int c, int i, double* W, long ldw, const double* vbuf,
                                const double* tauA, const double* dot) {
  const int m = n - c - 1;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) {
    const int r = c + 1 + k;
    const double al = -0.5 * tauA[c] * dot[c];
    W[r + (long)i * ldw] += al * vbuf[r];
  }
}

// P = [V W], Q = [W V] for the trailing update, rows r0..n-1.
__global__ void panel_pack_kernel(const double* a, long lda, int n, int r0, int j0, int pw,
                                  const double* W, long ldw, double* P, double* Q, long ldp) {
  const int p = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = r0 + k;
  if (r >= n) return;
  const double v = a[r + (long)(j0 + p) * lda];
  const double w = W[r + (long)p * ldw];
  P[k + (long)p * ldp] = v;
  P[k + (long)(p + pw) * ldp] = w;
  Q[k + (long)p * ldp] = w;
  Q[k + (long)(p + pw) * ldp] = v;
}

__global__ void panel_restore_kernel(double* a, long lda, int n, int j0, int pw, double* d,
                                     const double* e) {
  const int p = threadIdx.x;
  if (p >= pw) return;
  const int c = j0 + p;
  d[c] = a[c + (long)c * lda];
  if (c + 1 < n) a[(c + 1) + (long)c * lda] = e[c];
}

int sytrd_lower(hipStream_t s, double* a, long lda, int n, double* d, double* e, double* tau,
                double* work, size_t workSize) {
  DH_CHECK(workSize >= sytrd_work_size(n), "sytrd: workspace too small");
  SytrdWs ws;
  sytrd_layout(n, &ws, work);
  DH_TRY(fill<double>(s, ws.nrm2, n, 0.0));
  DH_TRY(fill<double>(s, ws.dot, n, 0.0));
  DH_TRY(fill<double>(s, ws.t12, (size_t)n * 2 * kNB, 0.0));
  DH_TRY(fill<double>(s, e, std::max(n - 1, 1), 0.0));
  DH_TRY(fill<double>(s, tau, std::max(n - 1, 1), 0.0));
  const long ldw = n;
  for (int j0 = 0; j0 < n; j0 += kNB) {
    const int pw = std::min(kNB, n - j0);
    for (int i = 0; i < pw; ++i) {
      const int c = j0 + i;
      hipLaunchKernelGGL(col_update_kernel, dim3(cdiv(n - c, 256)), dim3(256), 0, s, a, lda, n, c,
                         j0, ws.W, ldw, ws.nrm2);
      DH_LAUNCH_CHECK();
      const int m = n - c - 1;
      if (m <= 0) continue;
      const int nRB = (m + kTB - 1) / kTB;
      const int nCh = (nRB + kCH - 1) / kCH;
      SymvProfile& pf = g_symvProf;
      if (pf.on) {
        if (pf.ev.size() < 2 * (pf.used + 1)) {
          const size_t old = pf.ev.size();
          pf.ev.resize(2 * (pf.used + 1) + 2 * 1024);
          for (size_t q = old; q < pf.ev.size(); ++q) DH_HIP(hipEventCreate(&pf.ev[q]));
          pf.bytes.resize(pf.ev.size() / 2);
        }
        DH_HIP(hipEventRecord(pf.ev[2 * pf.used], s));
      }
      hipLaunchKernelGGL(symv_kernel, dim3(nRB, nCh), dim3(256), 0, s, a, lda, n, c, j0, ws.W,
                         ldw, ws.nrm2, ws.Y, ws.Z, ws.vbuf, ws.t12, ws.scal, e, tau);
      DH_LAUNCH_CHECK();
      if (pf.on) {
        DH_HIP(hipEventRecord(pf.ev[2 * pf.used + 1], s));
        // lower triangle of the m x m trailing matrix plus v in and the partial sums out
        pf.bytes[pf.used] = 8.0 * ((double)m * (m + 1) / 2 + 2.0 * m);
        ++pf.used;
      }
      hipLaunchKernelGGL(finalize_w_kernel, dim3(cdiv(m, 256)), dim3(256), 0, s, a, lda, n, c, j0,
                         ws.W, ldw, ws.Y, ws.Z, ws.vbuf, ws.t12, tau, ws.dot);
      DH_LAUNCH_CHECK();
      hipLaunchKernelGGL(finish_w_kernel, dim3(cdiv(m, 256)), dim3(256), 0, s, n, c, i, ws.W, ldw,
                         ws.vbuf, tau, ws.dot);
      DH_LAUNCH_CHECK();
    }
    const int r0 = j0 + pw;
    if (r0 < n) {
      const int mt = n - r0;
      hipLaunchKernelGGL(panel_pack_kernel, dim3(cdiv(mt, 256), pw), dim3(256), 0, s, a, lda, n,
                         r0, j0, pw, ws.W, ldw, ws.P, ws.Q, (long)mt);
      DH_LAUNCH_CHECK();
      double* a22 = a + r0 + (long)r0 * lda;
      DH_TRY(gemm<double>(s, false, true, mt, mt, 2 * pw, -1.0, ws.P, mt, ws.Q, mt, 1.0, a22, lda,
                          true));
    }
    hipLaunchKernelGGL(panel_restore_kernel, dim3(1), dim3(64), 0, s, a, lda, n, j0, pw, d, e);
    DH_LAUNCH_CHECK();
  }
  DH_TRY(symv_prof_collect(s));
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Back-transformation Z <- Q Z, Q = H(0) H(1) ... H(n-2), panels applied last to first with the
// compact WY form I - V T V^T (T upper triangular, LAPACK dlarft 'F','C').
// ---------------------------------------------------------------------------------------------
struct OrmWs {
  double* V;    // n x kNB
  double* G;    // kNB x kNB
  double* T;    // kNB x kNB
  double* W1;   // kNB x ncols
  double* W2;   // kNB x ncols
};

static size_t ormtr_layout(int n, int ncols, OrmWs* ws, double* base) {
  size_t off = 0;
  OrmWs dummy;
  OrmWs* w = ws ? ws : &dummy;
  auto take = [&](double** p, size_t cnt) {
    if (ws) *p = base + off;
    off += (cnt + 63) / 64 * 64;
  };
  take(&w->V, (size_t)n * kNB);
  take(&w->G, kNB * kNB);
  take(&w->T, kNB * kNB);
  take(&w->W1, (size_t)kNB * std::max(ncols, 1));
  take(&w->W2, (size_t)kNB * std::max(ncols, 1));
  return off;
}

size_t ormtr_work_size(int n, int ncols) { return ormtr_layout(n, ncols, nullptr, nullptr); }

// explicit V (rows j0+1..n-1) of the panel's reflectors
__global__ void form_v_kernel(const double* a, long lda, int n, int j0, int pw, double* V,
                              long ldv) {
  const int p = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = j0 + 1 + k;
  if (r >= n) return;
  const int c = j0 + p;
  double v;
  if (r < c + 1) v = 0.0;
  else if (r == c + 1) v = 1.0;
  else v = a[r + (long)c * lda];
  V[k + (long)p * ldv] = v;
}

// T(i,i) = tau_i, T(0:i,i) = -tau_i T(0:i,0:i) G(0:i,i), G = V^T V
__global__ void larft_kernel(const double* G, const double* tau, int j0, int pw, double* T) {
  __shared__ double Ts[kNB][kNB + 1];
  const int t = threadIdx.x;
  for (int e = t; e < kNB * kNB; e += blockDim.x) Ts[e % kNB][e / kNB] = 0.0;
  __syncthreads();
  for (int i = 0; i < pw; ++i) {
    const double ti = tau[j0 + i];
    if (t < i) {
      double acc = 0.0;
      for (int k = t; k < i; ++k) acc += Ts[t][k] * G[k + i * kNB];
      Ts[t][i] = -ti * acc;
    }
    if (t == 0) Ts[i][i] = ti;
    __syncthreads();
  }
  for (int e = t; e < kNB * kNB; e += blockDim.x) T[e] = Ts[e % kNB][e / kNB];
}

int ormtr_lower(hipStream_t s, const double* a, long lda, int n, const double* tau, double* z,
                long ldz, int ncols, double* work, size_t workSize) {
  DH_CHECK(workSize >= ormtr_work_size(n, ncols), "ormtr: workspace too small");
  if (n <= 1 || ncols <= 0) return 0;
  OrmWs ws;
  ormtr_layout(n, ncols, &ws, work);
  const int nref = n - 1;
  int lastPanel = ((nref - 1) / kNB) * kNB;
  for (int j0 = lastPanel; j0 >= 0; j0 -= kNB) {
    const int pw = std::min(kNB, nref - j0);
    const int m = n - j0 - 1;
    hipLaunchKernelGGL(form_v_kernel, dim3(cdiv(m, 256), pw), dim3(256), 0, s, a, lda, n, j0, pw,
                       ws.V, (long)m);
    DH_LAUNCH_CHECK();
    DH_TRY(gemm<double>(s, true, false, pw, pw, m, 1.0, ws.V, m, ws.V, m, 0.0, ws.G, kNB));
    hipLaunchKernelGGL(larft_kernel, dim3(1), dim3(64), 0, s, ws.G, tau, j0, pw, ws.T);
    DH_LAUNCH_CHECK();
    double* zr = z + (j0 + 1);
    DH_TRY(gemm<double>(s, true, false, pw, ncols, m, 1.0, ws.V, m, zr, ldz, 0.0, ws.W1, kNB));
    DH_TRY(gemm<double>(s, false, false, pw, ncols, pw, 1.0, ws.T, kNB, ws.W1, kNB, 0.0, ws.W2,
                        kNB));
    DH_TRY(gemm<double>(s, false, false, m, ncols, pw, -1.0, ws.V, m, ws.W2, kNB, 1.0, zr, ldz));
  }
  return 0;
}

}  // namespace dftbhip
