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
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "common.h"

namespace dftbhip {

constexpr int kNB = 64;      // panel width
// Panel products t1 = W^T v, t2 = V^T v of the column being reduced: one 128-byte line per entry
// (the SYMV's workgroups add into them atomically; entries sharing a line would serialize those
// atomics) and two columns in a ring, the next column's entries cleared by col_update.
constexpr int kT12Pad = 16;
__host__ __device__ inline long t12_at(int c, int p) {
  return ((long)(c & 1) * 2 * kNB + p) * kT12Pad;
}
constexpr long kT12Size = 2L * 2 * kNB * kT12Pad;

// Optional per-launch timing of the symmetric product kernel (HIP events on the launch stream):
// launches, summed duration and the algorithmic bytes (lower triangle of the trailing matrix).
struct SymvProfile {
  bool on = false;
  std::vector<hipEvent_t> ev;
  std::vector<double> bytes;
  std::vector<int> size;   // trailing size m of each recorded launch
  size_t used = 0;
  double launches = 0, seconds = 0, algBytes = 0;
  std::vector<std::pair<int, float>>* sink = nullptr;   // (m, ms) per launch, when set
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
    if (p.sink) p.sink->push_back({p.size[i], ms});
    p.launches += 1;
  }
  p.used = 0;
  return 0;
}
constexpr int kTB = 64;      // SYMV tile
constexpr int kSymvTiles8From = 7000;   // trailing sizes from which 8 tiles per workgroup
constexpr int kSymvTiles4From = 2500;   // ... and from which 4
constexpr int kCHMin = 2;    // fewest row tiles per SYMV workgroup (sizes the slot array)
// row tiles per SYMV workgroup (2, 4, 8 or 16; 0: chosen from the trailing size), set through
// set_symv_tiles
static int g_symvTiles = 0;
// bumped whenever the tile choice changes: captured launch sequences (below) are only replayed
// while it is unchanged
static unsigned g_tileGen = 0;

int set_symv_tiles(int tiles) {
  DH_CHECK(tiles == 0 || tiles == 2 || tiles == 4 || tiles == 8 || tiles == 16,
           "symv tiles must be 0 (by size), 2, 4, 8 or 16");
  g_symvTiles = tiles;
  ++g_tileGen;
  return 0;
}

// Tile choice by trailing size from the autotuner (symv_autotune), per arithmetic type
// ([0] FP64, [1] FP32): rows (first trailing size, tiles) in ascending size; empty: the fixed
// thresholds above.
static std::vector<std::pair<int, int>> g_tileTable[2];

int clear_symv_tables() {
  g_tileTable[0].clear();
  g_tileTable[1].clear();
  ++g_tileGen;
  return 0;
}

// Long tile chains keep more loads in flight per workgroup on a large trailing matrix; a small
// one has too few column blocks to fill the chip, and short chains there give more workgroups.
template <typename T>
static int symv_tiles(int m) {
  if (g_symvTiles != 0) return g_symvTiles;
  const auto& t = g_tileTable[std::is_same<T, float>::value ? 1 : 0];
  if (!t.empty()) {
    int v = t.front().second;
    for (const auto& r : t)
      if (m >= r.first) v = r.second;
    return v;
  }
  return m >= kSymvTiles8From ? 8 : (m >= kSymvTiles4From ? 4 : 2);
}

template <typename T>
struct SytrdWs {
  T* W;            // n x kNB
  T* Y;            // nSlot x n   partial sums of A22 v, slot-major: Y(row, slot) at slot n + row
  T* vbuf;         // n
  double* nrm2;    // n   ||x||^2 per column
  double* dot;     // n   w.v per column (before the rank-one correction of w)
  double* t12;     // panel products, see t12_at
  double* scal;    // n
  T* P;            // n x 2kNB
  T* Q;            // n x 2kNB
  int nSlot;
};

static int sytrd_slots(int n) {
  const int nRB = (n + kTB - 1) / kTB;
  return nRB + (nRB + kCHMin - 1) / kCHMin;
}

// byte layout of the workspace (each array 256-byte aligned)
template <typename T>
static size_t sytrd_layout(int n, SytrdWs<T>* ws, char* base) {
  size_t off = 0;
  auto take = [&](auto** p, size_t bytes) {
    if (ws) *p = reinterpret_cast<std::remove_reference_t<decltype(**p)>*>(base + off);
    off += (bytes + 255) / 256 * 256;
  };
  SytrdWs<T> dummy;
  SytrdWs<T>* w = ws ? ws : &dummy;
  w->nSlot = sytrd_slots(n);
  take(&w->W, sizeof(T) * n * kNB);
  take(&w->Y, sizeof(T) * n * w->nSlot);
  take(&w->vbuf, sizeof(T) * n);
  take(&w->nrm2, sizeof(double) * n);
  take(&w->dot, sizeof(double) * n);
  take(&w->t12, sizeof(double) * kT12Size);
  take(&w->scal, sizeof(double) * n);
  take(&w->P, sizeof(T) * n * 2 * kNB);
  take(&w->Q, sizeof(T) * n * 2 * kNB);
  return off;
}

// workspace bytes
size_t sytrd_work_bytes(int n, int elemSize) {
  return elemSize == 8 ? sytrd_layout<double>(n, nullptr, nullptr)
                       : sytrd_layout<float>(n, nullptr, nullptr);
}
size_t sytrd_work_size(int n) { return (sytrd_work_bytes(n, 8) + 7) / 8; }

// Several problems of one size in one launch: problem blockIdx.z finds its matrix, workspace
// block and vectors at these strides from those of problem 0 (all zero for a single problem).
struct ProblemStrides {
  long a;    // matrix elements
  long ws;   // workspace bytes
  long v;    // elements of d, e, tau
};

template <typename P>
__device__ __forceinline__ P* ws_of(P* p, long bytes) {
  return (P*)((const char*)p + (long)blockIdx.z * bytes);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
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

// W stores w before its rank-one correction: the true column p of the panel is
// W(:,p) + alpha_p V(:,p) with alpha_p = -tau_p (w_p . v_p) / 2, applied by every reader.
template <typename T>
__device__ __forceinline__ double wcorr(int cp, const T* tau, const double* dot) {
  return -0.5 * (double)tau[cp] * dot[cp];
}

// A(r,c) -= sum_p V(r,p) W(c,p) + W(r,p) V(c,p), rows r >= c; accumulate ||A(c+2:n,c)||^2.
// A block handles 64 rows (the lanes); wave g takes the reflectors p = 16 g .. 16 g + 15 with
// all their loads in one batch, and the four partial sums meet in LDS.
template <typename T>
__global__ __launch_bounds__(256) void col_update_kernel(T* a, long lda, int n, int c, int j0,
                                                         const T* W, long ldw, const T* tau,
                                                         const double* dot, double* nrm2,
                                                         double* t12, ProblemStrides ps) {
  __shared__ double vc[kNB], wc[kNB], al[kNB], part[4][64];
  a += (long)blockIdx.z * ps.a;
  tau += (long)blockIdx.z * ps.v;
  W = ws_of(W, ps.ws);
  dot = ws_of(dot, ps.ws);
  nrm2 = ws_of(nrm2, ps.ws);
  t12 = ws_of(t12, ps.ws);
  const int i = c - j0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (threadIdx.x < i) {
    const int p = threadIdx.x;
    al[p] = wcorr(j0 + p, tau, dot);
    vc[p] = a[c + (long)(j0 + p) * lda];
    wc[p] = W[c + (long)p * ldw] + al[p] * vc[p];
  }
  // the panel products of column c + 1 start from zero (their ring entry was column c - 1's,
  // whose readers have finished)
  if (blockIdx.x == 0)
    for (int e = threadIdx.x; e < 2 * kNB; e += blockDim.x) t12[t12_at(c + 1, e)] = 0.0;
  __syncthreads();
  const int r = c + blockIdx.x * 64 + lane;
  double d = 0.0;
  const int p0 = 16 * wv;
  if (r < n && p0 < i) {
    double vr[16], wr[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      vr[q] = 0.0;
      wr[q] = 0.0;
      if (p0 + q < i) {
        vr[q] = a[r + (long)(j0 + p0 + q) * lda];
        wr[q] = W[r + (long)(p0 + q) * ldw];
      }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int p = p0 + q;
      if (p < i) d += vr[q] * wc[p] + (wr[q] + al[p] * vr[q]) * vc[p];
    }
  }
  part[wv][lane] = d;
  __syncthreads();
  if (wv == 0) {
    double sq = 0.0;
    if (r < n) {
      double v = a[r + (long)c * lda] - ((part[0][lane] + part[1][lane]) +
                                         (part[2][lane] + part[3][lane]));
      a[r + (long)c * lda] = (T)v;
      v = (double)(T)v;
      if (r >= c + 2) sq = v * v;
    }
    const double tot = wave_sum(sq);
    if (lane == 0 && tot != 0.0) atomicAdd(&nrm2[c], tot);
  }
}

// Partial sums of y = A22 v (A22 = a(c+1:n, c+1:n), lower triangle read once).
// Workgroup (J, chunk) owns column block J and row blocks J + chunk*CH ... (CH of them):
//   y partial of row block I  -> Y(row, J)            (one per tile, reduced over 4 waves)
//   transposed sums (A^T v)_J -> Y(col, nRB + chunk)  (accumulated in registers over the tiles)
// Thread (ii = tid % 64, cg = tid / 64) holds rows ii of 16 columns cg*16 .. cg*16+15.
// The chunk-0 workgroup of each column block publishes v and the panel products t1, t2.
template <typename T, int CH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void symv_kernel(const T* a, long lda, int n, int c, int j0,
                                                   const T* W, long ldw, const double* nrm2,
                                                   const T* tauA, const double* dotA, T* Y,
                                                   int nSlot, T* vbuf, double* t12,
                                                   double* scalOut, T* eOut, T* tauOut,
                                                   ProblemStrides ps) {
  __shared__ T vJs[kTB];
  __shared__ T red[4][4][kTB];
  a += (long)blockIdx.z * ps.a;
  tauA += (long)blockIdx.z * ps.v;
  eOut += (long)blockIdx.z * ps.v;
  tauOut += (long)blockIdx.z * ps.v;
  W = ws_of(W, ps.ws);
  nrm2 = ws_of(nrm2, ps.ws);
  dotA = ws_of(dotA, ps.ws);
  Y = ws_of(Y, ps.ws);
  vbuf = ws_of(vbuf, ps.ws);
  t12 = ws_of(t12, ps.ws);
  scalOut = ws_of(scalOut, ps.ws);
  const int m = n - c - 1;
  const int nRB = (m + kTB - 1) / kTB;
  const int J = blockIdx.x, ch = blockIdx.y;
  const int Ib = J + ch * CH;
  if (J >= nRB || Ib >= nRB) return;
  const int Ie = min(Ib + CH, nRB);
  const int tid = threadIdx.x;
  const int ii = tid & 63, cg = tid >> 6;
  const double alpha = a[(c + 1) + (long)c * lda];
  double beta, tau, scal;
  reflector(alpha, nrm2[c], beta, tau, scal);
  if (J == 0 && ch == 0 && tid == 0) {
    scalOut[c] = scal;
    eOut[c] = (T)beta;
    tauOut[c] = (T)tau;
  }
  auto vval = [&](int k) -> T {
    if (k >= m) return T(0);
    return (k == 0) ? T(1) : (T)((double)a[(c + 1 + k) + (long)c * lda] * scal);
  };
  if (tid < kTB) vJs[tid] = vval(J * kTB + tid);
  __syncthreads();
  const int cols0 = c + 1 + J * kTB;
  if (ch == 0) {
    if (tid < kTB && J * kTB + tid < m) vbuf[cols0 + tid] = vJs[tid];
    // panel products over this block's rows: t12[p] += W(:, p)^T v, t12[i + p] += V(:, p)^T v
    // for p < i (wave cg takes p = cg + 4q). All loads are issued before the reductions and
    // atomics, which would otherwise order every load behind the previous atomic.
    const int i = c - j0;
    const bool rowOk = J * kTB + ii < m;
    const int r = cols0 + ii;
    const double vv = (double)vJs[ii];
    double pw[kNB / 4], pv[kNB / 4];
#pragma unroll
    for (int q = 0; q < kNB / 4; ++q) {
      const int p = cg + 4 * q;
      pw[q] = 0.0;
      pv[q] = 0.0;
      if (p < i && rowOk) {
        const double vr = a[r + (long)(j0 + p) * lda];
        pv[q] = vr * vv;
        pw[q] = ((double)W[r + (long)p * ldw] + wcorr(j0 + p, tauA, dotA) * vr) * vv;
      }
    }
#pragma unroll
    for (int q = 0; q < kNB / 4; ++q) {
      const int p = cg + 4 * q;
      if (p < i) {
        const double sw = wave_sum(pw[q]), sv = wave_sum(pv[q]);
        if (ii == 0 && sw != 0.0) atomicAdd(&t12[t12_at(c, p)], sw);
        if (ii == 0 && sv != 0.0) atomicAdd(&t12[t12_at(c, i + p)], sv);
      }
    }
  }
  const T* vj = vJs + cg * 16;   // v of this thread's 16 columns (LDS broadcast reads)
  T b0[16], b1[16], b2[16], zacc[16];
  T v0 = T(0), v1 = T(0), v2 = T(0);
#pragma unroll
  for (int e = 0; e < 16; ++e) zacc[e] = T(0);
  // tile I rows ii of this thread's 16 columns, and v of row ii (prefetched with the tile)
  auto load = [&](int I, T* dst, T& vdst) {
    const int gr = I * kTB + ii;
    const bool rowOk = gr < m;
    const T* src = a + (c + 1 + gr) + (long)cols0 * lda;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int jj = cg * 16 + e;
      const int gc = J * kTB + jj;
      dst[e] = (rowOk && gc < m && (I > J || ii >= jj)) ? src[(long)jj * lda] : T(0);
    }
    // raw column-c element of the row; scaled at use so that the prefetch is not waited on
    vdst = (rowOk && gr > 0) ? a[(c + 1 + gr) + (long)c * lda] : T(0);
  };
  // barrier for the LDS row sums only: __syncthreads would also wait for the prefetched global
  // loads of the next tiles (its workgroup fence waits on every outstanding memory operation)
  auto lds_barrier = [&]() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0); vmcnt, expcnt untouched (gfx9 encoding)
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  };
  // one tile: row partials (reduced over the 4 waves through LDS) and transposed sums
  auto tile = [&](int I, const T* t, T vraw) {
    const int gr = I * kTB + ii;
    const T vi = (gr >= m) ? T(0) : (gr == 0 ? T(1) : (T)((double)vraw * scal));
    T yp = T(0);
    if (I > J) {
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        yp += t[e] * vj[e];
        zacc[e] += t[e] * vi;
      }
    } else {
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        yp += t[e] * vj[e];
        if (ii > cg * 16 + e) zacc[e] += t[e] * vi;
      }
    }
    red[I & 3][cg][ii] = yp;
    // every second tile (and the last): one barrier, then the rows of the last two tiles are
    // summed over the waves (wave 0 and wave 1 take one tile each). A buffer is written again
    // four tiles later, after at least one more barrier.
    if (((I - Ib) & 1) == 0 && I + 1 < Ie) return;
    lds_barrier();
    const int Ir = (cg == 0) ? I : I - 1;
    if (cg < 2 && Ir >= Ib && (cg == 0 || ((I - Ib) & 1) == 1)) {
      const int k = Ir * kTB + ii;
      if (k < m)
        Y[(long)J * n + k] = red[Ir & 3][0][ii] + red[Ir & 3][1][ii] + red[Ir & 3][2][ii] +
                             red[Ir & 3][3][ii];
    }
  };
  // two tiles in flight ahead of the one being reduced
  load(Ib, b0, v0);
  if (Ib + 1 < Ie) load(Ib + 1, b1, v1);
  for (int I = Ib; I < Ie; I += 3) {
    if (I + 2 < Ie) load(I + 2, b2, v2);
    tile(I, b0, v0);
    if (I + 1 >= Ie) break;
    if (I + 3 < Ie) load(I + 3, b0, v0);
    tile(I + 1, b1, v1);
    if (I + 2 >= Ie) break;
    if (I + 4 < Ie) load(I + 4, b1, v1);
    tile(I + 2, b2, v2);
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const T z = wave_sum(zacc[e]);
    const int k = J * kTB + cg * 16 + e;
    if (ii == 0 && k < m) Y[(long)(nRB + ch) * n + k] = z;
  }
}

// y(r) from its slots, w = tau * (y - V t1 - W t2); store v into column c and accumulate w.v.
// Slots are stored slot-major (Y(row, slot) at slot * n + row): a block handles 64 rows (the
// lanes) and its kFinW waves split the slots and the panel correction between them (many
// waves per row block keep enough loads in flight: there are only m / 64 blocks).
constexpr int kFinW = 16;
template <typename T>
__global__ __launch_bounds__(64 * kFinW) void finalize_w_kernel(T* a, long lda, int n, int c,
                                                                int j0, T* W, long ldw, const T* Y,
                                                                int nSlot, const T* vbuf,
                                                                const double* t12, const T* tauA,
                                                                double* dot, int tiles,
                                                                ProblemStrides ps) {
  __shared__ double red[kFinW][64];
  a += (long)blockIdx.z * ps.a;
  tauA += (long)blockIdx.z * ps.v;
  W = ws_of(W, ps.ws);
  Y = ws_of(Y, ps.ws);
  vbuf = ws_of(vbuf, ps.ws);
  t12 = ws_of(t12, ps.ws);
  dot = ws_of(dot, ps.ws);
  const int m = n - c - 1;
  const int nRB = (m + kTB - 1) / kTB;
  const int i = c - j0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int k = blockIdx.x * 64 + lane;
  const int r = c + 1 + k;
  double acc = 0.0;
  if (k < m) {
    const int Ir = k / kTB;
    const int nch = (nRB - Ir + tiles - 1) / tiles;
    // this wave's slots (sl = wv, wv + kFinW, ...) with four independent sums
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int sl = wv;
    for (; sl + 3 * kFinW <= Ir; sl += 4 * kFinW) {
      s0 += Y[(long)sl * n + k];
      s1 += Y[(long)(sl + kFinW) * n + k];
      s2 += Y[(long)(sl + 2 * kFinW) * n + k];
      s3 += Y[(long)(sl + 3 * kFinW) * n + k];
    }
    for (; sl <= Ir; sl += kFinW) s0 += Y[(long)sl * n + k];
    const T* Yc = Y + (long)nRB * n;
    sl = wv;
    for (; sl + 3 * kFinW < nch; sl += 4 * kFinW) {
      s1 += Yc[(long)sl * n + k];
      s2 += Yc[(long)(sl + kFinW) * n + k];
      s3 += Yc[(long)(sl + 2 * kFinW) * n + k];
      s0 += Yc[(long)(sl + 3 * kFinW) * n + k];
    }
    for (; sl < nch; sl += kFinW) s1 += Yc[(long)sl * n + k];
    acc = (s0 + s1) + (s2 + s3);
    // panel correction of this wave's p = wv + kFinW q, loads ahead of the products
    double vr[kNB / kFinW], wr[kNB / kFinW];
#pragma unroll
    for (int q = 0; q < kNB / kFinW; ++q) {
      const int p = wv + kFinW * q;
      vr[q] = 0.0;
      wr[q] = 0.0;
      if (p < i) {
        vr[q] = a[r + (long)(j0 + p) * lda];
        wr[q] = W[r + (long)p * ldw];
      }
    }
#pragma unroll
    for (int q = 0; q < kNB / kFinW; ++q) {
      const int p = wv + kFinW * q;
      if (p < i) {
        const double w = wr[q] + wcorr(j0 + p, tauA, dot) * vr[q];
        acc -= vr[q] * t12[t12_at(c, p)] + w * t12[t12_at(c, i + p)];
      }
    }
  }
  red[wv][lane] = acc;
  __syncthreads();
  if (wv == 0) {
    double prod = 0.0;
    if (k < m) {
      double y = 0.0;
#pragma unroll
      for (int q = 0; q < kFinW; ++q) y += red[q][lane];
      const T w = (T)((double)tauA[c] * y);
      W[r + (long)i * ldw] = w;
      const T v = vbuf[r];
      a[r + (long)c * lda] = v;
      prod = (double)w * (double)v;
    }
    prod = wave_sum(prod);
    if (lane == 0 && prod != 0.0) atomicAdd(&dot[c], prod);
  }
}

// P = [V W], Q = [W V] for the trailing update, rows r0..n-1 (W with its rank-one correction).
template <typename T>
__global__ void panel_pack_kernel(const T* a, long lda, int n, int r0, int j0, int pw,
                                  const T* W, long ldw, const T* tau, const double* dot, T* P,
                                  T* Q, long ldp) {
  const int p = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = r0 + k;
  if (r >= n) return;
  const T v = a[r + (long)(j0 + p) * lda];
  const T w = (T)((double)W[r + (long)p * ldw] + wcorr(j0 + p, tau, dot) * (double)v);
  P[k + (long)p * ldp] = v;
  P[k + (long)(p + pw) * ldp] = w;
  Q[k + (long)p * ldp] = w;
  Q[k + (long)(p + pw) * ldp] = v;
}

template <typename T>
__global__ void panel_restore_kernel(T* a, long lda, int n, int j0, int pw, T* d, const T* e) {
  const int p = threadIdx.x;
  if (p >= pw) return;
  const int c = j0 + p;
  d[c] = a[c + (long)c * lda];
  if (c + 1 < n) a[(c + 1) + (long)c * lda] = e[c];
}

// K problems of size n at once: problem k's matrix at a + k sA, d, e, tau at + k sV and its
// workspace block at work + k wsStride bytes (wsStride >= sytrd_layout(n)). The column kernels
// run all problems in one launch each; the panel products and trailing updates run per problem.
template <typename T>
static int sytrd_launches(hipStream_t s, T* a, long lda, long sA, int n, T* d, T* e, T* tau,
                          long sV, void* work, size_t wsStride, int K) {
  std::vector<SytrdWs<T>> wk(K);
  for (int k = 0; k < K; ++k) {
    sytrd_layout<T>(n, &wk[k], static_cast<char*>(work) + (size_t)k * wsStride);
    DH_TRY(fill<double>(s, wk[k].nrm2, n, 0.0));
    DH_TRY(fill<double>(s, wk[k].dot, n, 0.0));
    DH_TRY(fill<double>(s, wk[k].t12, (size_t)kT12Size, 0.0));
    DH_TRY(fill<T>(s, e + k * sV, std::max(n - 1, 1), T(0)));
    DH_TRY(fill<T>(s, tau + k * sV, std::max(n, 1), T(0)));
  }
  const SytrdWs<T>& ws = wk[0];
  const ProblemStrides ps{K > 1 ? sA : 0, K > 1 ? (long)wsStride : 0, K > 1 ? sV : 0};
  const long ldw = n;
  for (int j0 = 0; j0 < n; j0 += kNB) {
    const int pw = std::min(kNB, n - j0);
    for (int i = 0; i < pw; ++i) {
      const int c = j0 + i;
      hipLaunchKernelGGL(col_update_kernel<T>, dim3(cdiv(n - c, 64), 1, K), dim3(256), 0, s, a,
                         lda, n, c, j0, ws.W, ldw, tau, ws.dot, ws.nrm2, ws.t12, ps);
      DH_LAUNCH_CHECK();
      const int m = n - c - 1;
      if (m <= 0) continue;
      const int nRB = (m + kTB - 1) / kTB;
      // K problems in one launch fill the chip like one problem of K times the row blocks
      const int tiles = symv_tiles<T>((int)std::min<long>((long)m * K, 1L << 30));
      const int nCh = (nRB + tiles - 1) / tiles;
      SymvProfile& pf = g_symvProf;
      if (pf.on) {
        if (pf.ev.size() < 2 * (pf.used + 1)) {
          const size_t old = pf.ev.size();
          pf.ev.resize(2 * (pf.used + 1) + 2 * 1024);
          for (size_t q = old; q < pf.ev.size(); ++q) DH_HIP(hipEventCreate(&pf.ev[q]));
          pf.bytes.resize(pf.ev.size() / 2);
          pf.size.resize(pf.ev.size() / 2);
        }
        DH_HIP(hipEventRecord(pf.ev[2 * pf.used], s));
      }
#define DH_SYMV(CH_)                                                                        \
  hipLaunchKernelGGL((symv_kernel<T, CH_>), dim3(nRB, nCh, K), dim3(256), 0, s, a, lda, n, c, \
                     j0, ws.W, ldw, ws.nrm2, tau, ws.dot, ws.Y, ws.nSlot, ws.vbuf, ws.t12,      \
                     ws.scal, e, tau, ps)
      if (tiles == 2) DH_SYMV(2);
      else if (tiles == 4) DH_SYMV(4);
      else if (tiles == 16) DH_SYMV(16);
      else DH_SYMV(8);
#undef DH_SYMV
      DH_LAUNCH_CHECK();
      if (pf.on) {
        DH_HIP(hipEventRecord(pf.ev[2 * pf.used + 1], s));
        // lower triangle of the m x m trailing matrix plus v in and the partial sums out, for
        // each problem
        pf.bytes[pf.used] = (double)K * sizeof(T) * ((double)m * (m + 1) / 2 + 2.0 * m);
        pf.size[pf.used] = m;
        ++pf.used;
      }
      hipLaunchKernelGGL(finalize_w_kernel<T>, dim3(cdiv(m, 64), 1, K), dim3(64 * kFinW), 0, s,
                         a, lda, n, c, j0, ws.W, ldw, ws.Y, ws.nSlot, ws.vbuf, ws.t12, tau,
                         ws.dot, tiles, ps);
      DH_LAUNCH_CHECK();
    }
    const int r0 = j0 + pw;
    for (int k = 0; k < K; ++k) {
      T* ak = a + k * sA;
      const T* tk = tau + k * sV;
      if (r0 < n) {
        const int mt = n - r0;
        hipLaunchKernelGGL(panel_pack_kernel<T>, dim3(cdiv(mt, 256), pw), dim3(256), 0, s, ak,
                           lda, n, r0, j0, pw, wk[k].W, ldw, tk, wk[k].dot, wk[k].P, wk[k].Q,
                           (long)mt);
        DH_LAUNCH_CHECK();
        T* a22 = ak + r0 + (long)r0 * lda;
        DH_TRY(gemm<T>(s, false, true, mt, mt, 2 * pw, T(-1), wk[k].P, mt, wk[k].Q, mt, T(1),
                       a22, lda, true));
      }
      hipLaunchKernelGGL(panel_restore_kernel<T>, dim3(1), dim3(64), 0, s, ak, lda, n, j0, pw,
                         d + k * sV, e + k * sV);
      DH_LAUNCH_CHECK();
    }
  }
  return 0;
}

// The launch sequence of one tridiagonalisation (3 launches per column and the panel updates:
// about 47 000 launches at n = 15769) depends only on the sizes, the buffers and the tile
// choice, and holds no host synchronisation: it is captured once into a HIP graph and replayed
// for every later call with the same arguments (an SCC cycle calls it once per iteration on
// the same buffers), which removes the host launch cost between the short column kernels
// (when switched on: set_sytrd_graphs).
// Cleared memory inside the sequence is written by fill kernels (no memset nodes). A few
// sequences are kept, the oldest replaced; not while the per-launch SYMV timing is on.
struct SytrdGraphKey {
  const void* a;
  long lda, sA, sV;
  int n, K, prec;
  const void *d, *e, *tau, *work;
  size_t wsStride;
  unsigned gen;
  bool operator==(const SytrdGraphKey& o) const {
    return a == o.a && lda == o.lda && sA == o.sA && sV == o.sV && n == o.n && K == o.K &&
           prec == o.prec && d == o.d && e == o.e && tau == o.tau && work == o.work &&
           wsStride == o.wsStride && gen == o.gen;
  }
};
struct SytrdGraph {
  SytrdGraphKey key;
  hipGraphExec_t exec = nullptr;
  unsigned long use = 0;
};
static std::vector<SytrdGraph> g_sytrdGraphs;
static std::mutex g_sytrdGraphMutex;
static unsigned long g_sytrdGraphUse = 0;
// off by default: the replay saves about 2 % of an FP32 iteration, and rocprofv3's kernel
// tracing takes the process down on the capture (a host fault inside the tool)
static bool g_sytrdGraphOn = false;
constexpr size_t kSytrdGraphsKept = 4;

int set_sytrd_graphs(bool on) {
  g_sytrdGraphOn = on;
  return 0;
}

void release_sytrd_graphs() {
  std::lock_guard<std::mutex> lock(g_sytrdGraphMutex);
  for (auto& g : g_sytrdGraphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g_sytrdGraphs.clear();
}

template <typename T>
int sytrd_lower_batched_t(hipStream_t s, T* a, long lda, long sA, int n, T* d, T* e, T* tau,
                          long sV, void* work, size_t wsStride, int K) {
  DH_CHECK(K >= 1, "sytrd: no problems");
  DH_CHECK(wsStride >= sytrd_layout<T>(n, nullptr, nullptr), "sytrd: workspace too small");
  DH_CHECK(K == 1 || (sA >= lda * (long)n && sV >= n && wsStride % 256 == 0),
           "sytrd: problems overlap or are misaligned");
  if (!g_sytrdGraphOn || g_symvProf.on || n < 2) {
    DH_TRY(sytrd_launches<T>(s, a, lda, sA, n, d, e, tau, sV, work, wsStride, K));
    DH_TRY(symv_prof_collect(s));
    return 0;
  }
  std::lock_guard<std::mutex> lock(g_sytrdGraphMutex);
  const SytrdGraphKey key{a, lda, sA, sV, n, K, (int)sizeof(T), d, e, tau, work, wsStride,
                          g_tileGen};
  for (auto& g : g_sytrdGraphs)
    if (g.key == key) {
      g.use = ++g_sytrdGraphUse;
      DH_HIP(hipGraphLaunch(g.exec, s));
      return 0;
    }
  DH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = sytrd_launches<T>(s, a, lda, sA, n, d, e, tau, sV, work, wsStride, K);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &graph);
  if (rc != 0) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    DH_HIP(ec);
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  DH_HIP(ei);
  if (g_sytrdGraphs.size() >= kSytrdGraphsKept) {
    auto old = std::min_element(g_sytrdGraphs.begin(), g_sytrdGraphs.end(),
                                [](const SytrdGraph& x, const SytrdGraph& y) {
                                  return x.use < y.use;
                                });
    (void)hipGraphExecDestroy(old->exec);
    g_sytrdGraphs.erase(old);
  }
  g_sytrdGraphs.push_back(SytrdGraph{key, exec, ++g_sytrdGraphUse});
  DH_HIP(hipGraphLaunch(exec, s));
  return 0;
}

template <typename T>
int sytrd_lower_t(hipStream_t s, T* a, long lda, int n, T* d, T* e, T* tau, void* work,
                  size_t workBytes) {
  return sytrd_lower_batched_t<T>(s, a, lda, 0, n, d, e, tau, 0, work, workBytes, 1);
}

// ---------------------------------------------------------------------------------------------
// Back-transformation Z <- Q Z, Q = H(0) H(1) ... H(n-2), panels applied last to first with the
// compact WY form I - V T V^T (T upper triangular, LAPACK dlarft 'F','C').
// ---------------------------------------------------------------------------------------------
template <typename T>
struct OrmWs {
  T* V;         // n x kOB
  T* G;         // kOB x kOB
  T* Tm;        // kOB x kOB
  T* W1;        // kOB x ncols
  T* W2;        // kOB x ncols
};

constexpr int kOB = 256;    // reflectors per block of the back-transformation
constexpr int kOBH = 128;   // half block: the compact-WY T of a half is formed in LDS

template <typename T>
static size_t ormtr_layout(int n, int ncols, OrmWs<T>* ws, char* base) {
  size_t off = 0;
  OrmWs<T> dummy;
  OrmWs<T>* w = ws ? ws : &dummy;
  auto take = [&](T** p, size_t cnt) {
    if (ws) *p = reinterpret_cast<T*>(base + off);
    off += (cnt * sizeof(T) + 255) / 256 * 256;
  };
  take(&w->V, (size_t)n * kOB);
  take(&w->G, kOB * kOB);
  take(&w->Tm, kOB * kOB);
  take(&w->W1, (size_t)kOB * std::max(ncols, 1));
  take(&w->W2, (size_t)kOB * std::max(ncols, 1));
  return off;
}

size_t ormtr_work_bytes(int n, int ncols, int elemSize) {
  return elemSize == 8 ? ormtr_layout<double>(n, ncols, nullptr, nullptr)
                       : ormtr_layout<float>(n, ncols, nullptr, nullptr);
}
size_t ormtr_work_size(int n, int ncols) { return (ormtr_work_bytes(n, ncols, 8) + 7) / 8; }

// explicit V (rows j0+shift..n-1) of the block's reflectors; reflector c has its unit element
// at row c + shift and its other elements below it in column c of a
template <typename T>
__global__ void form_v_kernel(const T* a, long lda, int n, int j0, int pw, int shift, T* V,
                              long ldv) {
  const int p = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = j0 + shift + k;
  if (r >= n) return;
  const int c = j0 + p;
  T v;
  if (r < c + shift) v = T(0);
  else if (r == c + shift) v = T(1);
  else v = a[r + (long)c * lda];
  V[k + (long)p * ldv] = v;
}

// T(i,i) = tau_i, T(0:i,i) = -tau_i T(0:i,0:i) G(0:i,i), G = V^T V   (LAPACK dlarft 'F','C').
// The strict upper triangle of G is read once into LDS; T is built in LDS, both packed by
// columns (entry (r, c), r <= c, at r + c(c+1)/2), so the recurrence makes no global accesses.
// pw <= kOBH reflectors; G and Tout with leading dimension ld (only the pw x pw block of Tout
// is written)
constexpr int kOBPacked = kOBH * (kOBH + 1) / 2;
template <typename T>
__global__ __launch_bounds__(128) void larft_kernel(const T* G, long ld, const T* tau, int j0,
                                                    int pw, T* Tout) {
  __shared__ T Tp[kOBPacked];
  __shared__ T Gp[kOBPacked];
  const int t = threadIdx.x;
  for (int e = t; e < pw * pw; e += blockDim.x) {
    const int r = e % pw, c = e / pw;
    if (r < c) Gp[r + c * (c + 1) / 2] = G[r + (long)c * ld];
  }
  __syncthreads();
  for (int i = 0; i < pw; ++i) {
    const T ti = tau[j0 + i];
    const int ci = i * (i + 1) / 2;
    if (t < i) {
      // eight independent partial sums, so that the LDS reads of a group overlap
      T acc[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] = T(0);
      int k = t;
      for (; k + 8 <= i; k += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] += Tp[t + (k + u) * (k + u + 1) / 2] * Gp[k + u + ci];
      }
      for (; k < i; ++k) acc[0] += Tp[t + k * (k + 1) / 2] * Gp[k + ci];
      Tp[t + ci] = -ti * (((acc[0] + acc[1]) + (acc[2] + acc[3])) +
                          ((acc[4] + acc[5]) + (acc[6] + acc[7])));
    } else if (t == i) {
      Tp[i + ci] = ti;
    }
    __syncthreads();
  }
  for (int e = t; e < pw * pw; e += blockDim.x) {
    const int r = e % pw, c = e / pw;
    Tout[r + (long)c * ld] = (r <= c) ? Tp[r + c * (c + 1) / 2] : T(0);
  }
}

template <typename T>
int ormtr_lower_t(hipStream_t s, const T* a, long lda, int n, const T* tau, T* z, long ldz,
                  int ncols, void* work, size_t workBytes, int shift, int nref) {
  DH_CHECK(workBytes >= ormtr_layout<T>(n, ncols, nullptr, nullptr), "ormtr: workspace too small");
  if (nref <= 0 || ncols <= 0) return 0;
  OrmWs<T> ws;
  ormtr_layout<T>(n, ncols, &ws, static_cast<char*>(work));
  const int lastBlock = ((nref - 1) / kOB) * kOB;
  for (int j0 = lastBlock; j0 >= 0; j0 -= kOB) {
    const int pw = std::min(kOB, nref - j0);
    const int m = n - j0 - shift;
    if (m <= 0) continue;
    hipLaunchKernelGGL(form_v_kernel<T>, dim3(cdiv(m, 256), pw), dim3(256), 0, s, a, lda, n, j0,
                       pw, shift, ws.V, (long)m);
    DH_LAUNCH_CHECK();
    const int sk = std::max(1, std::min(64, m / 512));
    DH_TRY(gemm<T>(s, true, false, pw, pw, m, T(1), ws.V, m, ws.V, m, T(0), ws.G, kOB, false,
                   kKFull, nullptr, sk));
    // T of each half in LDS, then the coupling block of H1 H2 = I - [V1 V2] T [V1 V2]^T:
    // T12 = -T11 (V1^T V2) T22 (the product goes through G's unused lower-left block)
    const int p1 = std::min(pw, kOBH), p2 = pw - p1;
    DH_TRY(fill<T>(s, ws.Tm, (size_t)kOB * kOB, T(0)));
    hipLaunchKernelGGL(larft_kernel<T>, dim3(1), dim3(128), 0, s, ws.G, (long)kOB, tau, j0, p1,
                       ws.Tm);
    DH_LAUNCH_CHECK();
    if (p2 > 0) {
      T* T22 = ws.Tm + p1 + (long)p1 * kOB;
      hipLaunchKernelGGL(larft_kernel<T>, dim3(1), dim3(128), 0, s, ws.G + p1 + (long)p1 * kOB,
                         (long)kOB, tau, j0 + p1, p2, T22);
      DH_LAUNCH_CHECK();
      T* X = ws.G + p1;
      DH_TRY(gemm<T>(s, false, false, p1, p2, p2, T(1), ws.G + (long)p1 * kOB, kOB, T22, kOB,
                     T(0), X, kOB));
      DH_TRY(gemm<T>(s, false, false, p1, p2, p1, T(-1), ws.Tm, kOB, X, kOB, T(0),
                     ws.Tm + (long)p1 * kOB, kOB));
    }
    T* zr = z + (j0 + shift);
    const int skz = std::max(1, std::min(16, m / 1024));
    DH_TRY(gemm<T>(s, true, false, pw, ncols, m, T(1), ws.V, m, zr, ldz, T(0), ws.W1, kOB, false,
                   kKFull, nullptr, skz));
    DH_TRY(gemm<T>(s, false, false, pw, ncols, pw, T(1), ws.Tm, kOB, ws.W1, kOB, T(0), ws.W2,
                   kOB));
    DH_TRY(gemm<T>(s, false, false, m, ncols, pw, T(-1), ws.V, m, ws.W2, kOB, T(1), zr, ldz));
  }
  return 0;
}

int sytrd_lower(hipStream_t s, double* a, long lda, int n, double* d, double* e, double* tau,
                double* work, size_t workSize) {
  return sytrd_lower_t<double>(s, a, lda, n, d, e, tau, work, workSize * sizeof(double));
}
int sytrd_lower_batched(hipStream_t s, double* a, long lda, long sA, int n, double* d, double* e,
                        double* tau, long sV, void* work, size_t wsStride, int K) {
  return sytrd_lower_batched_t<double>(s, a, lda, sA, n, d, e, tau, sV, work, wsStride, K);
}
int sytrd_lower_f32(hipStream_t s, float* a, long lda, int n, float* d, float* e, float* tau,
                    void* work, size_t workBytes) {
  return sytrd_lower_t<float>(s, a, lda, n, d, e, tau, work, workBytes);
}
// ---------------------------------------------------------------------------------------------
// Autotuner of the SYMV tile count: the tridiagonalisation of a synthetic symmetric n x n matrix
// is run once per tile option with every SYMV launch timed by HIP events; launches are grouped by
// trailing size in buckets of kTuneBucket rows and each bucket takes the option of least summed
// time, switching away from the option of the bucket below only for a gain of kTuneGain. The
// result is installed as the tile table of the arithmetic type.
// ---------------------------------------------------------------------------------------------
constexpr int kTuneBucket = 256;
constexpr int kTuneOptions[4] = {2, 4, 8, 16};
constexpr double kTuneGain = 0.02;

// lower triangle of a symmetric matrix from a hash of (i, j): values in [-1, 1), diagonal
// shifted by i / n
template <typename T>
__global__ void synth_sym_kernel(T* a, long lda, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n || i < j) return;
  uint32_t h = (uint32_t)i * 0x9E3779B1u ^ ((uint32_t)j * 0x85EBCA77u + 0x165667B1u);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  double v = (double)(h >> 8) * (1.0 / 8388608.0) - 1.0;
  if (i == j) v += (double)i / n;
  a[i + (long)j * lda] = (T)v;
}

template <typename T>
static int symv_autotune_t(hipStream_t s, int n, std::vector<std::pair<int, int>>& rows,
                           std::vector<double>* detail) {
  DH_CHECK(n >= kTuneBucket && n <= 65535, "autotune: n must be in [256, 65535]");
  DevBuf<T> a, d, e, tau;
  DevBuf<char> work;
  DH_TRY(a.alloc((size_t)n * n));
  DH_TRY(d.alloc(n));
  DH_TRY(e.alloc(n));
  DH_TRY(tau.alloc(n));
  const size_t wb = sytrd_layout<T>(n, nullptr, nullptr);
  DH_TRY(work.alloc(wb));
  const int nb = (n + kTuneBucket - 1) / kTuneBucket;
  std::vector<double> t(4 * (size_t)nb, 0.0);
  std::vector<std::pair<int, float>> rec;
  SymvProfile& pf = g_symvProf;
  const bool on0 = pf.on;
  const double l0 = pf.launches, s0 = pf.seconds, b0 = pf.algBytes;
  const int tiles0 = g_symvTiles;
  int rc = 0;
  for (int o = 0; o < 4 && rc == 0; ++o) {
    g_symvTiles = kTuneOptions[o];
    ++g_tileGen;
    // untimed pass on a small matrix first: the option's kernels are loaded before timing
    for (int nn : {kTuneBucket, n}) {
      hipLaunchKernelGGL(synth_sym_kernel<T>, dim3(cdiv(nn, 256), nn), dim3(256), 0, s, a.p,
                         (long)nn, nn);
      rc = hipGetLastError() == hipSuccess ? 0 : kErrDevice;
      if (rc) {
        set_error("autotune: launch of the test matrix kernel failed");
        break;
      }
      pf.on = nn == n;
      pf.sink = nn == n ? &rec : nullptr;
      rec.clear();
      rc = sytrd_lower_t<T>(s, a.p, nn, nn, d.p, e.p, tau.p, work.p, wb);
      pf.sink = nullptr;
      if (rc) break;
    }
    for (const auto& r : rec) t[(size_t)o * nb + r.first / kTuneBucket] += 1e-3 * r.second;
  }
  g_symvTiles = tiles0;
  ++g_tileGen;
  pf.on = on0;
  pf.launches = l0;
  pf.seconds = s0;
  pf.algBytes = b0;
  pf.sink = nullptr;
  if (rc) return rc;
  DH_HIP(hipStreamSynchronize(s));
  // a bucket keeps the option of the bucket below unless another is kTuneGain faster there
  // (buckets of small trailing sizes are launch bound and differ by noise only)
  rows.clear();
  int cur = -1;
  for (int b = 0; b < nb; ++b) {
    int best = 0;
    for (int o = 1; o < 4; ++o)
      if (t[(size_t)o * nb + b] < t[(size_t)best * nb + b]) best = o;
    if (cur < 0 || t[(size_t)best * nb + b] < (1.0 - kTuneGain) * t[(size_t)cur * nb + b]) {
      if (best != cur) rows.push_back({b * kTuneBucket, kTuneOptions[best]});
      cur = best;
    }
  }
  rows.front().first = 0;
  if (detail) *detail = t;
  g_tileTable[std::is_same<T, float>::value ? 1 : 0] = rows;
  ++g_tileGen;
  return 0;
}

int symv_autotune(hipStream_t s, int n, int precision, std::vector<std::pair<int, int>>& rows,
                  std::vector<double>* detail) {
  DH_CHECK(precision == 0 || precision == 1, "autotune: precision must be 0 (FP64) or 1 (FP32)");
  return precision == 0 ? symv_autotune_t<double>(s, n, rows, detail)
                        : symv_autotune_t<float>(s, n, rows, detail);
}

int ormtr_lower(hipStream_t s, const double* a, long lda, int n, const double* tau, double* z,
                long ldz, int ncols, double* work, size_t workSize) {
  return ormtr_lower_t<double>(s, a, lda, n, tau, z, ldz, ncols, work, workSize * sizeof(double),
                               1, n - 1);
}
int ormtr_lower_f32(hipStream_t s, const float* a, long lda, int n, const float* tau, float* z,
                    long ldz, int ncols, void* work, size_t workBytes) {
  return ormtr_lower_t<float>(s, a, lda, n, tau, z, ldz, ncols, work, workBytes, 1, n - 1);
}
// back-transformation with the reflectors of the band reduction (unit element `shift` rows
// below the column, nref reflectors)
template <typename T>
int ormqr_shift(hipStream_t s, const T* a, long lda, int n, const T* tau, T* z, long ldz,
                int ncols, void* work, size_t workBytes, int shift, int nref) {
  return ormtr_lower_t<T>(s, a, lda, n, tau, z, ldz, ncols, work, workBytes, shift, nref);
}
template int ormqr_shift<double>(hipStream_t, const double*, long, int, const double*, double*,
                                 long, int, void*, size_t, int, int);
template int ormqr_shift<float>(hipStream_t, const float*, long, int, const float*, float*, long,
                                int, void*, size_t, int, int);

}  // namespace dftbhip
