// MFMA GEMM core for CDNA4 (gfx950), FP64 (v_mfma_f64_16x16x4f64) and FP32
// (v_mfma_f32_16x16x4f32), and an FP16 product with FP32 accumulation
// (v_mfma_f32_16x16x32_f16) for the correction steps of the eigenpair refinement.
//
//   C = alpha * op(A) * diag(w) * op(B) + beta * C        (column-major, LAPACK convention)
//
// Workgroup tile 128x128, 256 threads = 4 wave64s in a 2x2 arrangement, each wave owns a 64x64
// sub-tile = 4x4 MFMA blocks of 16x16. K is staged through LDS in steps of 16 with a register
// prefetch of the next step. Options used by the eigensolver and density-matrix builds:
//  * lowerOut: only tiles on/below the diagonal are computed (SYRK/SYR2K-like outputs);
//  * klim: skip the zero part of a triangular operand;
//  * w: per-k weights folded into op(A) (C = X diag(f) X^T for the density matrices).
#include "common.h"
#include "mfma.h"

namespace dftbhip {

constexpr int kBM = 128, kBN = 128, kBK = 16, kNT = 256, kPad = 1;
static_assert(kBM == kGemmTile && kBN == kGemmTile, "band tiles");

// tile order of full-output GEMMs: 0 the launch grid's (x = tile row), 1 XCD-grouped
static int g_gemmOrder = 1;
int set_gemm_order(int order) {
  DH_CHECK(order == 0 || order == 1, "gemm order: 0 or 1");
  g_gemmOrder = order;
  return 0;
}

template <typename T, bool TA, bool TB>
__device__ __forceinline__ void gemm_tile(int tm, int tn, int M, int N, int K, T alpha,
                                          const T* __restrict__ A, long lda,
                                          const T* __restrict__ B, long ldb, T beta,
                                          T* __restrict__ C, long ldc, int klim,
                                          const T* __restrict__ w, int kz = 0, int nkz = 1) {
  using acc_t = typename Mfma<T>::acc_t;
  // LDS: A and B staging during the K loop, reused for the output staging afterwards
  constexpr int kStage = 2 * kBK * (kBM + kPad);
  constexpr int kOut = 4 * 16 * (64 + 1);
  __shared__ T smem[kStage > kOut ? kStage : kOut];
  T (*As)[kBM + kPad] = reinterpret_cast<T (*)[kBM + kPad]>(smem);
  T (*Bs)[kBN + kPad] = reinterpret_cast<T (*)[kBN + kPad]>(smem + kBK * (kBM + kPad));
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int i0 = tm * kBM, j0 = tn * kBN;
  int kbeg = 0, kend = K;
  if (klim == kKALower) kend = min(K, i0 + kBM);
  else if (klim == kKAUpper) kbeg = min(K, i0);
  else if (klim == kKBUpper) kend = min(K, j0 + kBN);
  else if (klim == kKBLower) kbeg = min(K, j0);
  // split-K: slice kz of nkz (slices are multiples of the K step); partial sums added atomically
  if (nkz > 1) {
    const int span = kend - kbeg;
    const int chunk = ((span + nkz - 1) / nkz + kBK - 1) / kBK * kBK;
    const int b = kbeg + kz * chunk;
    kend = min(kend, b + chunk);
    kbeg = b;
  }

  acc_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

  T ra[8], rb[8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + q * kNT;
      int r, c;
      if (!TA) { r = e % kBM; c = e / kBM; } else { c = e % kBK; r = e / kBK; }
      const int gi = i0 + r, gk = k0 + c;
      T v = T(0);
      if (gi < M && gk < kend) {
        v = TA ? A[gk + (long)gi * lda] : A[gi + (long)gk * lda];
        if (w) v *= w[gk];
      }
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + q * kNT;
      int r, c;
      if (!TB) { c = e % kBK; r = e / kBK; } else { r = e % kBN; c = e / kBN; }
      const int gj = j0 + r, gk = k0 + c;
      T v = T(0);
      if (gj < N && gk < kend) v = TB ? B[gj + (long)gk * ldb] : B[gk + (long)gj * ldb];
      rb[q] = v;
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + q * kNT;
      int r, c;
      if (!TA) { r = e % kBM; c = e / kBM; } else { c = e % kBK; r = e / kBK; }
      As[c][r] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + q * kNT;
      int r, c;
      if (!TB) { c = e % kBK; r = e / kBK; } else { r = e % kBN; c = e / kBN; }
      Bs[c][r] = rb[q];
    }
  };

  if (kbeg < kend) {
    gload(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += kBK) {
      __syncthreads();
      sstore();
      __syncthreads();
      if (k0 + kBK < kend) gload(k0 + kBK);
#pragma unroll
      for (int kk = 0; kk < kBK / 4; ++kk) {
        T fa[4], fb[4];
        const int kr = kk * 4 + (lane >> 4);
#pragma unroll
        for (int m = 0; m < 4; ++m) fa[m] = As[kr][wm * 64 + m * 16 + (lane & 15)];
#pragma unroll
        for (int n = 0; n < 4; ++n) fb[n] = Bs[kr][wn * 64 + n * 16 + (lane & 15)];
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int n = 0; n < 4; ++n) acc[m][n] = Mfma<T>::op(fa[m], fb[n], acc[m][n]);
      }
    }
  }
  // Output layout of the 16x16 accumulator, column l%16 for lane l; element r sits in row
  // l/16 + 4r for FP64 and in row 4(l/16) + r for FP32 (pinned by the layout probes below).
  // Each wave stages 16 columns x 64 rows of its sub-tile in its own LDS area and writes them
  // back one column per store instruction (64 consecutive rows). The area is private to the
  // wave, so only the first staging waits for the other waves (their K loop reads the same
  // LDS); within the wave an LDS-count wait orders the staging before the reads. The 16 C
  // values a lane updates are loaded before any of its stores.
  __syncthreads();
  T* Cs = smem + wave * 16 * 65;
  auto wave_lds_sync = [&]() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0); vmcnt, expcnt untouched (gfx9 encoding)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  };
  const int gi = i0 + wm * 64 + lane;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int gj0 = j0 + wn * 64 + n * 16;
    T cold[16];
    if (beta != T(0) && nkz == 1) {
#pragma unroll
      for (int cc = 0; cc < 16; ++cc)
        cold[cc] = (gi < M && gj0 + cc < N) ? C[gi + (long)(gj0 + cc) * ldc] : T(0);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = Mfma<T>::kRowMajorQuad ? (lane >> 4) * 4 + r : (lane >> 4) + 4 * r;
        Cs[(lane & 15) * 65 + m * 16 + rr] = acc[m][n][r];
      }
    wave_lds_sync();
    if (gi < M) {
#pragma unroll
      for (int cc = 0; cc < 16; ++cc) {
        const int gj = gj0 + cc;
        if (gj < N) {
          T v = alpha * Cs[cc * 65 + lane];
          T* dst = C + gi + (long)gj * ldc;
          if (nkz > 1) {
            atomicAdd(dst, v);
          } else {
            if (beta != T(0)) v += beta * cold[cc];
            *dst = v;
          }
        }
      }
    }
    wave_lds_sync();   // this group's reads before the next group's staging
  }
}

__device__ __forceinline__ void lower_tile_index(int x, int& tm, int& tn) {
  int t = (int)((sqrt(8.0 * (double)x + 1.0) - 1.0) * 0.5);
  while ((t + 1) * (t + 2) / 2 <= x) ++t;
  while (t * (t + 1) / 2 > x) --t;
  tm = t;
  tn = x - t * (t + 1) / 2;
}

// XCD-grouped tile order of a full output: workgroup IDs are dealt to the 8 XCDs round-robin,
// so the IDs of one XCD (congruent mod 8) are given a contiguous range of a tile sequence that
// walks groups of kGroupRows tile rows column by column. The tiles an XCD runs at one time then
// share a few A row panels and B column panels in its L2, instead of every tile of the chip
// reading its own pair of panels. nPad = tiles rounded up to a multiple of 8; IDs past the last
// tile return.
constexpr int kXcds = 8, kGroupRows = 16;
__device__ __forceinline__ bool grouped_tile_index(int w, int gm, int gn, int nPad, int& tm,
                                                   int& tn) {
  const int per = nPad / kXcds;
  const int o = (w % kXcds) * per + w / kXcds;
  if (o >= gm * gn) return false;
  const int band = gn * kGroupRows;
  const int g = o / band, rem = o % band;
  const int rows = min(kGroupRows, gm - g * kGroupRows);
  tn = rem / rows;
  tm = g * kGroupRows + rem % rows;
  return true;
}

template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void gemm_kernel(int M, int N, int K, T alpha, const T* A,
                                                   long lda, const T* B, long ldb, T beta, T* C,
                                                   long ldc, int lowerOut, int klim, const T* w,
                                                   int grouped, int band,
                                                   const int2* __restrict__ tileList) {
  int tm, tn;
  if (tileList) {
    const int2 t = tileList[blockIdx.x];
    tm = t.x;
    tn = t.y;
  } else if (band >= 0) {
    // the diagonal tiles and the band tile rows below them: tile (tn + b, tn), b = 0 .. band
    tn = (int)blockIdx.x / (band + 1);
    tm = tn + (int)blockIdx.x % (band + 1);
    if (tm * kBM >= M) return;
  } else if (lowerOut) {
    lower_tile_index(blockIdx.x, tm, tn);
  } else if (grouped) {
    const int gm = (M + kBM - 1) / kBM, gn = (N + kBN - 1) / kBN;
    if (!grouped_tile_index((int)blockIdx.x, gm, gn, (int)gridDim.x, tm, tn)) return;
  } else {
    tm = blockIdx.x;
    tn = blockIdx.y;
  }
  gemm_tile<T, TA, TB>(tm, tn, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, klim, w,
                       (int)blockIdx.z, (int)gridDim.z);
}

template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void gemm_batched_kernel(const GemmDesc<T>* desc, T alpha,
                                                           T beta) {
  const GemmDesc<T> d = desc[blockIdx.z];
  const int tm = blockIdx.x, tn = blockIdx.y;
  if (tm * kBM >= d.M || tn * kBN >= d.N) return;
  gemm_tile<T, TA, TB>(tm, tn, d.M, d.N, d.K, alpha, d.A, d.lda, d.B, d.ldb, beta, d.C, d.ldc,
                       kKFull, nullptr);
}

template <typename T>
int gemm(hipStream_t s, bool ta, bool tb, int M, int N, int K, T alpha, const T* A, long lda,
         const T* B, long ldb, T beta, T* C, long ldc, bool lowerOut, GemmKLimit klim,
         const T* w, int splitK) {
  if (M <= 0 || N <= 0) return 0;
  DH_CHECK(!lowerOut || M == N, "gemm: lower output needs a square C");
  if (splitK > 1) {
    DH_CHECK(beta == T(0) && !lowerOut, "gemm: split-K needs beta = 0 and a full output");
    splitK = std::min<int>(splitK, std::max(1, K / kBK));
    DH_TRY(fill2d<T>(s, C, ldc, M, N, T(0)));
  }
  splitK = std::max(splitK, 1);
  dim3 grid;
  // grouped only for full-K products: with a triangular k-limit the work of a tile row grows
  // or shrinks along the rows, and XCDs given contiguous bands of rows would be unbalanced
  const bool grouped = !lowerOut && klim == kKFull && g_gemmOrder == 1;
  if (lowerOut) {
    const long t = cdiv(M, kBM);
    grid = dim3((unsigned)(t * (t + 1) / 2), 1, 1);
  } else if (grouped) {
    const long tiles = (long)cdiv(M, kBM) * cdiv(N, kBN);
    grid = dim3((unsigned)((tiles + kXcds - 1) / kXcds * kXcds), 1, splitK);
  } else {
    grid = dim3(cdiv(M, kBM), cdiv(N, kBN), splitK);
  }
#define DH_GEMM_CASE(TA_, TB_)                                                           \
  hipLaunchKernelGGL((gemm_kernel<T, TA_, TB_>), grid, dim3(kNT), 0, s, M, N, K, alpha, A, \
                     lda, B, ldb, beta, C, ldc, (int)lowerOut, (int)klim, w, (int)grouped, -1, \
                     (const int2*)nullptr)
  if (!ta && !tb) DH_GEMM_CASE(false, false);
  else if (!ta && tb) DH_GEMM_CASE(false, true);
  else if (ta && !tb) DH_GEMM_CASE(true, false);
  else DH_GEMM_CASE(true, true);
#undef DH_GEMM_CASE
  DH_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int gemm_band(hipStream_t s, bool ta, bool tb, int n, int K, T alpha, const T* A, long lda,
              const T* B, long ldb, T* C, long ldc, int band, int splitK) {
  if (n <= 0) return 0;
  DH_CHECK(band >= 0, "gemm_band: negative band");
  const long t = cdiv(n, kBM);
  // few tiles with a long K: split K so that the chip fills (partial sums added atomically)
  splitK = std::max(1, std::min(splitK, K / (4 * kBK)));
  if (splitK > 1) DH_TRY(fill2d<T>(s, C, ldc, n, n, T(0)));
  const dim3 grid((unsigned)(t * (band + 1)), 1, (unsigned)splitK);
#define DH_GEMM_BAND(TA_, TB_)                                                           \
  hipLaunchKernelGGL((gemm_kernel<T, TA_, TB_>), grid, dim3(kNT), 0, s, n, n, K, alpha, A, \
                     lda, B, ldb, T(0), C, ldc, 0, (int)kKFull, (const T*)nullptr, 0, band,   \
                     (const int2*)nullptr)
  if (!ta && !tb) DH_GEMM_BAND(false, false);
  else if (!ta && tb) DH_GEMM_BAND(false, true);
  else if (ta && !tb) DH_GEMM_BAND(true, false);
  else DH_GEMM_BAND(true, true);
#undef DH_GEMM_BAND
  DH_LAUNCH_CHECK();
  return 0;
}
template int gemm_band<double>(hipStream_t, bool, bool, int, int, double, const double*, long,
                               const double*, long, double*, long, int, int);

template <typename T>
int gemm_tiles(hipStream_t s, bool ta, bool tb, int M, int N, int K, T alpha, const T* A,
               long lda, const T* B, long ldb, T beta, T* C, long ldc, const int2* tiles,
               int ntiles, const T* w) {
  if (M <= 0 || N <= 0 || ntiles <= 0) return 0;
  DH_CHECK(tiles, "gemm_tiles: null tile list");
#define DH_GEMM_TILES(TA_, TB_)                                                              \
  hipLaunchKernelGGL((gemm_kernel<T, TA_, TB_>), dim3((unsigned)ntiles), dim3(kNT), 0, s, M, N, \
                     K, alpha, A, lda, B, ldb, beta, C, ldc, 0, (int)kKFull, w, 0, -1, tiles)
  if (!ta && !tb) DH_GEMM_TILES(false, false);
  else if (!ta && tb) DH_GEMM_TILES(false, true);
  else if (ta && !tb) DH_GEMM_TILES(true, false);
  else DH_GEMM_TILES(true, true);
#undef DH_GEMM_TILES
  DH_LAUNCH_CHECK();
  return 0;
}
template int gemm_tiles<double>(hipStream_t, bool, bool, int, int, int, double, const double*,
                                long, const double*, long, double, double*, long, const int2*,
                                int, const double*);
template int gemm_tiles<float>(hipStream_t, bool, bool, int, int, int, float, const float*, long,
                               const float*, long, float, float*, long, const int2*, int,
                               const float*);

template <typename T>
int gemm_batched(hipStream_t s, bool ta, bool tb, const GemmDesc<T>* dDesc, int nBatch, int maxM,
                 int maxN, T alpha, T beta) {
  if (nBatch <= 0 || maxM <= 0 || maxN <= 0) return 0;
  dim3 grid(cdiv(maxM, kBM), cdiv(maxN, kBN), nBatch);
#define DH_GEMMB_CASE(TA_, TB_)                                                            \
  hipLaunchKernelGGL((gemm_batched_kernel<T, TA_, TB_>), grid, dim3(kNT), 0, s, dDesc, alpha, \
                     beta)
  if (!ta && !tb) DH_GEMMB_CASE(false, false);
  else if (!ta && tb) DH_GEMMB_CASE(false, true);
  else if (ta && !tb) DH_GEMMB_CASE(true, false);
  else DH_GEMMB_CASE(true, true);
#undef DH_GEMMB_CASE
  DH_LAUNCH_CHECK();
  return 0;
}

// ------------------------------------------------------------------------------------------
// FP16 product C = alpha diag(rs) A^T B (FP32 accumulation and output), A (K x M) and B (K x N)
// column-major with K contiguous, so that every MFMA operand (8 consecutive k of one row or
// column) is one 16-byte load. The operands are padded (f16_ld, f16_cols): no bounds checks in
// the K loop. Workgroup tile 128 x 128, 4 waves of 64 x 64 (4 x 4 MFMA blocks of 16 x 16), K
// staged through LDS in steps of 64 with a register prefetch of the next step; rows of the LDS
// tiles padded by 16 bytes so that the 16 rows one operand load reads fall on distinct banks.
// ------------------------------------------------------------------------------------------
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
constexpr int kHB = 128, kHK = 64, kHRow = kHK + 8;

__device__ __forceinline__ void wave_lds_wait() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0); vmcnt, expcnt untouched (gfx9 encoding)
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__global__ __launch_bounds__(256) void gemm_f16_tn_kernel(int M, int N, int Kp, float alpha,
                                                          const _Float16* __restrict__ A, long lda,
                                                          const _Float16* __restrict__ B, long ldb,
                                                          float* __restrict__ C, long ldc,
                                                          const float* __restrict__ rs, int gm,
                                                          int gn) {
  int tm, tn;
  if (!grouped_tile_index((int)blockIdx.x, gm, gn, (int)gridDim.x, tm, tn)) return;
  __shared__ __attribute__((aligned(16))) _Float16 smem[2 * kHB * kHRow];
  _Float16* As = smem;
  _Float16* Bs = smem + kHB * kHRow;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const long i0 = (long)tm * kHB, j0 = (long)tn * kHB;
  v4f acc[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[m][n] = v4f{0.f, 0.f, 0.f, 0.f};
  // 16-byte chunk c = tid + 256 q of a tile: row c / 8, k offset 8 (c % 8)
  v8h ra[4], rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = tid + 256 * q, row = c >> 3, kc = (c & 7) * 8;
      ra[q] = *reinterpret_cast<const v8h*>(A + (i0 + row) * lda + k0 + kc);
      rb[q] = *reinterpret_cast<const v8h*>(B + (j0 + row) * ldb + k0 + kc);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = tid + 256 * q, row = c >> 3, kc = (c & 7) * 8;
      *reinterpret_cast<v8h*>(As + row * kHRow + kc) = ra[q];
      *reinterpret_cast<v8h*>(Bs + row * kHRow + kc) = rb[q];
    }
  };
  gload(0);
  for (int k0 = 0; k0 < Kp; k0 += kHK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + kHK < Kp) gload(k0 + kHK);
#pragma unroll
    for (int kk = 0; kk < kHK / 32; ++kk) {
      const int ko = kk * 32 + (lane >> 4) * 8;
      v8h fa[4], fb[4];
#pragma unroll
      for (int m = 0; m < 4; ++m)
        fa[m] = *reinterpret_cast<const v8h*>(As + (wm * 64 + m * 16 + (lane & 15)) * kHRow + ko);
#pragma unroll
      for (int n = 0; n < 4; ++n)
        fb[n] = *reinterpret_cast<const v8h*>(Bs + (wn * 64 + n * 16 + (lane & 15)) * kHRow + ko);
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n)
          acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[m], fb[n], acc[m][n], 0, 0, 0);
    }
  }
  // accumulator element r of lane l: D(4 (l / 16) + r, l % 16) of its 16 x 16 block (the FP32
  // layout, pinned by the probe below); staged per wave as 16 columns x 64 rows and written one
  // column per store (64 consecutive rows)
  __syncthreads();
  float* Cs = reinterpret_cast<float*>(smem) + wave * 16 * 65;
  const long gi = i0 + wm * 64 + lane;
  const float sc = alpha * ((rs && gi < M) ? rs[gi] : 1.0f);
#pragma unroll
  for (int n = 0; n < 4; ++n) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) Cs[(lane & 15) * 65 + m * 16 + (lane >> 4) * 4 + r] = acc[m][n][r];
    wave_lds_wait();
    if (gi < M) {
#pragma unroll
      for (int cc = 0; cc < 16; ++cc) {
        const long gj = j0 + wn * 64 + n * 16 + cc;
        if (gj < N) C[gi + gj * ldc] = sc * Cs[cc * 65 + lane];
      }
    }
    wave_lds_wait();
  }
}

int gemm_f16_tn(hipStream_t s, int M, int N, int K, float alpha, const _Float16* A, long lda,
                const _Float16* B, long ldb, float* C, long ldc, const float* rowScale) {
  if (M <= 0 || N <= 0) return 0;
  const long Kp = f16_ld(K);
  DH_CHECK(lda % kF16LdMult == 0 && ldb % kF16LdMult == 0 && lda >= Kp && ldb >= Kp && ldc >= M,
           "gemm_f16_tn: leading dimensions");
  if (Kp == 0) return fill2d<float>(s, C, ldc, M, N, 0.0f);
  const int gm = (int)cdiv(M, kHB), gn = (int)cdiv(N, kHB);
  const long tiles = (long)gm * gn;
  const long grid = (tiles + kXcds - 1) / kXcds * kXcds;
  DH_CHECK(grid < (1L << 31), "gemm_f16_tn: grid too large");
  hipLaunchKernelGGL(gemm_f16_tn_kernel, dim3((unsigned)grid), dim3(256), 0, s, M, N, (int)Kp,
                     alpha, A, lda, B, ldb, C, ldc, rowScale, gm, gn);
  DH_LAUNCH_CHECK();
  return 0;
}

template int gemm<double>(hipStream_t, bool, bool, int, int, int, double, const double*, long,
                          const double*, long, double, double*, long, bool, GemmKLimit,
                          const double*, int);
template int gemm<float>(hipStream_t, bool, bool, int, int, int, float, const float*, long,
                         const float*, long, float, float*, long, bool, GemmKLimit,
                         const float*, int);
template int gemm_batched<double>(hipStream_t, bool, bool, const GemmDesc<double>*, int, int, int,
                                  double, double);
template int gemm_batched<float>(hipStream_t, bool, bool, const GemmDesc<float>*, int, int, int,
                                 float, float);

}  // namespace dftbhip

namespace dftbhip {
// Lane-layout check of the FP64 MFMA accumulator: one wave computes D = A B with
// A(i,k) = i [k==0] + [k==1], B(k,j) = [k==0] + 100 j [k==1], so D(i,j) = i + 100 j, and each
// lane stores its four accumulator elements: out[4*lane + r].
__global__ void mfma_layout_probe_kernel(double* out) {
  const int lane = threadIdx.x;
  const int i = lane & 15, k = lane >> 4;
  const double a = (k == 0) ? (double)i : (k == 1 ? 1.0 : 0.0);
  const int j = lane & 15;
  const double b = (k == 0) ? 1.0 : (k == 1 ? 100.0 * j : 0.0);
  v4d acc = {0, 0, 0, 0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[4 * lane + r] = acc[r];
}

__global__ void mfma_layout_probe_f32_kernel(double* out) {
  const int lane = threadIdx.x;
  const int i = lane & 15, k = lane >> 4;
  const float a = (k == 0) ? (float)i : (k == 1 ? 1.0f : 0.0f);
  const int j = lane & 15;
  const float b = (k == 0) ? 1.0f : (k == 1 ? 100.0f * j : 0.0f);
  v4f acc = {0, 0, 0, 0};
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[256 + 4 * lane + r] = acc[r];
}

// The same for the FP16 16x16x32 MFMA (FP32 accumulator): element e of the A operand of lane l
// is taken as A(l % 16, 8 (l / 16) + e) and of the B operand as B(8 (l / 16) + e, l % 16); a
// product only needs A and B to agree on the k of each (lane, element), which they do here.
__global__ void mfma_layout_probe_f16_kernel(double* out) {
  const int lane = threadIdx.x;
  const int i = lane & 15, j = lane & 15;
  v8h a, b;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = 8 * (lane >> 4) + e;
    a[e] = (_Float16)(k == 0 ? (float)i : (k == 1 ? 1.0f : 0.0f));
    b[e] = (_Float16)(k == 0 ? 1.0f : (k == 1 ? 100.0f * j : 0.0f));
  }
  v4f acc = {0, 0, 0, 0};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[4 * lane + r] = acc[r];
}
}  // namespace dftbhip

extern "C" int dftbhip_probe_mfma_f16(double* out) {
  using namespace dftbhip;
  DH_CHECK(out, "probe: null pointer");
  DevBuf<double> d;
  DH_TRY(d.alloc(256));
  hipLaunchKernelGGL(mfma_layout_probe_f16_kernel, dim3(1), dim3(64), 0, 0, d.p);
  DH_LAUNCH_CHECK();
  DH_HIP(hipMemcpy(out, d.p, 256 * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

// FP16 product on host arrays (FP16 bit patterns): c = alpha a^T b, a (k x m), b (k x n), c
// (m x n), all column-major
extern "C" int dftbhip_gemm_f16_tn(int m, int n, int k, float alpha, const uint16_t* a, int lda,
                                   const uint16_t* b, int ldb, float* c, int ldc) {
  using namespace dftbhip;
  DH_CHECK(m >= 0 && n >= 0 && k >= 0, "gemm_f16_tn: negative size");
  DH_CHECK(a && b && c, "gemm_f16_tn: null pointer");
  DH_CHECK(lda >= std::max(1, k) && ldb >= std::max(1, k) && ldc >= std::max(1, m),
           "gemm_f16_tn: bad leading dimension");
  if (m == 0 || n == 0) return 0;
  const long ldp = std::max<long>(f16_ld(k), kF16LdMult);
  DevBuf<_Float16> A, B;
  DevBuf<float> Cd;
  DH_TRY(A.alloc((size_t)ldp * f16_cols(m)));
  DH_TRY(B.alloc((size_t)ldp * f16_cols(n)));
  DH_TRY(Cd.alloc((size_t)m * n));
  DH_HIP(hipMemset(A.p, 0, A.n * sizeof(_Float16)));
  DH_HIP(hipMemset(B.p, 0, B.n * sizeof(_Float16)));
  if (k > 0) {
    DH_HIP(hipMemcpy2D(A.p, ldp * 2, a, (size_t)lda * 2, (size_t)k * 2, m, hipMemcpyHostToDevice));
    DH_HIP(hipMemcpy2D(B.p, ldp * 2, b, (size_t)ldb * 2, (size_t)k * 2, n, hipMemcpyHostToDevice));
  }
  DH_TRY(gemm_f16_tn(0, m, n, k, alpha, A.p, ldp, B.p, ldp, Cd.p, m, nullptr));
  DH_HIP(hipMemcpy2D(c, (size_t)ldc * 4, Cd.p, (size_t)m * 4, (size_t)m * 4, n,
                     hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int dftbhip_probe_mfma_f64(double* out) {
  using namespace dftbhip;
  DH_CHECK(out, "probe: null pointer");
  DevBuf<double> d;
  DH_TRY(d.alloc(512));
  hipLaunchKernelGGL(mfma_layout_probe_kernel, dim3(1), dim3(64), 0, 0, d.p);
  DH_LAUNCH_CHECK();
  hipLaunchKernelGGL(mfma_layout_probe_f32_kernel, dim3(1), dim3(64), 0, 0, d.p);
  DH_LAUNCH_CHECK();
  DH_HIP(hipMemcpy(out, d.p, 512 * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}
