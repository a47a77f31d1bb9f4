// Shared declarations of the dftbhip library: error state, device buffers, launch helpers and
// the internal (C++) interfaces between the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace dftbhip {

// ------------------------------------------------------------------------------------------
// Error handling: every entry point returns 0 on success, a negative code on argument/device
// errors and a LAPACK-style positive info for solver failures. The message of the last error is
// kept per thread and returned by dftbhip_last_error().
// ------------------------------------------------------------------------------------------
void set_error(const std::string& msg);
const std::string& last_error();

constexpr int kErrArg = -1;
constexpr int kErrDevice = -2;
constexpr int kErrState = -3;

#define DH_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t err_ = (call);                                                               \
    if (err_ != hipSuccess) {                                                               \
      (void)hipGetLastError();                                                              \
      ::dftbhip::set_error(std::string(#call) + ": " + hipGetErrorString(err_) + " at " +    \
                           __FILE__ + ":" + std::to_string(__LINE__));                     \
      return ::dftbhip::kErrDevice;                                                         \
    }                                                                                       \
  } while (0)

#define DH_CHECK(cond, msg)                                                                 \
  do {                                                                                      \
    if (!(cond)) {                                                                          \
      ::dftbhip::set_error(msg);                                                            \
      return ::dftbhip::kErrArg;                                                            \
    }                                                                                       \
  } while (0)

#define DH_TRY(expr)                                                                        \
  do {                                                                                      \
    int rc_ = (expr);                                                                       \
    if (rc_ != 0) return rc_;                                                               \
  } while (0)

#define DH_LAUNCH_CHECK()                                                                   \
  do {                                                                                      \
    hipError_t err_ = hipGetLastError();                                                    \
    if (err_ != hipSuccess) {                                                               \
      ::dftbhip::set_error(std::string("kernel launch: ") + hipGetErrorString(err_) + " at " + \
                           __FILE__ + ":" + std::to_string(__LINE__));                     \
      return ::dftbhip::kErrDevice;                                                         \
    }                                                                                       \
  } while (0)

// Owning device allocation (hipMalloc / hipFree).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count) {
    if (count <= n && p) return 0;
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      set_error(std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) + " bytes: " +
                hipGetErrorString(e));
      return kErrDevice;
    }
    n = count;
    return 0;
  }
  int upload(const T* h, size_t count, hipStream_t s) {
    DH_TRY(alloc(count));
    DH_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    return 0;
  }
};

inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// A fixed set of HIP events, destroyed with the owner (also on early error returns).
template <int N>
struct Events {
  hipEvent_t e[N] = {};
  int create() {
    for (auto& x : e) DH_HIP(hipEventCreate(&x));
    return 0;
  }
  Events() = default;
  Events(const Events&) = delete;
  Events& operator=(const Events&) = delete;
  ~Events() {
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  }
  hipEvent_t& operator[](int i) { return e[i]; }
};

// ------------------------------------------------------------------------------------------
// GEMM core (gemm.hip):  C = alpha * op(A) * diag(w) * op(B) + beta * C, column-major.
// ------------------------------------------------------------------------------------------
enum GemmKLimit {
  kKFull = 0,
  kKALower = 1,   // op(A)(i,k) == 0 for k > i  -> k < row tile end
  kKAUpper = 2,   // op(A)(i,k) == 0 for k < i  -> k >= row tile start
  kKBUpper = 3,   // op(B)(k,j) == 0 for k > j  -> k < column tile end
  kKBLower = 4,   // op(B)(k,j) == 0 for k < j  -> k >= column tile start
};

template <typename T>
struct GemmDesc {
  int M, N, K;
  const T* A;
  long lda;
  const T* B;
  long ldb;
  T* C;
  long ldc;
};

template <typename T>
int gemm(hipStream_t s, bool ta, bool tb, int M, int N, int K, T alpha, const T* A, long lda,
         const T* B, long ldb, T beta, T* C, long ldc, bool lowerOut = false,
         GemmKLimit klim = kKFull, const T* w = nullptr, int splitK = 1);

// C = alpha op(A) op(B) (n x n, K) on the 128 x 128 diagonal tiles and the `band` tile rows
// below them only: element (i, j) with i / 128 - j / 128 in [0, band] is written, the rest of C
// is untouched (splitK > 1: K split over that many workgroups per tile, C first cleared whole)
template <typename T>
int gemm_band(hipStream_t s, bool ta, bool tb, int n, int K, T alpha, const T* A, long lda,
              const T* B, long ldb, T* C, long ldc, int band, int splitK = 1);
constexpr int kGemmTile = 128;
// the same product on the 128 x 128 tiles of C in tiles[0 .. ntiles-1] = (tile row, tile
// column) only (w: per-k weights as in gemm); the rest of C is untouched
template <typename T>
int gemm_tiles(hipStream_t s, bool ta, bool tb, int M, int N, int K, T alpha, const T* A,
               long lda, const T* B, long ldb, T beta, T* C, long ldc, const int2* tiles,
               int ntiles, const T* w = nullptr);

// tile order of full-output GEMMs: 0 the launch grid's, 1 XCD-grouped (the default)
int set_gemm_order(int order);

// FP16 product with FP32 accumulation: C = alpha diag(rowScale) A^T B (rowScale may be null),
// A (K x M) and B (K x N) column-major FP16 with K contiguous. The operands are padded: lda and
// ldb multiples of kF16LdMult and >= f16_ld(K), at least f16_cols(M) (f16_cols(N)) columns, the
// padding (rows K .. f16_ld(K) - 1 and the columns past M or N) zero.
constexpr int kF16LdMult = 64, kF16ColMult = 128;
inline long f16_ld(long k) { return (k + kF16LdMult - 1) / kF16LdMult * kF16LdMult; }
inline long f16_cols(long m) { return (m + kF16ColMult - 1) / kF16ColMult * kF16ColMult; }
int gemm_f16_tn(hipStream_t s, int M, int N, int K, float alpha, const _Float16* A, long lda,
                const _Float16* B, long ldb, float* C, long ldc, const float* rowScale);

// Batched GEMM over device-resident descriptors; maxM/maxN bound the tile grid.
template <typename T>
int gemm_batched(hipStream_t s, bool ta, bool tb, const GemmDesc<T>* dDesc, int nBatch, int maxM,
                 int maxN, T alpha, T beta);

// ------------------------------------------------------------------------------------------
// Dense helpers (dense.hip)
// ------------------------------------------------------------------------------------------
template <typename T>
int fill(hipStream_t s, T* p, size_t n, T v);
template <typename T>
int fill2d(hipStream_t s, T* p, long ld, int m, int n, T v);
template <typename T>
int copy2d(hipStream_t s, const T* src, long lds, T* dst, long ldd, int m, int n);
// dst(:, j) = src(:, j) * sqrt(w(j)) for w >= 0 (the sqrt(f)-scaled columns of a SYRK)
template <typename T>
int scale_cols_sqrt(hipStream_t s, const T* src, long lds, const T* w, T* dst, long ldd, int m,
                    int n);
// dst(:, j) = src(:, j) * w(j)
template <typename T>
int scale_cols(hipStream_t s, const T* src, long lds, const T* w, T* dst, long ldd, int m, int n);
template <typename Tin, typename Tout>
int convert2d(hipStream_t s, const Tin* src, long lds, Tout* dst, long ldd, int m, int n);
// interleaved complex (re, im) matrix z (leading dimension ldz in complex elements) to the planes
// re, im (ld) and back; merge mode 0: every element, 1: the lower triangle and zeros above it,
// 2: the lower triangle only (the strict upper part of z untouched)
int split_complex(hipStream_t s, const double* z, long ldz, double* re, double* im, long ld,
                  int m, int n);
int merge_complex(hipStream_t s, const double* re, const double* im, long ld, double* z, long ldz,
                  int m, int n, int mode);
// dst(j, i) = src(i, j) for the m x n matrix src (dst is n x m), converted to Tout
template <typename Tin, typename Tout>
int transpose2d(hipStream_t s, const Tin* src, long lds, Tout* dst, long ldd, int m, int n);
// FP16 copies of the m x n matrix x, or of the residual x - diag(lam) y (y given), rows scaled
// by rsc (given): same(i, j) (ld1) and/or the transpose tr(j, i) (ld2); either may be null
int f16_convert(hipStream_t s, const double* x, const double* y, const double* lam,
                const double* rsc, long ldx, int m, int n, _Float16* same, long ld1, _Float16* tr,
                long ld2);
// dst = a + b with b in FP32 (m x n)
int add_f32(hipStream_t s, const double* a, long lda, const float* b, long ldb, double* dst,
            long ldd, int m, int n);
int symmetrize_lower(hipStream_t s, double* a, long lda, int n);
int symmetrize_lower(hipStream_t s, float* a, long lda, int n);
int zero_strict_upper(hipStream_t s, double* a, long lda, int n);


// ------------------------------------------------------------------------------------------
// Symmetric eigensolver pieces (sytrd.hip, stedc.hip)
// ------------------------------------------------------------------------------------------
int sytrd_lower(hipStream_t s, double* a, long lda, int n, double* d, double* e, double* tau,
                double* work, size_t workSize);
size_t sytrd_work_size(int n);
// K problems of size n in one pass (matrices at a + k sA, vectors at + k sV, workspace blocks at
// work + k wsStride bytes, wsStride >= sytrd_work_bytes(n, 8) and a multiple of 256)
int sytrd_lower_batched(hipStream_t s, double* a, long lda, long sA, int n, double* d, double* e,
                        double* tau, long sV, void* work, size_t wsStride, int K);
// per-launch timing of the tridiagonalisation's symmetric product: launches, seconds, bytes
void symv_profiling(bool on);
void symv_profile_stats(double* out);
int ormtr_lower(hipStream_t s, const double* a, long lda, int n, const double* tau, double* z,
                long ldz, int ncols, double* work, size_t workSize);
size_t ormtr_work_size(int n, int ncols);
// FP32 variants (workspace in bytes)
int sytrd_lower_f32(hipStream_t s, float* a, long lda, int n, float* d, float* e, float* tau,
                    void* work, size_t workBytes);
int ormtr_lower_f32(hipStream_t s, const float* a, long lda, int n, const float* tau, float* z,
                    long ldz, int ncols, void* work, size_t workBytes);
size_t sytrd_work_bytes(int n, int elemSize);
int set_symv_tiles(int tiles);   // row tiles per SYMV workgroup: 2, 4, 8, 16 or 0 (by size)
// capture the tridiagonalisation's launch sequence in a HIP graph and replay it for repeated
// calls on the same buffers (off by default); release_sytrd_graphs frees the kept graphs
int set_sytrd_graphs(bool on);
void release_sytrd_graphs();
// two-stage tridiagonalisation: bulge chasing on one workgroup of an ordinary launch (the path
// taken when the device refuses the cooperative launch) instead of the cooperative grid
int set_sb2st_single(bool on);
// autotuned tile tables of the SYMV (FP64: precision 0, FP32: 1), rows (first size, tiles);
// detail: summed seconds per option (2, 4, 8, 16) and 256-row bucket of trailing sizes
int symv_autotune(hipStream_t s, int n, int precision, std::vector<std::pair<int, int>>& rows,
                  std::vector<double>* detail);
int clear_symv_tables();
// two-stage tridiagonalisation (band.hip): full symmetric input, workspace in bytes
size_t band_work_bytes(int n, int elemSize);
int sytrd_2stage(hipStream_t s, double* a, long lda, int n, double* d, double* e, void* work,
                 size_t workBytes);
int sytrd_2stage_f32(hipStream_t s, float* a, long lda, int n, float* d, float* e, void* work,
                     size_t workBytes);
// Z <- Q Z; work is the tridiagonalisation's workspace, work2 that of ormtr (n, ncols)
int ormtr_2stage(hipStream_t s, const double* a, long lda, int n, double* z, long ldz, int ncols,
                 void* work, size_t workBytes, void* work2, size_t work2Bytes);
int ormtr_2stage_f32(hipStream_t s, const float* a, long lda, int n, float* z, long ldz,
                     int ncols, void* work, size_t workBytes, void* work2, size_t work2Bytes);
size_t ormtr_work_bytes(int n, int ncols, int elemSize);
// ws: device buffers kept between calls (null: allocated for this call)
struct StedcWs;
StedcWs* stedc_ws_new();
void stedc_ws_delete(StedcWs* w);
int stedc(hipStream_t s, int n, const double* dHost, const double* eHost, double* w, double* z,
          long ldz, double* work, size_t workSize, int* info, StedcWs* ws = nullptr);
size_t stedc_work_size(int n);

}  // namespace dftbhip
