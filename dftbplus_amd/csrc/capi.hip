// C ABI of the dftbhip library and the resident SCC context.
//
// Stateless drop-ins (sygvd, density matrices, sparse kernels) take host arrays with the
// reference's conventions and stage them through device memory. The context keeps everything
// resident: geometry, neighbour blocks, H0/S, the gamma matrix, the dense matrices and the
// eigensolver workspaces, and runs the SCC cycle of src/dftbp/dftbplus/main.F90:1429-1563 with
// the charge mixing (Broyden, src/dftbp/mixer/broydenmixer.F90:180-342) and the Fermi filling
// (src/dftbp/dftb/etemp.F90:59-455) on the host, everything of O(N^2) or more on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dftbhip.h"
#include "common.h"
#include "batch.h"
#include "neigh.h"
#include "eigen.h"
#include "sparse.h"
#include "zhegv.h"

namespace dftbhip {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const std::string& last_error() { return g_err; }

// Library state of the stateless drop-ins: the device, its stream, the eigensolver workspaces
// and the staging matrices of sygvd and the density matrices, kept between calls (grown on
// demand) and released by dftbhip_finalize. The object is never destroyed: device memory must
// not be freed from a static destructor, which runs after the HIP runtime has shut down.
struct Runtime {
  int device = -1;
  hipStream_t stream = nullptr;
  EigenSolver eig;
  ZhegvSolver zeig;
  DevBuf<double> stageA, stageB, stageC, stageW;
  void release() {
    eig.release();
    zeig.release();
    stageA.release();
    stageB.release();
    stageC.release();
    stageW.release();
  }
};
static Runtime& g_rt = *new Runtime();

// Workspace of the allocation-free entry point dftbhip_dsytrd_noalloc: a fixed device array
// that every GPU the library runs on holds for the life of the process. Host code never hands
// the symbol itself to a kernel (in host code it is only a stand-in): hipGetSymbolAddress gives
// each device's own address of it once, with that device current, and that address is used.
constexpr size_t kScratchBytes = 96ull << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void* g_scratchPtr[kMaxDevices] = {};
static std::mutex& g_scratchMutex = *new std::mutex();

// the current device's address of the workspace array (looked up on the first call per device)
static int scratch_ptr(void** out) {
  int dev = 0;
  DH_HIP(hipGetDevice(&dev));
  DH_CHECK(dev >= 0 && dev < kMaxDevices, "scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchPtr[dev]) {
    void* p = nullptr;
    DH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DH_CHECK(p != nullptr, "scratch: no device address");
    g_scratchPtr[dev] = p;
  }
  *out = g_scratchPtr[dev];
  return 0;
}

static int ensure_runtime() {
  if (g_rt.stream) return 0;
  int dev = 0;
  int n = 0;
  DH_HIP(hipGetDeviceCount(&n));
  if (n <= 0) {
    set_error("dftbhip: no HIP device available");
    return kErrDevice;
  }
  if (g_rt.device >= 0) dev = g_rt.device;
  DH_HIP(hipSetDevice(dev));
  g_rt.device = dev;
  DH_HIP(hipStreamCreate(&g_rt.stream));
  return 0;
}

// y = G^T x for a dense (n x n) column-major G: one wave per output element.
__global__ void gemv_t_kernel(int n, const double* G, const double* x, double* y) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (j >= n) return;
  const double* col = G + (long)j * n;
  double acc = 0.0;
  for (int i = lane; i < n; i += 64) acc += col[i] * x[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if (lane == 0) y[j] = acc;
}

__global__ void expand_kernel(int nat, const long* iAtomStart, const double* va, double* vo) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nat) return;
  for (long o = iAtomStart[a]; o < iAtomStart[a + 1]; ++o) vo[o] = va[a];
}

__global__ void scale_kernel(int n, const double* a, const double* b, double* c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = a[i] * b[i];
}

// ------------------------------------------------------------------------------------------
// Slater-Koster parameters on the device
// ------------------------------------------------------------------------------------------
struct SkTables {
  SkDev dev{};
  DevBuf<double> htab, stab, repdata;
  DevBuf<RepDev> dreps;
  int nsp = 0;
  std::vector<int> norb;
  std::vector<double> hubbu, q0;
  std::vector<double> repcut;
  double skcut = 0.0;
  std::vector<double> gcut;   // expGamma cutoffs (sp2 * nsp + sp1)
  DevBuf<double> dhubbu, dgcut;
};

static double expgamma_h(double r, double ua, double ub) {
  const double ta = 3.2 * ua, tb = 3.2 * ub;
  auto gsub = [](double rr, double t1, double t2) {
    const double d2 = t1 * t1 - t2 * t2;
    return std::exp(-t1 * rr) * ((0.5 * std::pow(t2, 4) * t1 / (d2 * d2)) -
                                 (std::pow(t2, 6) - 3.0 * std::pow(t2, 4) * t1 * t1) /
                                     (rr * d2 * d2 * d2));
  };
  if (r < 1.0e-5) {
    if (std::fabs(ua - ub) < 1.0e-10) return -0.5 * (ua + ub);
    return -0.5 * ((ta * tb) / (ta + tb) + (ta * tb) * (ta * tb) / std::pow(ta + tb, 3));
  }
  if (std::fabs(ua - ub) < 1.0e-10) {
    const double tm = 0.5 * (ta + tb);
    return std::exp(-tm * r) * (1.0 / r + 0.6875 * tm + 0.1875 * r * tm * tm +
                                0.02083333333333333333 * r * r * tm * tm * tm);
  }
  return gsub(r, ta, tb) + gsub(r, tb, ta);
}

// expGammaCutoff (src/dftbp/dftb/shortgammafuncs.F90:27-92)
static double expgamma_cutoff_h(double u2, double u1) {
  const double minv = 1.0e-10, tol = 1.0e-10;
  double rab = 1.0;
  while (expgamma_h(rab, u2, u1) > minv) rab *= 2.0;
  double mincut = rab + 0.1, maxcut = 0.5 * rab - 0.1;
  const double maxg = expgamma_h(maxcut, u2, u1), ming = expgamma_h(mincut, u2, u1);
  double lowerg = expgamma_h(mincut, u2, u1);
  double cut = maxcut + 0.1;
  double g = expgamma_h(cut, u2, u1);
  while ((g - lowerg) > tol) {
    maxcut = 0.5 * (cut + mincut);
    if ((maxg >= ming) == (minv >= expgamma_h(maxcut, u2, u1))) {
      mincut = maxcut;
      lowerg = expgamma_h(mincut, u2, u1);
    } else {
      cut = maxcut;
      g = expgamma_h(cut, u2, u1);
    }
  }
  return mincut;
}

static int sk_setup(SkTables& t, const dftbhip_params* p, hipStream_t s) {
  DH_CHECK(p->nspecies > 0 && p->nspecies <= kMaxSpecies, "params: bad number of species");
  DH_CHECK(p->nrows >= 9 && p->dist > 0.0, "params: SK grid needs at least 9 rows");
  const int ns = p->nspecies;
  t.nsp = ns;
  t.dev.nsp = ns;
  t.dev.nrows = p->nrows;
  t.dev.dist = p->dist;
  t.norb.assign(ns, 0);
  t.hubbu.assign(ns, 0.0);
  t.q0.assign(ns, 0.0);
  for (int a = 0; a < ns; ++a) {
    SpeciesDev& sp = t.dev.sp[a];
    DH_CHECK(p->nshell[a] >= 1 && p->nshell[a] <= 2, "params: 1 or 2 shells per species");
    sp.nshell = p->nshell[a];
    int no = 0;
    for (int sh = 0; sh < sp.nshell; ++sh) {
      const int l = p->angshell[a][sh];
      DH_CHECK(l == 0 || l == 1, "params: only s and p shells are supported");
      sp.ang[sh] = l;
      for (int m = 0; m < 2 * l + 1; ++m) {
        DH_CHECK(no < 4, "params: at most 4 orbitals per atom");
        sp.onsite[no++] = p->onsite[a][sh];
      }
      t.q0[a] += p->occ[a][sh];
    }
    sp.norb = no;
    t.norb[a] = no;
    t.hubbu[a] = p->hubbu[a];
  }
  std::vector<double> hall, sall;
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b < ns; ++b) {
      const int pr = a * ns + b;
      int nint = 0;
      for (int i = 0; i < t.dev.sp[a].nshell; ++i)
        for (int j = 0; j < t.dev.sp[b].nshell; ++j)
          nint += std::min(t.dev.sp[a].ang[i], t.dev.sp[b].ang[j]) + 1;
      DH_CHECK(p->ninteg[pr] == nint, "params: ninteg does not match the shells");
      DH_CHECK(p->htab[pr] && p->stab[pr], "params: missing SK table");
      t.dev.nint[pr] = nint;
      t.dev.tabOff[pr] = (long)hall.size();
      hall.insert(hall.end(), p->htab[pr], p->htab[pr] + (size_t)p->nrows * nint);
      sall.insert(sall.end(), p->stab[pr], p->stab[pr] + (size_t)p->nrows * nint);
    }
  DH_TRY(t.htab.upload(hall.data(), hall.size(), s));
  DH_TRY(t.stab.upload(sall.data(), sall.size(), s));
  t.dev.htab = t.htab.p;
  t.dev.stab = t.stab.p;
  t.skcut = p->nrows * p->dist + 1.0;
  // repulsives
  std::vector<double> rd;
  std::vector<RepDev> reps(ns * ns);
  std::vector<long> offs(ns * ns * 3);
  t.repcut.assign(ns * ns, 0.0);
  for (int pr = 0; pr < ns * ns; ++pr) {
    const int ni = p->rep_nint[pr];
    DH_CHECK(ni >= 1 && p->rep_xs[pr] && p->rep_last[pr], "params: missing repulsive");
    reps[pr].nint = ni;
    reps[pr].cutoff = p->rep_cutoff[pr];
    t.repcut[pr] = p->rep_cutoff[pr];
    for (int c = 0; c < 3; ++c) reps[pr].expc[c] = p->rep_expc[pr][c];
    offs[3 * pr] = (long)rd.size();
    rd.insert(rd.end(), p->rep_xs[pr], p->rep_xs[pr] + ni);
    offs[3 * pr + 1] = (long)rd.size();
    if (ni > 1) rd.insert(rd.end(), p->rep_coef[pr], p->rep_coef[pr] + 4 * (ni - 1));
    offs[3 * pr + 2] = (long)rd.size();
    rd.insert(rd.end(), p->rep_last[pr], p->rep_last[pr] + 6);
  }
  DH_TRY(t.repdata.upload(rd.data(), rd.size(), s));
  for (int pr = 0; pr < ns * ns; ++pr) {
    reps[pr].xs = t.repdata.p + offs[3 * pr];
    reps[pr].coef = t.repdata.p + offs[3 * pr + 1];
    reps[pr].last = t.repdata.p + offs[3 * pr + 2];
  }
  DH_TRY(t.dreps.upload(reps.data(), reps.size(), s));
  t.gcut.assign(ns * ns, 0.0);
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b < ns; ++b) t.gcut[b * ns + a] = expgamma_cutoff_h(t.hubbu[b], t.hubbu[a]);
  DH_TRY(t.dhubbu.upload(t.hubbu.data(), ns, s));
  DH_TRY(t.dgcut.upload(t.gcut.data(), ns * ns, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace dftbhip

struct dftbhip_sk {
  dftbhip::SkTables t;
};

using namespace dftbhip;

// ------------------------------------------------------------------------------------------
// device management
// ------------------------------------------------------------------------------------------
extern "C" int dftbhip_get_gpus_available(int* ngpus) {
  DH_CHECK(ngpus, "ngpus is null");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *ngpus = (e == hipSuccess) ? n : 0;
  return 0;
}

extern "C" int dftbhip_init(int device) {
  if (g_rt.stream) return 0;
  g_rt.device = device;
  DH_TRY(ensure_runtime());
  void* p = nullptr;
  return scratch_ptr(&p);   // the workspace address of this device, looked up at set-up
}

extern "C" int dftbhip_finalize(void) {
  release_sytrd_graphs();
  g_rt.release();
  if (g_rt.stream) (void)hipStreamDestroy(g_rt.stream);
  g_rt.stream = nullptr;
  return 0;
}

extern "C" int dftbhip_last_error(char* buf, int len) {
  if (!buf || len <= 0) return kErrArg;
  std::strncpy(buf, g_err.c_str(), len - 1);
  buf[len - 1] = 0;
  return 0;
}

// ------------------------------------------------------------------------------------------
// generalized eigensolver and density matrices (stateless drop-ins)
// ------------------------------------------------------------------------------------------
extern "C" int dftbhip_sygvd(int itype, char jobz, char uplo, int n, double* a, int lda,
                             double* b, int ldb, double* w) {
  DH_CHECK(itype == 1, "sygvd: only itype = 1");
  DH_CHECK(uplo == 'L' || uplo == 'l', "sygvd: only uplo = 'L'");
  DH_CHECK(jobz == 'V' || jobz == 'v' || jobz == 'N' || jobz == 'n', "sygvd: bad jobz");
  DH_CHECK(n >= 0 && lda >= std::max(1, n) && ldb >= std::max(1, n), "sygvd: bad dimensions");
  DH_CHECK(a && b && w, "sygvd: null pointer");
  if (n == 0) return 0;
  DH_TRY(ensure_runtime());
  hipStream_t s = g_rt.stream;
  EigenSolver& E = g_rt.eig;
  DH_TRY(E.ensure(n));
  DevBuf<double>& H = g_rt.stageA;
  DevBuf<double>& S = g_rt.stageB;
  DH_TRY(H.alloc((size_t)n * n));
  DH_TRY(S.alloc((size_t)n * n));
  DH_HIP(hipMemcpy2DAsync(H.p, n * sizeof(double), a, lda * sizeof(double), n * sizeof(double), n,
                          hipMemcpyHostToDevice, s));
  DH_HIP(hipMemcpy2DAsync(S.p, n * sizeof(double), b, ldb * sizeof(double), n * sizeof(double), n,
                          hipMemcpyHostToDevice, s));
  DH_TRY(symmetrize_lower(s, H.p, n, n));
  int info = 0;
  E.factored = false;
  DH_TRY(E.factor_overlap(s, S.p, n, &info));
  if (info != 0) {
    set_error("sygvd: overlap matrix not positive definite");
    return n + info;
  }
  std::vector<double> wh;
  DH_TRY(E.solve(s, H.p, n, wh, &info));
  if (info != 0) return info;
  std::copy(wh.begin(), wh.end(), w);
  if (jobz == 'V' || jobz == 'v') {
    DH_TRY(E.vectors(s, n, H.p, n));
    DH_HIP(hipMemcpy2DAsync(a, lda * sizeof(double), H.p, n * sizeof(double), n * sizeof(double),
                            n, hipMemcpyDeviceToHost, s));
  }
  DH_HIP(hipMemcpy2DAsync(b, ldb * sizeof(double), E.L.p, n * sizeof(double), n * sizeof(double),
                          n, hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

static int dm_common(int n, int nlev, const double* evec, int lde, const double* wts, double* dm,
                     int lddm) {
  DH_CHECK(n >= 0 && nlev >= 0 && nlev <= n && lde >= std::max(1, n) && lddm >= std::max(1, n),
           "density matrix: bad dimensions");
  if (n == 0) return 0;
  DH_TRY(ensure_runtime());
  hipStream_t s = g_rt.stream;
  DevBuf<double>& X = g_rt.stageA;
  DevBuf<double>& P = g_rt.stageB;
  DevBuf<double>& f = g_rt.stageW;
  DH_TRY(X.alloc((size_t)n * std::max(nlev, 1)));
  DH_TRY(P.alloc((size_t)n * n));
  DH_TRY(f.upload(wts, std::max(nlev, 1), s));
  if (nlev > 0)
    DH_HIP(hipMemcpy2DAsync(X.p, n * sizeof(double), evec, lde * sizeof(double),
                            n * sizeof(double), nlev, hipMemcpyHostToDevice, s));
  DH_TRY(gemm<double>(s, false, true, n, n, nlev, 1.0, X.p, n, X.p, n, 0.0, P.p, n, true, kKFull,
                      f.p));
  DH_HIP(hipMemcpy2DAsync(dm, lddm * sizeof(double), P.p, n * sizeof(double), n * sizeof(double),
                          n, hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

extern "C" int dftbhip_hegvd(int itype, char jobz, char uplo, int n, double* a, int lda,
                             double* b, int ldb, double* w) {
  DH_CHECK(itype == 1, "hegvd: only itype = 1");
  DH_CHECK(uplo == 'L' || uplo == 'l', "hegvd: only uplo = 'L'");
  DH_CHECK(jobz == 'V' || jobz == 'v' || jobz == 'N' || jobz == 'n', "hegvd: bad jobz");
  DH_CHECK(n >= 0 && lda >= std::max(1, n) && ldb >= std::max(1, n), "hegvd: bad dimensions");
  DH_CHECK(a && b && w, "hegvd: null pointer");
  if (n == 0) return 0;
  DH_TRY(ensure_runtime());
  hipStream_t s = g_rt.stream;
  ZhegvSolver& Z = g_rt.zeig;
  DH_TRY(Z.ensure(n));
  const size_t nn = (size_t)n * n;
  DevBuf<double>& A = g_rt.stageA;
  DevBuf<double>& B = g_rt.stageB;
  DH_TRY(A.alloc(2 * nn));
  DH_TRY(B.alloc(2 * nn));
  DH_HIP(hipMemcpy2DAsync(A.p, 2 * n * sizeof(double), a, 2 * (size_t)lda * sizeof(double),
                          2 * n * sizeof(double), n, hipMemcpyHostToDevice, s));
  DH_HIP(hipMemcpy2DAsync(B.p, 2 * n * sizeof(double), b, 2 * (size_t)ldb * sizeof(double),
                          2 * n * sizeof(double), n, hipMemcpyHostToDevice, s));
  const bool vec = jobz == 'V' || jobz == 'v';
  std::vector<double> wh;
  int info = 0;
  DH_TRY(Z.solve(s, A.p, n, B.p, n, vec, wh, &info));
  if (info > n) {
    set_error("hegvd: overlap matrix not positive definite");
    return info;
  }
  if (info != 0) {
    set_error("hegvd: divide and conquer failed");
    return info;
  }
  std::copy(wh.begin(), wh.end(), w);
  if (vec) {
    DH_TRY(merge_complex(s, Z.Hr.p, Z.Hi.p, n, A.p, n, n, n, 0));
    DH_HIP(hipMemcpy2DAsync(a, 2 * (size_t)lda * sizeof(double), A.p, 2 * n * sizeof(double),
                            2 * n * sizeof(double), n, hipMemcpyDeviceToHost, s));
  }
  // the Cholesky factor into the lower triangle of b, its strict upper part as given
  DH_TRY(merge_complex(s, Z.Lr.p, Z.Li.p, n, B.p, n, n, n, 2));
  DH_HIP(hipMemcpy2DAsync(b, 2 * (size_t)ldb * sizeof(double), B.p, 2 * n * sizeof(double),
                          2 * n * sizeof(double), n, hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

// P = C diag(w) C^H of complex C (interleaved, n x nlev, lde complex elements) through the
// planes: Re P = Cr W Cr^T + Ci W Ci^T, Im P = Ci W Cr^T - Cr W Ci^T (lower-output GEMMs with
// the weights applied in the loads); the lower triangle of P goes to dm (interleaved, lddm),
// zeros above it.
static int dm_cmplx_common(int n, int nlev, const double* evec, int lde, const double* wts,
                           double* dm, int lddm) {
  DH_CHECK(n >= 0 && nlev >= 0 && nlev <= n && lde >= std::max(1, n) && lddm >= std::max(1, n),
           "complex density matrix: bad dimensions");
  if (n == 0) return 0;
  DH_TRY(ensure_runtime());
  hipStream_t s = g_rt.stream;
  const size_t nn = (size_t)n * n, nl = (size_t)n * std::max(nlev, 1);
  DevBuf<double>& Z = g_rt.stageA;    // interleaved C in, interleaved P out
  DevBuf<double>& C = g_rt.stageB;    // Cr | Ci
  DevBuf<double>& P = g_rt.stageC;    // Pr | Pi
  DevBuf<double>& f = g_rt.stageW;
  DH_TRY(Z.alloc(2 * nn));
  DH_TRY(C.alloc(2 * nl));
  DH_TRY(P.alloc(2 * nn));
  DH_TRY(f.upload(wts, std::max(nlev, 1), s));
  double *Cr = C.p, *Ci = C.p + nl, *Pr = P.p, *Pi = P.p + nn;
  if (nlev > 0) {
    DH_HIP(hipMemcpy2DAsync(Z.p, 2 * n * sizeof(double), evec, 2 * (size_t)lde * sizeof(double),
                            2 * n * sizeof(double), nlev, hipMemcpyHostToDevice, s));
    DH_TRY(split_complex(s, Z.p, n, Cr, Ci, n, n, nlev));
  }
  DH_TRY(gemm<double>(s, false, true, n, n, nlev, 1.0, Cr, n, Cr, n, 0.0, Pr, n, true, kKFull,
                      f.p));
  DH_TRY(gemm<double>(s, false, true, n, n, nlev, 1.0, Ci, n, Ci, n, 1.0, Pr, n, true, kKFull,
                      f.p));
  DH_TRY(gemm<double>(s, false, true, n, n, nlev, 1.0, Ci, n, Cr, n, 0.0, Pi, n, true, kKFull,
                      f.p));
  DH_TRY(gemm<double>(s, false, true, n, n, nlev, -1.0, Cr, n, Ci, n, 1.0, Pi, n, true, kKFull,
                      f.p));
  DH_TRY(merge_complex(s, Pr, Pi, n, Z.p, n, n, n, 1));
  DH_HIP(hipMemcpy2DAsync(dm, 2 * (size_t)lddm * sizeof(double), Z.p, 2 * n * sizeof(double),
                          2 * n * sizeof(double), n, hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

extern "C" int dftbhip_dm_cmplx(int n, int nlevels, const double* evec, int lde,
                                const double* filling, double* dm, int lddm) {
  DH_CHECK(evec && filling && dm, "dm_cmplx: null pointer");
  return dm_cmplx_common(n, nlevels, evec, lde, filling, dm, lddm);
}

extern "C" int dftbhip_edm_cmplx(int n, int nlevels, const double* evec, int lde,
                                 const double* filling, const double* eigen, double* edm,
                                 int lddm) {
  DH_CHECK(evec && filling && eigen && edm, "edm_cmplx: null pointer");
  std::vector<double> w(std::max(nlevels, 1));
  for (int i = 0; i < nlevels; ++i) w[i] = filling[i] * eigen[i];
  return dm_cmplx_common(n, nlevels, evec, lde, w.data(), edm, lddm);
}

extern "C" int dftbhip_dm_real(int n, int nlevels, const double* evec, int lde,
                               const double* filling, double* dm, int lddm) {
  DH_CHECK(evec && filling && dm, "dm_real: null pointer");
  return dm_common(n, nlevels, evec, lde, filling, dm, lddm);
}

extern "C" int dftbhip_edm_real(int n, int nlevels, const double* evec, int lde,
                                const double* filling, const double* eigen, double* edm,
                                int lddm) {
  DH_CHECK(evec && filling && eigen && edm, "edm_real: null pointer");
  std::vector<double> w(std::max(nlevels, 1));
  for (int i = 0; i < nlevels; ++i) w[i] = filling[i] * eigen[i];
  return dm_common(n, nlevels, evec, lde, w.data(), edm, lddm);
}

// ------------------------------------------------------------------------------------------
// SK parameters and sparse drop-ins
// ------------------------------------------------------------------------------------------
extern "C" int dftbhip_sk_create(dftbhip_sk** sk, const dftbhip_params* p) {
  DH_CHECK(sk && p, "sk_create: null pointer");
  DH_TRY(ensure_runtime());
  auto* h = new dftbhip_sk();
  const int rc = sk_setup(h->t, p, g_rt.stream);
  if (rc != 0) {
    delete h;
    return rc;
  }
  *sk = h;
  return 0;
}

extern "C" int dftbhip_sk_destroy(dftbhip_sk* sk) {
  delete sk;
  return 0;
}

namespace {

// Flattened blocks from the reference's neighbour arrays (1-based atom/image indices).
struct HostBlocks {
  std::vector<int> at1, img2, at2;
  std::vector<long> start;
};

int blocks_from_reference(int nAtom, int nAll, const int* img2CentCell, int mNeighbour,
                          const int* iNeighbour, const int* nNeighbourSK, const int* iPair,
                          int sparseSize, const std::vector<int>& norbAtom, HostBlocks& hb) {
  const long ld = (long)mNeighbour + 1;
  for (int a = 0; a < nAtom; ++a) {
    DH_CHECK(nNeighbourSK[a] >= 0 && nNeighbourSK[a] <= mNeighbour, "nNeighbourSK out of range");
    for (int k = 0; k <= nNeighbourSK[a]; ++k) {
      const int img = iNeighbour[k + a * ld] - 1;
      DH_CHECK(img >= 0 && img < nAll, "iNeighbour out of range");
      const int jc = img2CentCell[img] - 1;
      DH_CHECK(jc >= 0 && jc < nAtom, "img2CentCell out of range");
      const long st = iPair[k + a * ld];
      DH_CHECK(st >= 0 && st + (long)norbAtom[a] * norbAtom[jc] <= sparseSize,
               "iPair out of range");
      hb.at1.push_back(a);
      hb.img2.push_back(img);
      hb.at2.push_back(jc);
      hb.start.push_back(st);
    }
  }
  return 0;
}

struct DevBlocks {
  DevBuf<int> at1, img2, at2, species, norb;
  DevBuf<long> start, atomStart;
  PairsDev pr{};
  // host copies of the orbital layout, for the tile mask of the packed density
  std::vector<int> hNorb;
  std::vector<long> hStart;
  // 128 x 128 tiles (row tile >= column tile) of the dense matrices that pack reads: the
  // density and energy-density products are formed on these tiles only
  DevBuf<int2> packTiles;
  int nPackTiles = 0;
  int upload(const HostBlocks& hb, const std::vector<int>& species, const std::vector<int>& norbAtom,
             const std::vector<long>& atomStart, hipStream_t s) {
    const size_t n = hb.at1.size();
    DH_TRY(at1.upload(hb.at1.data(), std::max<size_t>(n, 1), s));
    DH_TRY(img2.upload(hb.img2.data(), std::max<size_t>(n, 1), s));
    DH_TRY(at2.upload(hb.at2.data(), std::max<size_t>(n, 1), s));
    DH_TRY(start.upload(hb.start.data(), std::max<size_t>(n, 1), s));
    if (!species.empty()) DH_TRY(this->species.upload(species.data(), species.size(), s));
    DH_TRY(norb.upload(norbAtom.data(), norbAtom.size(), s));
    DH_TRY(this->atomStart.upload(atomStart.data(), atomStart.size(), s));
    hNorb = norbAtom;
    hStart = atomStart;
    pr.nblk = (int)n;
    pr.at1 = at1.p;
    pr.img2 = img2.p;
    pr.at2 = at2.p;
    pr.start = start.p;
    pr.species = this->species.p;
    rowsValid = false;
    return 0;
  }
  // the same from lists already on the device (start null: zeros)
  int from_device(int n, const int* dAt1, const int* dImg2, const int* dAt2, const long* dStart,
                  const std::vector<int>& species, const std::vector<int>& norbAtom,
                  const std::vector<long>& atomStart, hipStream_t s) {
    const size_t m = std::max(n, 1);
    DH_TRY(at1.alloc(m));
    DH_TRY(img2.alloc(m));
    DH_TRY(at2.alloc(m));
    DH_TRY(start.alloc(m));
    if (n > 0) {
      DH_HIP(hipMemcpyAsync(at1.p, dAt1, n * sizeof(int), hipMemcpyDeviceToDevice, s));
      DH_HIP(hipMemcpyAsync(img2.p, dImg2, n * sizeof(int), hipMemcpyDeviceToDevice, s));
      DH_HIP(hipMemcpyAsync(at2.p, dAt2, n * sizeof(int), hipMemcpyDeviceToDevice, s));
      if (dStart)
        DH_HIP(hipMemcpyAsync(start.p, dStart, n * sizeof(long), hipMemcpyDeviceToDevice, s));
      else
        DH_TRY(fill<long>(s, start.p, n, 0L));
    }
    if (!species.empty()) DH_TRY(this->species.upload(species.data(), species.size(), s));
    DH_TRY(norb.upload(norbAtom.data(), norbAtom.size(), s));
    DH_TRY(this->atomStart.upload(atomStart.data(), atomStart.size(), s));
    hNorb = norbAtom;
    hStart = atomStart;
    pr.nblk = n;
    pr.at1 = at1.p;
    pr.img2 = img2.p;
    pr.at2 = at2.p;
    pr.start = start.p;
    pr.species = this->species.p;
    rowsValid = false;
    return 0;
  }
  // row view of the blocks for the block-sparse products (built on first use per geometry):
  // a counting sort of the block list by row atom, on the host
  DevBuf<int> rowPtr, gPtr;
  DevBuf<int2> rowEnt, rowEntS, grp;
  RowBlocksDev rows{};
  bool rowsValid = false;
  int nGroups = 0;   // groups (row atom, column atom) of the row view
  int build_rows(int nat, hipStream_t s) {
    if (rowsValid) return 0;
    const int nb = pr.nblk;
    std::vector<int> h1(std::max(nb, 1)), h2(std::max(nb, 1));
    if (nb > 0) {
      DH_HIP(hipMemcpyAsync(h1.data(), pr.at1, nb * sizeof(int), hipMemcpyDeviceToHost, s));
      DH_HIP(hipMemcpyAsync(h2.data(), pr.at2, nb * sizeof(int), hipMemcpyDeviceToHost, s));
      DH_HIP(hipStreamSynchronize(s));
    }
    std::vector<int> ptr(nat + 1, 0);
    for (int k = 0; k < nb; ++k) {
      DH_CHECK(h1[k] >= 0 && h1[k] < nat && h2[k] >= 0 && h2[k] < nat,
               "row view: block atom out of range");
      ptr[h2[k] + 1] += 1;
      if (h1[k] != h2[k]) ptr[h1[k] + 1] += 1;
    }
    for (int a = 0; a < nat; ++a) ptr[a + 1] += ptr[a];
    // tiles of the dense matrices that pack reads: block k at rows of at2, columns of at1
    // (at2 >= at1 in the half list, so on or below the diagonal)
    if (hNorb.size() == (size_t)nat && hStart.size() >= (size_t)nat) {
      const long nOrb = hStart[nat - 1] + hNorb[nat - 1];
      const int nt = (int)cdiv(nOrb, kGemmTile);
      std::vector<char> mark((size_t)nt * nt, 0);
      for (int k = 0; k < nb; ++k) {
        const long r0 = hStart[h2[k]], r1 = r0 + hNorb[h2[k]] - 1;
        const long c0 = hStart[h1[k]], c1 = c0 + hNorb[h1[k]] - 1;
        for (long tr = r0 / kGemmTile; tr <= r1 / kGemmTile; ++tr)
          for (long tc = c0 / kGemmTile; tc <= c1 / kGemmTile; ++tc)
            mark[(size_t)std::max(tr, tc) * nt + std::min(tr, tc)] = 1;
      }
      std::vector<int2> tl;
      for (int tr = 0; tr < nt; ++tr)
        for (int tc = 0; tc <= tr; ++tc)
          if (mark[(size_t)tr * nt + tc]) tl.push_back(make_int2(tr, tc));
      nPackTiles = (int)tl.size();
      if (!tl.empty()) DH_TRY(packTiles.upload(tl.data(), tl.size(), s));
    } else {
      nPackTiles = 0;
    }
    std::vector<int2> ent(std::max(ptr[nat], 1));
    std::vector<int> fillPos(ptr.begin(), ptr.end() - 1);
    for (int k = 0; k < nb; ++k) {
      ent[fillPos[h2[k]]++] = make_int2(2 * k, h1[k]);
      if (h1[k] != h2[k]) ent[fillPos[h1[k]]++] = make_int2(2 * k + 1, h2[k]);
    }
    // a copy with each row's entries by column atom (block order kept within one), and the
    // groups of entries with one column atom in it; the products keep the block order, in which
    // the measured block-sparse product runs more than twice as fast
    std::vector<int2> ents(ent);
    std::vector<int> gp(nat + 1, 0);
    std::vector<int2> gr;
    gr.reserve(ptr[nat] + 1);
    for (int a = 0; a < nat; ++a) {
      std::stable_sort(ents.begin() + ptr[a], ents.begin() + ptr[a + 1],
                       [](const int2& x, const int2& y) { return x.y < y.y; });
      for (int p = ptr[a]; p < ptr[a + 1]; ++p)
        if (p == ptr[a] || ents[p].y != ents[p - 1].y) gr.push_back(make_int2(p, p + 1));
        else gr.back().y = p + 1;
      gp[a + 1] = (int)gr.size();
    }
    if (gr.empty()) gr.push_back(make_int2(0, 0));
    DH_TRY(rowPtr.upload(ptr.data(), ptr.size(), s));
    DH_TRY(rowEnt.upload(ent.data(), ent.size(), s));
    DH_TRY(rowEntS.upload(ents.data(), ents.size(), s));
    DH_TRY(gPtr.upload(gp.data(), gp.size(), s));
    DH_TRY(grp.upload(gr.data(), gr.size(), s));
    DH_HIP(hipStreamSynchronize(s));
    rows = RowBlocksDev{nat, rowPtr.p, rowEnt.p, rowEntS.p, gPtr.p, grp.p};
    nGroups = gp[nat];
    rowsValid = true;
    return 0;
  }
};

int atom_layout(int nAtom, const int* iAtomStart, std::vector<int>& norb, std::vector<long>& off) {
  norb.resize(nAtom);
  off.resize(nAtom + 1);
  for (int a = 0; a <= nAtom; ++a) off[a] = iAtomStart[a] - 1;
  for (int a = 0; a < nAtom; ++a) {
    norb[a] = (int)(off[a + 1] - off[a]);
    DH_CHECK(norb[a] >= 1 && norb[a] <= 4, "iAtomStart: 1..4 orbitals per atom");
  }
  return 0;
}

}  // namespace

extern "C" int dftbhip_build_h0_s(const dftbhip_sk* sk, int nAtom, int nAll, const double* coords,
                                  const int* species, const int* img2CentCell, int mNeighbour,
                                  const int* iNeighbour, const int* nNeighbourSK,
                                  const int* iPair, int sparseSize, double* h0, double* over) {
  DH_CHECK(sk && coords && species && img2CentCell && iNeighbour && nNeighbourSK && iPair && h0 &&
               over,
           "build_h0_s: null pointer");
  DH_CHECK(nAtom > 0 && nAll >= nAtom && sparseSize > 0, "build_h0_s: bad sizes");
  DH_TRY(ensure_runtime());
  hipStream_t s = g_rt.stream;
  std::vector<int> sp(nAtom), norb(nAtom);
  std::vector<long> off(nAtom + 1, 0);
  for (int a = 0; a < nAtom; ++a) {
    sp[a] = species[a] - 1;
    DH_CHECK(sp[a] >= 0 && sp[a] < sk->t.nsp, "species out of range");
    norb[a] = sk->t.norb[sp[a]];
    off[a + 1] = off[a] + norb[a];
  }
  HostBlocks hb;
  DH_TRY(blocks_from_reference(nAtom, nAll, img2CentCell, mNeighbour, iNeighbour, nNeighbourSK,
                               iPair, sparseSize, norb, hb));
  DevBlocks db;
  DH_TRY(db.upload(hb, sp, norb, off, s));
  DevBuf<double> c, dh0, ds;
  DH_TRY(c.upload(coords, 3 * (size_t)nAll, s));
  DH_TRY(dh0.alloc(sparseSize));
  DH_TRY(ds.alloc(sparseSize));
  DH_TRY(fill<double>(s, dh0.p, sparseSize, 0.0));
  DH_TRY(fill<double>(s, ds.p, sparseSize, 0.0));
  DH_TRY(launch_build_h0s(s, sk->t.dev, db.pr, c.p, dh0.p, ds.p));
  DH_HIP(hipMemcpyAsync(h0, dh0.p, sparseSize * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(over, ds.p, sparseSize * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

extern "C" int dftbhip_unpack_hs_real(int nAtom, const int* iAtomStart, const int* img2CentCell,
                                      int mNeighbour, const int* iNeighbour,
                                      const int* nNeighbourSK, const int* iPair, int sparseSize,
                                      const double* orig, double* square, int ldsq) {
  DH_CHECK(iAtomStart && img2CentCell && iNeighbour && nNeighbourSK && iPair && orig && square,
           "unpack: null pointer");
  DH_TRY(ensure_runtime());
  hipStream_t s = g_rt.stream;
  std::vector<int> norb;
  std::vector<long> off;
  DH_TRY(atom_layout(nAtom, iAtomStart, norb, off));
  const int n = (int)off[nAtom];
  DH_CHECK(ldsq >= n, "unpack: ldsq too small");
  int nAll = 0;
  for (int a = 0; a < nAtom; ++a)
    for (int k = 0; k <= nNeighbourSK[a]; ++k)
      nAll = std::max(nAll, iNeighbour[k + a * ((long)mNeighbour + 1)]);
  HostBlocks hb;
  DH_TRY(blocks_from_reference(nAtom, nAll, img2CentCell, mNeighbour, iNeighbour, nNeighbourSK,
                               iPair, sparseSize, norb, hb));
  DevBlocks db;
  DH_TRY(db.upload(hb, {}, norb, off, s));
  DevBuf<double> x, D;
  DH_TRY(x.upload(orig, sparseSize, s));
  DH_TRY(D.alloc((size_t)n * n));
  DH_TRY(db.build_rows(nAtom, s));
  DH_TRY(launch_unpack_rows<double>(s, db.rows, db.norb.p, db.atomStart.p, db.start.p, x.p, D.p,
                                    n, n));
  DH_HIP(hipMemcpy2DAsync(square, ldsq * sizeof(double), D.p, n * sizeof(double),
                          n * sizeof(double), n, hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

extern "C" int dftbhip_pack_hs_real(int nAtom, const int* iAtomStart, const int* img2CentCell,
                                    int mNeighbour, const int* iNeighbour,
                                    const int* nNeighbourSK, const int* iPair, int sparseSize,
                                    const double* square, int ldsq, double* primitive) {
  DH_CHECK(iAtomStart && img2CentCell && iNeighbour && nNeighbourSK && iPair && square &&
               primitive,
           "pack: null pointer");
  DH_TRY(ensure_runtime());
  hipStream_t s = g_rt.stream;
  std::vector<int> norb;
  std::vector<long> off;
  DH_TRY(atom_layout(nAtom, iAtomStart, norb, off));
  const int n = (int)off[nAtom];
  DH_CHECK(ldsq >= n, "pack: ldsq too small");
  int nAll = 0;
  for (int a = 0; a < nAtom; ++a)
    for (int k = 0; k <= nNeighbourSK[a]; ++k)
      nAll = std::max(nAll, iNeighbour[k + a * ((long)mNeighbour + 1)]);
  HostBlocks hb;
  DH_TRY(blocks_from_reference(nAtom, nAll, img2CentCell, mNeighbour, iNeighbour, nNeighbourSK,
                               iPair, sparseSize, norb, hb));
  DevBlocks db;
  DH_TRY(db.upload(hb, {}, norb, off, s));
  DevBuf<double> x, D;
  DH_TRY(D.alloc((size_t)n * n));
  DH_HIP(hipMemcpy2DAsync(D.p, n * sizeof(double), square, ldsq * sizeof(double),
                          n * sizeof(double), n, hipMemcpyHostToDevice, s));
  DH_TRY(x.alloc(sparseSize));
  DH_TRY(fill<double>(s, x.p, sparseSize, 0.0));
  DH_TRY(launch_pack<double>(s, db.pr, db.norb.p, db.atomStart.p, D.p, n, x.p));
  DH_HIP(hipMemcpyAsync(primitive, x.p, sparseSize * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

extern "C" int dftbhip_mulliken(int nAtom, const int* iAtomStart, const int* img2CentCell,
                                int mNeighbour, const int* iNeighbour, const int* nNeighbourSK,
                                const int* iPair, int sparseSize, const double* over,
                                const double* rho, double* qorb) {
  DH_CHECK(iAtomStart && img2CentCell && iNeighbour && nNeighbourSK && iPair && over && rho &&
               qorb,
           "mulliken: null pointer");
  DH_TRY(ensure_runtime());
  hipStream_t s = g_rt.stream;
  std::vector<int> norb;
  std::vector<long> off;
  DH_TRY(atom_layout(nAtom, iAtomStart, norb, off));
  const int n = (int)off[nAtom];
  int nAll = 0;
  for (int a = 0; a < nAtom; ++a)
    for (int k = 0; k <= nNeighbourSK[a]; ++k)
      nAll = std::max(nAll, iNeighbour[k + a * ((long)mNeighbour + 1)]);
  HostBlocks hb;
  DH_TRY(blocks_from_reference(nAtom, nAll, img2CentCell, mNeighbour, iNeighbour, nNeighbourSK,
                               iPair, sparseSize, norb, hb));
  DevBlocks db;
  DH_TRY(db.upload(hb, {}, norb, off, s));
  DevBuf<double> ds, dr, q;
  DH_TRY(ds.upload(over, sparseSize, s));
  DH_TRY(dr.upload(rho, sparseSize, s));
  DH_TRY(q.alloc(n));
  DH_TRY(db.build_rows(nAtom, s));
  DH_TRY(launch_mulliken_rows(s, db.rows, db.norb.p, db.atomStart.p, db.start.p, ds.p, nullptr,
                              dr.p, q.p, nullptr));
  DH_HIP(hipMemcpyAsync(qorb, q.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

// ------------------------------------------------------------------------------------------
// Resident SCC context
// ------------------------------------------------------------------------------------------
namespace dftbhip {

// Half neighbour list (src/dftbp/dftb/periodic.F90:446-696): atom a lists the images of atoms
// b >= a within the cutoff, sorted by distance; images are numbered after the central atoms in
// (translation, atom) order. Candidate images are binned on a grid of cutoff-sized cells.
struct NeighbourList {
  std::vector<double> coordsAll;               // 3 * nAll
  std::vector<int> img2cent;                   // nAll
  std::vector<std::vector<int>> nb;            // per atom: images (self first)
  std::vector<std::vector<double>> d2;
};


static int build_neighbours(int nat, const double* coords, const double* lat, double cutoff,
                            NeighbourList& nl) {
  std::vector<double> tr;
  cell_translations(lat, cutoff, tr);
  const int ntr = (int)tr.size() / 3;
  double lo[3], hi[3];
  for (int c = 0; c < 3; ++c) {
    lo[c] = 1e300;
    hi[c] = -1e300;
  }
  for (int a = 0; a < nat; ++a)
    for (int c = 0; c < 3; ++c) {
      lo[c] = std::min(lo[c], coords[3 * a + c]);
      hi[c] = std::max(hi[c], coords[3 * a + c]);
    }
  // candidate points (translation, atom) inside the box grown by the cutoff
  struct Pt {
    double x[3];
    int t, a;
  };
  std::vector<Pt> pts;
  for (int t = 0; t < ntr; ++t)
    for (int a = 0; a < nat; ++a) {
      Pt p;
      bool in = true;
      for (int c = 0; c < 3; ++c) {
        p.x[c] = coords[3 * a + c] + tr[3 * t + c];
        if (p.x[c] < lo[c] - cutoff || p.x[c] > hi[c] + cutoff) in = false;
      }
      if (!in) continue;
      p.t = t;
      p.a = a;
      pts.push_back(p);
    }
  int nbin[3];
  for (int c = 0; c < 3; ++c)
    nbin[c] = std::max(1, (int)std::floor((hi[c] - lo[c] + 2 * cutoff) / cutoff));
  auto binOf = [&](const double* x, int* b) {
    for (int c = 0; c < 3; ++c) {
      int v = (int)std::floor((x[c] - (lo[c] - cutoff)) / cutoff);
      b[c] = std::min(std::max(v, 0), nbin[c] - 1);
    }
  };
  std::vector<std::vector<int>> bins((size_t)nbin[0] * nbin[1] * nbin[2]);
  for (size_t i = 0; i < pts.size(); ++i) {
    int b[3];
    binOf(pts[i].x, b);
    bins[((size_t)b[0] * nbin[1] + b[1]) * nbin[2] + b[2]].push_back((int)i);
  }
  const double cut2 = cutoff * cutoff;
  std::vector<std::vector<std::pair<double, int>>> found(nat);   // (d2, point index)
  for (int a2 = 0; a2 < nat; ++a2) {
    int b[3];
    binOf(coords + 3 * a2, b);
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          const int bx = b[0] + dx, by = b[1] + dy, bz = b[2] + dz;
          if (bx < 0 || by < 0 || bz < 0 || bx >= nbin[0] || by >= nbin[1] || bz >= nbin[2])
            continue;
          for (int pi : bins[((size_t)bx * nbin[1] + by) * nbin[2] + bz]) {
            const Pt& p = pts[pi];
            if (p.a < a2) continue;
            double d2 = 0.0;
            for (int c = 0; c < 3; ++c) d2 += (p.x[c] - coords[3 * a2 + c]) * (p.x[c] - coords[3 * a2 + c]);
            if (d2 > cut2) continue;
            if (d2 < 1.0e-4) {
              if (p.t == 0 && p.a == a2) continue;
              set_error("Atoms " + std::to_string(p.a + 1) + " and " + std::to_string(a2 + 1) +
                        " too close together");
              return kErrArg;
            }
            found[a2].push_back({d2, pi});
          }
        }
  }
  // image numbering: central atoms first, then used images in (translation, atom) order
  std::vector<int> used(pts.size(), -1);
  std::vector<int> order;
  for (int a2 = 0; a2 < nat; ++a2)
    for (auto& f : found[a2]) {
      const Pt& p = pts[f.second];
      if (p.t != 0 && used[f.second] < 0) {
        used[f.second] = 0;
        order.push_back(f.second);
      }
    }
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    return pts[x].t != pts[y].t ? pts[x].t < pts[y].t : pts[x].a < pts[y].a;
  });
  nl.coordsAll.assign(coords, coords + 3 * nat);
  nl.img2cent.resize(nat);
  std::iota(nl.img2cent.begin(), nl.img2cent.end(), 0);
  for (int pi : order) {
    used[pi] = (int)nl.img2cent.size();
    nl.img2cent.push_back(pts[pi].a);
    for (int c = 0; c < 3; ++c) nl.coordsAll.push_back(pts[pi].x[c]);
  }
  nl.nb.assign(nat, {});
  nl.d2.assign(nat, {});
  for (int a2 = 0; a2 < nat; ++a2) {
    auto& f = found[a2];
    std::sort(f.begin(), f.end(), [&](const std::pair<double, int>& x, const std::pair<double, int>& y) {
      if (x.first != y.first) return x.first < y.first;
      const Pt &px = pts[x.second], &py = pts[y.second];
      return px.t != py.t ? px.t < py.t : px.a < py.a;
    });
    nl.nb[a2].push_back(a2);
    nl.d2[a2].push_back(0.0);
    for (auto& x : f) {
      const Pt& p = pts[x.second];
      nl.nb[a2].push_back(p.t == 0 ? p.a : used[x.second]);
      nl.d2[a2].push_back(x.first);
    }
  }
  return 0;
}

// Ewald parameters (src/dftbp/dftb/coulomb.F90:1824-2037) and reciprocal vectors.
struct EwaldSetup {
  double alpha = 0, maxr = 0, maxg = 0, vol = 0;
  std::vector<double> G, w, tr;
};

static double gterm(double g, double al, double vol) {
  return 4.0 * M_PI * (std::exp(-0.25 * g * g / (al * al)) / (vol * g * g));
}
static double rterm(double r, double al) { return std::erfc(al * r) / r; }

template <typename F>
static double bisect_max(F fn, double minv) {
  double xx = 1.0e-8, yy = fn(xx);
  while (yy > minv) {
    xx *= 2.0;
    yy = fn(xx);
  }
  double xl = 0.5 * xx, yl = fn(xl), xr = xx, yr = yy;
  int it = 1;
  while (yl - yr > minv && it <= 30) {
    xx = 0.5 * (xl + xr);
    yy = fn(xx);
    if (yy >= minv) {
      xl = xx;
      yl = yy;
    } else {
      xr = xx;
      yr = yy;
    }
    ++it;
  }
  return xx;
}

static void ewald_setup(const double* lat, double tol, EwaldSetup& ew) {
  const double* a = lat;
  const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
  ew.vol = std::fabs(det);
  // reciprocal vectors rec_i (rows) = 2 pi * (inv(lat)^T)_i
  double inv[9];
  inv[0] = (a[4] * a[8] - a[5] * a[7]) / det;
  inv[1] = -(a[1] * a[8] - a[2] * a[7]) / det;
  inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
  inv[3] = -(a[3] * a[8] - a[5] * a[6]) / det;
  inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
  inv[5] = -(a[0] * a[5] - a[2] * a[3]) / det;
  inv[6] = (a[3] * a[7] - a[4] * a[6]) / det;
  inv[7] = -(a[0] * a[7] - a[1] * a[6]) / det;
  inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
  double rec[9];
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) rec[3 * i + c] = 2.0 * M_PI * inv[3 * c + i];
  double ming = 1e300, minr = 1e300;
  for (int i = 0; i < 3; ++i) {
    ming = std::min(ming, std::sqrt(rec[3 * i] * rec[3 * i] + rec[3 * i + 1] * rec[3 * i + 1] +
                                    rec[3 * i + 2] * rec[3 * i + 2]));
    minr = std::min(minr, std::sqrt(a[3 * i] * a[3 * i] + a[3 * i + 1] * a[3 * i + 1] +
                                    a[3 * i + 2] * a[3 * i + 2]));
  }
  const double vol = ew.vol;
  auto diff = [&](double al) {
    return (gterm(4.0 * ming, al, vol) - gterm(5.0 * ming, al, vol)) -
           (rterm(2.0 * minr, al) - rterm(3.0 * minr, al));
  };
  double alpha = 1.0e-8, d = diff(alpha);
  while (d < -tol) {
    alpha *= 2.0;
    d = diff(alpha);
  }
  double aleft = 0.5 * alpha;
  while (d < tol) {
    alpha *= 2.0;
    d = diff(alpha);
  }
  double aright = alpha;
  alpha = 0.5 * (aleft + aright);
  d = diff(alpha);
  int it = 0;
  while (std::fabs(d) > tol && it <= 30) {
    if (d < 0) aleft = alpha;
    else aright = alpha;
    alpha = 0.5 * (aleft + aright);
    d = diff(alpha);
    ++it;
  }
  ew.alpha = alpha;
  ew.maxr = bisect_max([&](double x) { return rterm(x, alpha); }, tol);
  ew.maxg = bisect_max([&](double x) { return gterm(x, alpha, vol); }, tol);
  int nmax[3];
  for (int i = 0; i < 3; ++i) {
    const double nr = std::sqrt(rec[3 * i] * rec[3 * i] + rec[3 * i + 1] * rec[3 * i + 1] +
                                rec[3 * i + 2] * rec[3 * i + 2]);
    nmax[i] = (int)std::ceil(ew.maxg / nr) + 1;
  }
  ew.G.clear();
  ew.w.clear();
  for (int x = -nmax[0]; x <= nmax[0]; ++x)
    for (int y = -nmax[1]; y <= nmax[1]; ++y)
      for (int z = -nmax[2]; z <= nmax[2]; ++z) {
        // one of each +-G pair: lexicographically positive (x, y, z)
        if (x < 0 || (x == 0 && (y < 0 || (y == 0 && z <= 0)))) continue;
        double g[3];
        for (int c = 0; c < 3; ++c) g[c] = x * rec[c] + y * rec[3 + c] + z * rec[6 + c];
        const double g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
        if (g2 > ew.maxg * ew.maxg) continue;
        ew.G.insert(ew.G.end(), g, g + 3);
        ew.w.push_back(2.0 * 4.0 * M_PI / vol * std::exp(-g2 / (4.0 * alpha * alpha)) / g2);
      }
  cell_translations(lat, ew.maxr, ew.tr);
}

// Fermi filling of one spin-unpolarised k-point (etemp.F90 Efilling/electronFill/middleGap,
// main.F90:4544-4651): fillings include the factor 2.
struct FillResult {
  double ef, eband, ts, e0;
};

static double electron_count(double ef, const std::vector<double>& e, double kT) {
  double s = 0.0, comp = 0.0;
  for (double x : e) {
    const double arg = (x - ef) / kT;
    const double v = 1.0 / (1.0 + std::exp(arg));
    const double y = v - comp;
    const double t = s + y;
    comp = (t - s) - y;
    s = t;
  }
  return s;
}

static void electron_fill(double ef, const std::vector<double>& e, double kT,
                          std::vector<double>& f, double& eband, double& ts) {
  const double eps = 2.220446049250313e-16;
  const int n = (int)e.size();
  f.assign(n, 0.0);
  double eb = 0.0, cb = 0.0, tsum = 0.0, ct = 0.0;
  auto kahan = [](double& s, double v, double& c) {
    const double y = v - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  };
  for (int j = 0; j < n; ++j) {
    const double x = (e[j] - ef) / kT;
    f[j] = x > 700.0 ? 0.0 : 1.0 / (1.0 + std::exp(x));
    if (f[j] <= 1.0e-15) break;
    if (f[j] > eps && f[j] < 1.0 - eps)
      kahan(tsum, -(f[j] * std::log(f[j]) + (1.0 - f[j]) * std::log(1.0 - f[j])), ct);
    kahan(eb, f[j] * e[j], cb);
  }
  eband = eb;
  ts = tsum * kT;
}

static int fermi_filling(const std::vector<double>& e, double nel, double kT,
                         std::vector<double>& f, FillResult& r) {
  kT = std::max(kT, 1.0e-8);
  const double ne = 0.5 * nel;
  const int n = (int)e.size();
  const double eps2 = 2.0 * 2.220446049250313e-16;
  double ef, eb, ts;
  auto finish = [&](double scale) {
    electron_fill(ef, e, kT, f, eb, ts);
    if (scale != 1.0)
      for (double& x : f) x *= scale;
    for (double& x : f) x *= 2.0;
    r.ef = ef;
    r.eband = 2.0 * eb;
    r.ts = 2.0 * ts;
    r.e0 = 2.0 * (eb - 0.5 * ts);
  };
  if (ne < 2.220446049250313e-16 || n == 0) {
    // no electrons: empty levels, the Fermi level well below the lowest one
    f.assign(n, 0.0);
    r.ef = (n ? e.front() : 0.0) - 1000.0 * (kT + 1.1920928955078125e-07);
    r.eband = r.ts = r.e0 = 0.0;
    return 0;
  }
  if (n <= ne) {
    ef = e.back() + 1000.0 * (kT + 1.1920928955078125e-07);
    finish(1.0);
    return 0;
  }
  if (std::fabs(ne - std::round(ne)) <= 1.0e-15) {
    int k = (int)std::ceil(ne - 1e-12);
    k = std::min(k, n - 1);
    ef = 0.5 * (e[k] + e[k - 1]);
    if (std::fabs(ne - electron_count(ef, e, kT)) <= 1.0e-7) {
      finish(1.0);
      return 0;
    }
  }
  double upper = e.back() + 0.01, lower = e.front() - 0.01;
  double nmax = electron_count(upper, e, kT), nmin = electron_count(lower, e, kT);
  while (nmin > ne) {
    lower = 2.0 * (lower - upper) + lower;
    nmin = electron_count(lower, e, kT);
  }
  while (nmax < ne) {
    upper = 2.0 * (upper - lower) + lower;
    nmax = electron_count(upper, e, kT);
  }
  ef = 0.5 * (upper + lower);
  double nc = electron_count(ef, e, kT);
  while (std::fabs(ne - nc) > 1.0e-15 &&
         std::fabs(upper - lower) >= std::max(std::fabs(ef) * eps2, eps2)) {
    if ((nmax >= nmin) == (ne >= nc)) {
      lower = ef;
      nmin = nc;
    } else {
      upper = ef;
      nmax = nc;
    }
    ef = 0.5 * (upper + lower);
    nc = electron_count(ef, e, kT);
  }
  if (std::fabs(ne - nc) > 1.0e-7) {
    set_error("Fermi level search did not converge.");
    return kErrState;
  }
  finish((ne > eps2 && std::fabs(ne - nc) > eps2) ? ne / nc : 1.0);
  return 0;
}

// Fermi filling over k-points with normalised weights kw (Efilling with kWeight,
// etemp.F90:59-194; middleGap :481-520): e[k] ascending levels of k-point k. Fillings carry the
// factor 2 of the spin-unpolarised case; Eband and TS are k-weighted.
static int fermi_filling_k(const std::vector<std::vector<double>>& e,
                           const std::vector<double>& kw, double nel, double kT,
                           std::vector<std::vector<double>>& f, FillResult& r) {
  kT = std::max(kT, 1.0e-8);
  const double ne = 0.5 * nel;
  const int nk = (int)e.size();
  const int nlev = nk ? (int)e[0].size() : 0;
  const double eps2 = 2.0 * 2.220446049250313e-16;
  auto count = [&](double ef) {
    double s = 0.0, comp = 0.0;
    for (int k = 0; k < nk; ++k)
      for (double x : e[k]) {
        const double v = kw[k] / (1.0 + std::exp((x - ef) / kT));
        const double y = v - comp;
        const double t = s + y;
        comp = (t - s) - y;
        s = t;
      }
    return s;
  };
  double ef = 0.0;
  auto finish = [&](double scale) {
    f.assign(nk, std::vector<double>());
    double eb = 0.0, ts = 0.0, cb = 0.0, ct = 0.0;
    auto kahan = [](double& sum, double v, double& c) {
      const double y = v - c;
      const double t = sum + y;
      c = (t - sum) - y;
      sum = t;
    };
    for (int k = 0; k < nk; ++k) {
      double ebk, tsk;
      electron_fill(ef, e[k], kT, f[k], ebk, tsk);
      kahan(eb, kw[k] * ebk, cb);
      kahan(ts, kw[k] * tsk, ct);
      for (double& x : f[k]) x *= 2.0 * scale;
    }
    r.ef = ef;
    r.eband = 2.0 * eb;
    r.ts = 2.0 * ts;
    r.e0 = 2.0 * (eb - 0.5 * ts);
  };
  double emin = HUGE_VAL, emax = -HUGE_VAL;
  for (int k = 0; k < nk; ++k)
    for (double x : e[k]) {
      emin = std::min(emin, x);
      emax = std::max(emax, x);
    }
  if (ne < 2.220446049250313e-16 || nlev == 0) {
    // no electrons: empty levels, the Fermi level well below the lowest one
    f.assign(nk, std::vector<double>(nlev, 0.0));
    r.ef = (nlev ? emin : 0.0) - 1000.0 * (kT + 1.1920928955078125e-07);
    r.eband = r.ts = r.e0 = 0.0;
    return 0;
  }
  if (nlev <= ne) {
    ef = emax + 1000.0 * (kT + 1.1920928955078125e-07);
    finish(1.0);
    return 0;
  }
  if (std::fabs(ne - std::round(ne)) <= 1.0e-15) {
    // all levels sorted, weights accumulated until the electron count is reached
    std::vector<std::pair<double, int>> all;
    all.reserve((size_t)nk * nlev);
    for (int k = 0; k < nk; ++k)
      for (int j = 0; j < nlev; ++j) all.push_back({e[k][j], k});
    std::stable_sort(all.begin(), all.end(),
                     [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                       return a.first < b.first;
                     });
    double acc = 0.0;
    size_t ind = 0;
    while (acc < ne && ind < all.size()) acc += kw[all[ind++].second];
    ind = std::min(all.size() - 1, ind);
    ef = 0.5 * (all[ind].first + all[ind - 1].first);
    if (std::fabs(ne - count(ef)) <= 1.0e-7) {
      finish(1.0);
      return 0;
    }
  }
  double upper = emax + 0.01, lower = emin - 0.01;
  double nmax = count(upper), nmin = count(lower);
  while (nmin > ne) {
    lower = 2.0 * (lower - upper) + lower;
    nmin = count(lower);
  }
  while (nmax < ne) {
    upper = 2.0 * (upper - lower) + lower;
    nmax = count(upper);
  }
  ef = 0.5 * (upper + lower);
  double nc = count(ef);
  while (std::fabs(ne - nc) > 1.0e-15 &&
         std::fabs(upper - lower) >= std::max(std::fabs(ef) * eps2, eps2)) {
    if ((nmax >= nmin) == (ne >= nc)) {
      lower = ef;
      nmin = nc;
    } else {
      upper = ef;
      nmax = nc;
    }
    ef = 0.5 * (upper + lower);
    nc = count(ef);
  }
  if (std::fabs(ne - nc) > 1.0e-7) {
    set_error("Fermi level search did not converge.");
    return kErrState;
  }
  finish((ne > eps2 && std::fabs(ne - nc) > eps2) ? ne / nc : 1.0);
  return 0;
}

// Modified Broyden mixing (src/dftbp/mixer/broydenmixer.F90:204-342).
struct Broyden {
  int n = 0, it = 0;
  double alpha, omega0, minw, maxw, wfac;
  std::vector<double> ww, qlast, dlast;
  std::vector<std::vector<double>> aa, dF, uu;
  void reset(int n_, const dftbhip_options& o, int maxIter) {
    n = n_;
    it = 0;
    alpha = o.mix_param;
    omega0 = o.omega0;
    minw = o.min_weight;
    maxw = o.max_weight;
    wfac = o.weight_fac;
    ww.assign(maxIter + 1, 0.0);
    aa.assign(maxIter + 1, std::vector<double>(maxIter + 1, 0.0));
    dF.clear();
    uu.clear();
    qlast.assign(n, 0.0);
    dlast.assign(n, 0.0);
  }
  static double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
  }
  void mix(std::vector<double>& q, const std::vector<double>& d) {
    ++it;
    const int nn = it;
    if (nn == 1) {
      qlast = q;
      dlast = d;
      for (int i = 0; i < n; ++i) q[i] += alpha * d[i];
      return;
    }
    const int n1 = nn - 1;
    double w = std::sqrt(dot(d, d));
    w = (w > wfac / maxw) ? wfac / w : maxw;
    w = std::max(w, minw);
    ww[n1 - 1] = w;
    std::vector<double> dfu(n);
    for (int i = 0; i < n; ++i) dfu[i] = d[i] - dlast[i];
    double inv = std::max(std::sqrt(dot(dfu, dfu)), 2.220446049250313e-16);
    inv = 1.0 / inv;
    for (double& x : dfu) x *= inv;
    std::vector<double> cc(n1);
    for (int ii = 0; ii < nn - 2; ++ii) {
      aa[ii][n1 - 1] = dot(dF[ii], dfu);
      aa[n1 - 1][ii] = aa[ii][n1 - 1];
      cc[ii] = ww[ii] * dot(dF[ii], d);
    }
    aa[n1 - 1][n1 - 1] = 1.0;
    cc[n1 - 1] = ww[n1 - 1] * dot(dfu, d);
    // beta(i,j) = w_i w_j a(i,j) + omega0^2 delta; solve beta^T x = cc (LU with pivoting)
    std::vector<std::vector<double>> B(n1, std::vector<double>(n1));
    for (int i = 0; i < n1; ++i)
      for (int j = 0; j < n1; ++j)
        B[i][j] = ww[i] * ww[j] * aa[i][j] + (i == j ? omega0 * omega0 : 0.0);
    // B is symmetric: beta^T = beta
    std::vector<int> piv(n1);
    for (int k = 0; k < n1; ++k) {
      int p = k;
      for (int i = k + 1; i < n1; ++i)
        if (std::fabs(B[i][k]) > std::fabs(B[p][k])) p = i;
      std::swap(B[k], B[p]);
      std::swap(cc[k], cc[p]);
      for (int i = k + 1; i < n1; ++i) {
        const double f = B[i][k] / B[k][k];
        for (int j = k; j < n1; ++j) B[i][j] -= f * B[k][j];
        cc[i] -= f * cc[k];
      }
    }
    for (int k = n1 - 1; k >= 0; --k) {
      double s = cc[k];
      for (int j = k + 1; j < n1; ++j) s -= B[k][j] * cc[j];
      cc[k] = s / B[k][k];
    }
    dF.push_back(dfu);
    std::vector<double> u(n);
    for (int i = 0; i < n; ++i) u[i] = alpha * dfu[i] + inv * (q[i] - qlast[i]);
    qlast = q;
    dlast = d;
    for (int i = 0; i < n; ++i) q[i] += alpha * d[i];
    for (int ii = 0; ii < nn - 2; ++ii)
      for (int i = 0; i < n; ++i) q[i] -= ww[ii] * cc[ii] * uu[ii][i];
    for (int i = 0; i < n; ++i) q[i] -= ww[n1 - 1] * cc[n1 - 1] * u[i];
    uu.push_back(u);
  }
};

enum Timer {
  kTSetup = 0, kTHam, kTReduce, kTTridiag, kTDC, kTBack, kTDM, kTPackMull, kTHost, kTForces,
  kTIter, kTRefine, kTCount
};

struct SccState {
  std::vector<double> qin;
  Broyden mixer;
  int it = 0;
  bool conv = false, active = false;
  double err = 0.0, nel = 0.0, eNonScc = 0.0, eScc = 0.0;
  double prevErr = HUGE_VAL;   // charge error of the previous iteration (mixed-path switch)
  // electronic energy of the last and of the previous iteration, input electrons of the last
  double eElec = 0.0, eElecPrev = std::numeric_limits<double>::quiet_NaN(), qinSum = 0.0;
  FillResult fr{};
};

// k-points of one real size tridiagonalised together (1, the default) or one by one (0); set
// through dftbhip_set_tuning parameter 2
static int g_kBatch = 1;
// neighbour list, pairs and SK blocks of a geometry built on the device (1, the default) or on
// the host (0); set through dftbhip_set_tuning parameter 3
static int g_devNeighbours = 1;
// the Broyden history restarts at the iteration that seeds the mixed-precision tail (1) or is
// kept (0, the default); set through dftbhip_set_tuning parameter 7
static int g_tailMixReset = 0;
// arithmetic of a step of the sparse refinement (dftbhip_set_tuning parameter 8):
//  0: M = X^T H X and G = X^T S X in FP64, the correction product X E in FP64;
//  1: the same with X E in FP32;
//  2: residual form (the default): G in FP64, the Rayleigh quotients and the residuals
//     R = H X - S X diag(lam) in FP64, X^T R and X E as FP16 products (FP32 accumulation) of
//     row-scaled operands.
static int g_tailUpdate32 = 2;
// residual form: tile rows of X^T S X below its diagonal tiles that each step forms (128 x 128
// tiles; -1: the whole lower triangle); dftbhip_set_tuning parameter 10
static int g_tailBand = 1;
// residual form: iterations whose previous charge error is at most this use arithmetic 1
// (0: never, the default); dftbhip_set_tuning parameter 11 (value e: 10^-e, 0: never)
static double g_tailFineBelow = 0.0;

struct Context {
  dftbhip_options opt{};
  SccState st;
  SkTables sk;
  hipStream_t s = nullptr;
  int nat = 0, norbTot = 0;
  bool periodic = false;
  double lat[9]{};
  std::vector<double> coords;
  std::vector<int> species;
  NeighbourList nl;
  std::vector<int> norbAtom;
  std::vector<long> atomStart;
  HostBlocks blk, nbp;
  DevBlocks dblk, dnb;
  NeighDev nbdev{};
  long nnz = 0;
  EwaldSetup ew;
  DevBuf<double> coordsAll, coordsC, h0, ov, hs, rho, erho, gamma, vorb, vatom, dq, qorb, eorb;
  DevBuf<double> Hd, Sd, X, P, f, feps, grad, eatom, ewG, ewW, ewTr, ewCS, ewCSq, ww2;
  DevBuf<float> f32;   // FP32 filling weights of the mixed path
  // k-points: fractional k (3 nk), normalised weights; per block fractional image translation;
  // per k-point the inverse Cholesky factor of the embedded S(k) and all eigenvectors
  int nk = 0;
  std::vector<double> kpts, kw;
  DevBuf<double> blkT;
  std::vector<DevBuf<double>> LinvK, XK;
  DevNeighbourList dnl;      // neighbour list built on the device
  bool nlOnDevice = false;
  // k-points tridiagonalised together (all real, one size): reduced matrices and reflectors,
  // d, e, tau and the workspace blocks of every point
  DevBuf<double> kA, kD, kE, kTau;
  DevBuf<char> kWs;
  // their divide and conquer side by side: per point a stream (default priority), the D&C
  // buffers and the eigenvalues; the eigenvectors go straight to XK
  struct KPointDc {
    hipStream_t st = nullptr;
    std::shared_ptr<StedcWs> ws;
    DevBuf<double> work, w;
    KPointDc() = default;
    KPointDc(const KPointDc&) = delete;
    KPointDc& operator=(const KPointDc&) = delete;
    ~KPointDc() {
      if (st) (void)hipStreamDestroy(st);
    }
  };
  std::vector<std::unique_ptr<KPointDc>> kdc;
  std::vector<std::vector<double>> eigK, fillK;
  std::vector<int> noccK;
  EigenSolver eig;
  std::vector<double> q0atom, eigv, fill, qout, dqout;
  std::vector<double> qorbLast;   // output orbital populations of the last SCC iteration
  double energies[8]{};
  int nocc = 0;
  bool xf64 = true;     // the eigenvectors in X are FP64 (else FP32, mixed path)
  bool xFull = false;   // X holds all n refined FP64 eigenvectors of the last iteration
  // Ogita-Aishima refinement on the sparse H and S: the n eigenvectors transposed (two buffers,
  // tCur the current one), the products (H X)^T, (S X)^T, M = X^T H X, G = X^T S X, the
  // Rayleigh quotients
  DevBuf<double> tXt[2], tYH, tYS, tM, tG, tLam;
  DevBuf<int> tClip;
  // residual form: FP16 operands (padded, f16_ld x f16_cols), row reductions and scales
  DevBuf<_Float16> hX, hXt, hR, hE;
  DevBuf<double> tPart, tRsc;
  DevBuf<float> tRinv;
  DevBuf<int> tFar;
  int lastFar = 0;   // pairs beyond the band of G in the last refinement (whole G formed)
  std::vector<double> qinGiven;   // input charges for the next scc_begin (set_input_charges)
  bool tailOff = false;   // the tail failed once in this cycle: no more tail iterations
  DevBuf<double> tMb, tCt;   // band tiles of X^T H X; in-band couplings (transposed)
  // descriptors of the in-band deflation YH <- YH - C^T YS, one batch per tile offset
  struct DeflBatch {
    DevBuf<GemmDesc<double>> desc;
    int count = 0;
  };
  std::vector<DeflBatch> tDefl;
  // folded 4 x 4 blocks of H and S per row-view group for the block-sparse products
  DevBuf<double> tFold;
  DevBuf<int> tColOff;   // pairs beyond the band of G in the last refinement (whole G formed)
  int tCur = 0;
  // the same per k-point (all time-reversal invariant): the transposed eigenvectors of point k
  // in tXtK[2 k + tCurK[k]]
  std::vector<DevBuf<double>> tXtK;
  std::vector<int> tCurK;
  bool xFullK = false;   // XK holds all n FP64 eigenvectors of every point (the tail's seed)
  int lastClipped = 0;   // pairs outside the first-order range in the last refinement step
  bool sccDone = false;
  // SCC cycles start from the input charges the previous cycle ended with (the reference's
  // geometry loop keeps qInput between geometry steps, main.F90:219-357) instead of the
  // neutral reference charges
  bool warmStart = false;
  double times[kTCount]{};
  int ncalls = 0;
};

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int ctx_geometry(Context& c, int nat, const double* coords, const int* species,
                        const double* latvecs) {
  hipStream_t s = c.s;
  const double t0 = now();
  c.nat = nat;
  c.coords.assign(coords, coords + 3 * nat);
  c.species.assign(species, species + nat);
  c.periodic = latvecs != nullptr;
  if (latvecs) std::copy(latvecs, latvecs + 9, c.lat);
  const int ns = c.sk.nsp;
  c.norbAtom.resize(nat);
  c.atomStart.assign(nat + 1, 0);
  c.q0atom.resize(nat);
  for (int a = 0; a < nat; ++a) {
    DH_CHECK(species[a] >= 0 && species[a] < ns, "set_geometry: species out of range");
    c.norbAtom[a] = c.sk.norb[species[a]];
    c.atomStart[a + 1] = c.atomStart[a] + c.norbAtom[a];
    c.q0atom[a] = c.sk.q0[species[a]];
  }
  c.norbTot = (int)c.atomStart[nat];
  double cutoff = c.sk.skcut;
  for (double r : c.sk.repcut) cutoff = std::max(cutoff, r);
  for (double g : c.sk.gcut) cutoff = std::max(cutoff, g);
  c.nlOnDevice = g_devNeighbours != 0;
  if (c.nlOnDevice) {
    // neighbour list, pairs and SK blocks built on the device
    DH_TRY(device_neighbours(s, nat, coords, c.periodic ? c.lat : nullptr, cutoff, c.sk.skcut,
                             c.norbAtom.data(), c.dnl));
    c.nnz = c.dnl.nnz;
    DH_TRY(c.dblk.from_device(c.dnl.nBlk, c.dnl.bAt1.p, c.dnl.bImg2.p, c.dnl.bAt2.p,
                              c.dnl.bStart.p, c.species, c.norbAtom, c.atomStart, s));
    DH_TRY(c.dnb.from_device(c.dnl.nPairs, c.dnl.pAt1.p, c.dnl.pImg2.p, c.dnl.pAt2.p, nullptr,
                             c.species, c.norbAtom, c.atomStart, s));
    DH_TRY(c.coordsAll.alloc(3 * (size_t)c.dnl.nAll));
    DH_HIP(hipMemcpyAsync(c.coordsAll.p, c.dnl.coordsAll.p, 3 * (size_t)c.dnl.nAll * sizeof(double),
                          hipMemcpyDeviceToDevice, s));
  } else {
    DH_TRY(build_neighbours(nat, coords, c.periodic ? c.lat : nullptr, cutoff, c.nl));
    // SK blocks and all neighbour pairs
    c.blk = HostBlocks();
    c.nbp = HostBlocks();
    long st = 0;
    const double sk2 = c.sk.skcut * c.sk.skcut;
    for (int a = 0; a < nat; ++a) {
      const auto& nb = c.nl.nb[a];
      const auto& d2 = c.nl.d2[a];
      for (size_t k = 0; k < nb.size(); ++k) {
        const int img = nb[k], jc = c.nl.img2cent[img];
        c.nbp.at1.push_back(a);
        c.nbp.img2.push_back(img);
        c.nbp.at2.push_back(jc);
        c.nbp.start.push_back(0);
        if (d2[k] <= sk2) {
          c.blk.at1.push_back(a);
          c.blk.img2.push_back(img);
          c.blk.at2.push_back(jc);
          c.blk.start.push_back(st);
          st += (long)c.norbAtom[a] * c.norbAtom[jc];
        }
      }
    }
    c.nnz = st;
    DH_TRY(c.dblk.upload(c.blk, c.species, c.norbAtom, c.atomStart, s));
    DH_TRY(c.dnb.upload(c.nbp, c.species, c.norbAtom, c.atomStart, s));
    DH_TRY(c.coordsAll.upload(c.nl.coordsAll.data(), c.nl.coordsAll.size(), s));
  }
  DH_TRY(c.dblk.build_rows(nat, s));   // row view of the blocks (Mulliken, tail products)
  c.nbdev.n = c.dnb.pr.nblk;
  c.nbdev.at1 = c.dnb.pr.at1;
  c.nbdev.img2 = c.dnb.pr.img2;
  c.nbdev.at2 = c.dnb.pr.at2;
  c.nbdev.species = c.dnb.pr.species;
  DH_TRY(c.coordsC.upload(coords, 3 * (size_t)nat, s));
  const int n = c.norbTot;
  DH_TRY(c.h0.alloc(c.nnz));
  DH_TRY(c.ov.alloc(c.nnz));
  DH_TRY(c.hs.alloc(c.nnz));
  DH_TRY(c.rho.alloc(c.nnz));
  DH_TRY(c.erho.alloc(c.nnz));
  DH_TRY(c.vorb.alloc(n));
  DH_TRY(c.vatom.alloc(nat));
  DH_TRY(c.dq.alloc(nat));
  DH_TRY(c.qorb.alloc(n));
  DH_TRY(c.eorb.alloc(n));
  DH_TRY(c.f.alloc(n));
  DH_TRY(c.feps.alloc(n));
  DH_TRY(c.grad.alloc(3 * (size_t)nat));
  DH_TRY(c.eatom.alloc(nat));
  // k-points: complex problems as their real embedding of size 2n; time-reversal-invariant
  // points (2k integral) as real n x n problems
  int nd = n;
  for (int k = 0; k < c.nk; ++k) nd = std::max(nd, kpoint_dim(&c.kpts[3 * k], n));
  DH_CHECK(c.nk == 0 || c.periodic, "set_geometry: k-points need lattice vectors");
  DH_TRY(c.Hd.alloc((size_t)nd * nd));
  DH_TRY(c.Sd.alloc((size_t)nd * nd));
  DH_TRY(c.X.alloc((size_t)nd * nd));
  DH_TRY(c.P.alloc((size_t)nd * nd));
  if (c.nk > 0) {
    // fractional cell translation of each block's image: T = (r_img - r_centre) L^-1
    const double* L = c.lat;
    const double det = L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) +
                       L[2] * (L[3] * L[7] - L[4] * L[6]);
    double inv[9];
    inv[0] = (L[4] * L[8] - L[5] * L[7]) / det;
    inv[1] = -(L[1] * L[8] - L[2] * L[7]) / det;
    inv[2] = (L[1] * L[5] - L[2] * L[4]) / det;
    inv[3] = -(L[3] * L[8] - L[5] * L[6]) / det;
    inv[4] = (L[0] * L[8] - L[2] * L[6]) / det;
    inv[5] = -(L[0] * L[5] - L[2] * L[3]) / det;
    inv[6] = (L[3] * L[7] - L[4] * L[6]) / det;
    inv[7] = -(L[0] * L[7] - L[1] * L[6]) / det;
    inv[8] = (L[0] * L[4] - L[1] * L[3]) / det;
    if (c.nlOnDevice) {
      DH_TRY(c.blkT.alloc(3 * (size_t)std::max(c.dnl.nBlk, 1)));
      DH_TRY(block_translations(s, c.dnl.nBlk, c.dnl.bImg2.p, c.dnl.bAt2.p, c.coordsAll.p,
                                c.coordsC.p, inv, c.blkT.p));
    } else {
      const size_t nb = c.blk.at1.size();
      std::vector<double> T(3 * std::max<size_t>(nb, 1), 0.0);
      for (size_t k = 0; k < nb; ++k) {
        const int img = c.blk.img2[k], jc = c.blk.at2[k];
        double d[3];
        for (int q = 0; q < 3; ++q) d[q] = c.nl.coordsAll[3 * img + q] - coords[3 * jc + q];
        for (int q = 0; q < 3; ++q)
          T[3 * k + q] = std::nearbyint(d[0] * inv[q] + d[1] * inv[3 + q] + d[2] * inv[6 + q]);
      }
      DH_TRY(c.blkT.upload(T.data(), T.size(), s));
    }
  }
  DH_TRY(c.gamma.alloc((size_t)nat * nat));
  // H0, S
  DH_TRY(launch_build_h0s(s, c.sk.dev, c.dblk.pr, c.coordsAll.p, c.h0.p, c.ov.p));
  // gamma matrix
  DH_TRY(fill<double>(s, c.gamma.p, (size_t)nat * nat, 0.0));
  if (!c.periodic) {
    DH_TRY(launch_invr_cluster(s, nat, c.coordsC.p, c.gamma.p));
  } else {
    ewald_setup(c.lat, c.opt.ewald_tol, c.ew);
    const int ng = (int)c.ew.w.size();
    const int ntr = (int)c.ew.tr.size() / 3;
    DH_TRY(c.ewTr.upload(c.ew.tr.data(), c.ew.tr.size(), s));
    DH_TRY(launch_ewald_real(s, nat, c.coordsC.p, ntr, c.ewTr.p, c.ew.alpha, c.ew.maxr,
                             -M_PI / (c.ew.vol * c.ew.alpha * c.ew.alpha),
                             -2.0 * c.ew.alpha / std::sqrt(M_PI), c.gamma.p));
    if (ng > 0) {
      DH_TRY(c.ewG.upload(c.ew.G.data(), c.ew.G.size(), s));
      DH_TRY(c.ewW.upload(c.ew.w.data(), ng, s));
      std::vector<double> w2(2 * ng);
      for (int q = 0; q < ng; ++q) w2[q] = w2[q + ng] = c.ew.w[q];
      DH_TRY(c.ww2.upload(w2.data(), 2 * ng, s));
      DH_TRY(c.ewCS.alloc((size_t)nat * 2 * ng));
      DH_TRY(c.ewCSq.alloc(2 * (size_t)ng));
      DH_TRY(launch_strucfac(s, nat, c.coordsC.p, ng, c.ewG.p, c.ewCS.p));
      DH_TRY(gemm<double>(s, false, true, nat, nat, 2 * ng, 1.0, c.ewCS.p, nat, c.ewCS.p, nat, 1.0,
                          c.gamma.p, nat, false, kKFull, c.ww2.p));
      DH_HIP(hipStreamSynchronize(s));
    }
  }
  DH_TRY(launch_short_gamma(s, c.nbdev, c.coordsAll.p, c.sk.dhubbu.p, c.sk.dgcut.p, ns, nat,
                            c.gamma.p));
  // overlap factorisation (S fixed during the SCC cycle); per k-point with k-points
  DH_TRY(c.eig.ensure(nd));
  c.eig.factored = false;
  c.eig.twoStage = c.opt.tridiag == 0;
  int info = 0;
  if (c.nk == 0) {
    DH_TRY(launch_unpack_rows<double>(s, c.dblk.rows, c.dblk.norb.p, c.dblk.atomStart.p,
                                      c.dblk.start.p, c.ov.p, c.Sd.p,
                                 n, n));
    DH_TRY(c.eig.factor_overlap(s, c.Sd.p, n, &info));
  } else {
    c.LinvK.resize(c.nk);
    c.XK.resize(c.nk);
    c.eigK.assign(c.nk, std::vector<double>());
    c.fillK.assign(c.nk, std::vector<double>());
    c.noccK.assign(c.nk, 0);
    for (int k = 0; k < c.nk && info == 0; ++k) {
      const int ndk = kpoint_dim(&c.kpts[3 * k], n);
      DH_TRY(launch_unpack_kpt(s, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p, c.blkT.p,
                               &c.kpts[3 * k], c.ov.p, c.Sd.p, ndk, n));
      DH_TRY(c.eig.ensure(ndk));
      DH_TRY(c.eig.factor_overlap(s, c.Sd.p, ndk, &info));
      if (info != 0) break;
      DH_TRY(c.LinvK[k].alloc((size_t)ndk * ndk));
      DH_TRY(c.XK[k].alloc((size_t)ndk * ndk));
      DH_TRY(copy2d<double>(s, c.eig.Linv.p, ndk, c.LinvK[k].p, ndk, ndk, ndk));
    }
  }
  if (info != 0) {
    set_error("set_geometry: overlap matrix not positive definite (minor " +
              std::to_string(info) + ")");
    return nd + info;
  }
  if (c.opt.precision == 1) {
    // k-points: the eigensolves stay FP64 (refine_steps = 0), the mixed form is the tail
    DH_CHECK(c.nk == 0 || (c.opt.refine_steps == 0 && c.opt.tail_below > 0.0),
             "set_geometry: with k-points the mixed precision runs FP64 eigensolves and the "
             "tail only (refine_steps 0, tail_below > 0)");
    if (c.nk == 0) {
      DH_TRY(c.eig.prepare_mixed(s));
      DH_TRY(c.f32.alloc(n));
    }
  }
  if (c.nk > 0) DH_TRY(c.f.alloc(nd));
  DH_HIP(hipStreamSynchronize(s));
  c.sccDone = false;
  c.times[kTSetup] = now() - t0;
  return 0;
}

// C = A B^T (n x n, K = k, w: per-k weights) for the density and energy-density matrices,
// formed only on the tiles pack reads (the sparse blocks' tiles) when the row view holds them,
// else on the whole lower triangle
template <typename T>
static int pack_product(Context& c, int n, int k, const T* A, const T* B, T* C,
                        const T* w = nullptr) {
  if (c.dblk.rowsValid && c.dblk.nPackTiles > 0)
    return gemm_tiles<T>(c.s, false, true, n, n, k, T(1), A, n, B, n, T(0), C, n,
                         c.dblk.packTiles.p, c.dblk.nPackTiles, w);
  return gemm<T>(c.s, false, true, n, n, k, T(1), A, n, B, n, T(0), C, n, true, kKFull, w);
}

static int gamma_times(Context& c, const std::vector<double>& dqa, std::vector<double>* vout) {
  hipStream_t s = c.s;
  DH_HIP(hipMemcpyAsync(c.dq.p, dqa.data(), c.nat * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(gemv_t_kernel, dim3(cdiv(c.nat, 4)), dim3(256), 0, s, c.nat, c.gamma.p, c.dq.p,
                     c.vatom.p);
  DH_LAUNCH_CHECK();
  hipLaunchKernelGGL(expand_kernel, dim3(cdiv(c.nat, 256)), dim3(256), 0, s, c.nat,
                     c.dblk.atomStart.p, c.vatom.p, c.vorb.p);
  DH_LAUNCH_CHECK();
  if (vout) {
    vout->resize(c.nat);
    DH_HIP(hipMemcpyAsync(vout->data(), c.vatom.p, c.nat * sizeof(double), hipMemcpyDeviceToHost, s));
    DH_HIP(hipStreamSynchronize(s));
  }
  return 0;
}

// Levels entering the density matrices: up to the last one with a filling of at least
// machine epsilon (fullDensityMatrix_real, densitymatrix.F90:549-554).
static int occupied_levels(const std::vector<double>& f) {
  for (int i = (int)f.size() - 1; i >= 0; --i)
    if (std::fabs(f[i]) >= std::numeric_limits<double>::epsilon()) return i + 1;
  return 0;
}

// Ogita-Aishima refinement (Ogita & Aishima 2018) of all n generalized eigenpairs held
// transposed in xt[cur] (FP64, n x n; c.tXt at the Gamma point, c.tXtK[k] at a
// time-reversal-invariant k-point, whose matrices are the real sums of the blocks with the
// phases +-1 of kf), against the sparse H (c.hs) and S (c.ov) of the current iteration. Per step: YH^T = (H X)^T and YS^T = (S X)^T by the block-sparse product
// (H and S stay sparse: 2 nnz n flops instead of 2 n^3), M = X^T H X and G = X^T S X as
// lower-output GEMMs (n^3 each), the correction E written transposed, and X <- X (I + E), i.e.
// Xt <- Xt + E^T Xt (2 n^3). The refined vectors end in Xout (normal layout), the Rayleigh
// quotients of the last input basis (unsorted) in lamHost.
static int oa_refine_sparse_on(Context& c, int steps, DevBuf<double>* xt, int& cur, double* Xout,
                               const double* kf, std::vector<double>& lamHost) {
  hipStream_t s = c.s;
  const int n = c.norbTot;
  const size_t nn = (size_t)n * n;
  DH_TRY(c.dblk.build_rows(c.nat, s));
  DH_TRY(xt[1 - cur].alloc(nn));
  DH_TRY(c.tYH.alloc(nn));
  DH_TRY(c.tYS.alloc(nn));
  DH_TRY(c.tM.alloc(nn));
  DH_TRY(c.tG.alloc(nn));
  DH_TRY(c.tLam.alloc(n));
  DH_TRY(c.tClip.alloc(1));
  const int ng = c.dblk.nGroups;
  DH_TRY(c.tFold.alloc((size_t)std::max(ng, 1) * 32));
  DH_TRY(c.tColOff.alloc((size_t)std::max(ng, 1)));
  DH_TRY(launch_spmm_fold(s, c.dblk.rows, c.dblk.norb.p, c.dblk.atomStart.p, c.dblk.start.p,
                          c.hs.p, c.ov.p, kf ? c.blkT.p : nullptr, kf, c.tFold.p, c.tColOff.p));
  const int mode =
      (g_tailUpdate32 == 2 && c.st.prevErr <= g_tailFineBelow) ? 1 : g_tailUpdate32;
  const long ldh = f16_ld(n);
  if (mode == 2) {
    // the padding of the FP16 operands must be zero: cleared once per call (the steps write the
    // n x n parts only)
    const size_t nh = (size_t)ldh * f16_cols(n);
    for (DevBuf<_Float16>* b : {&c.hX, &c.hXt, &c.hR, &c.hE}) {
      DH_TRY(b->alloc(nh));
      DH_TRY(fill<_Float16>(s, b->p, nh, (_Float16)0.0f));
    }
    DH_TRY(c.tPart.alloc((size_t)64 * n));
    DH_TRY(c.tFar.alloc(1));
    c.lastFar = 0;
    const int band = g_tailBand;
    if (band >= 0) {
      DH_TRY(c.tMb.alloc(nn));
      DH_TRY(c.tCt.alloc(nn));
      const int nt = (int)cdiv(n, kGemmTile);
      c.tDefl.resize(2 * band + 1);
      for (int o = -band; o <= band; ++o) {
        std::vector<GemmDesc<double>> d;
        for (int bj = 0; bj < nt; ++bj) {
          const int bi = bj + o;
          if (bi < 0 || bi >= nt) continue;
          const int m = std::min(kGemmTile, n - bj * kGemmTile);
          const int k = std::min(kGemmTile, n - bi * kGemmTile);
          d.push_back(GemmDesc<double>{m, n, k,
                                       c.tCt.p + (size_t)bj * kGemmTile +
                                           (size_t)bi * kGemmTile * n,
                                       (long)n, c.tYS.p + (size_t)bi * kGemmTile, (long)n,
                                       c.tYH.p + (size_t)bj * kGemmTile, (long)n});
        }
        auto& b = c.tDefl[o + band];
        b.count = (int)d.size();
        if (!d.empty()) DH_TRY(b.desc.upload(d.data(), d.size(), s));
      }
    }

    DH_TRY(c.tRsc.alloc(n));
    DH_TRY(c.tRinv.alloc(n));
  }
  for (int st = 0; st < std::max(steps, 1); ++st) {
    DH_TRY(fill<int>(s, c.tClip.p, 1, 0));
    const double* Xt = xt[cur].p;
    double* Xn = xt[1 - cur].p;
    DH_TRY(launch_spmm_hs_t(s, c.dblk.rows, c.dblk.norb.p, c.dblk.atomStart.p, n, c.tFold.p,
                            c.tColOff.p, Xt, n, n, c.tYH.p, c.tYS.p, n));
    if (mode == 2) {
      // Residual form. G = X^T S X and M = X^T H X on their diagonal tile band (FP64), the
      // Rayleigh quotients lam_j = M_jj / G_jj and the in-band couplings C_ij = x_i^T r_j =
      // M_ij - lam_j G_ij (r_j = (H - lam_j S) x_j). The in-band part of each residual,
      // sum_i S x_i C_ij, is removed from H X in FP64: what is left is small once the cycle
      // converges (near-degenerate pairs outside the first-order range keep C_ij of the order of
      // their level spacing, which would otherwise stay in every residual). The deflated
      // residuals are scaled per vector by a power of two into the FP16 range; Ft = R^T X
      // (FP16 operands, FP32 sums, unscaled per row) gives the couplings beyond the band; E
      // from C (in the band), Ft (beyond) and G, per column scaled into FP16; Xt <- Xt + E^T Xt
      // with the product in FP16. The rounding of the FP16 products is a relative 2^-11-level
      // error of the correction, removed by the next step like any other basis error.
      const int band = g_tailBand;
      int gband = band;
      if (band >= 0) {
        DH_TRY(gemm_band<double>(s, false, true, n, n, 1.0, Xt, n, c.tYS.p, n, c.tG.p, n, band,
                                 8));
        DH_TRY(gemm_band<double>(s, false, true, n, n, 1.0, Xt, n, c.tYH.p, n, c.tMb.p, n, band,
                                 8));
        DH_TRY(oa_band_coupling(s, c.tMb.p, c.tG.p, n, n, band, c.tLam.p, c.tCt.p, n));
        // YH <- YH - C^T YS over the band, one batched launch per tile offset (disjoint outputs)
        for (int o = -band; o <= band; ++o) {
          auto& d = c.tDefl[o + band];
          if (d.count > 0)
            DH_TRY(gemm_batched<double>(s, false, false, d.desc.p, d.count, kGemmTile, n, -1.0,
                                        1.0));
        }
      } else {
        DH_TRY(gemm<double>(s, false, true, n, n, n, 1.0, Xt, n, c.tYS.p, n, 0.0, c.tG.p, n,
                            true));
        DH_TRY(oa_rayleigh(s, Xt, c.tYH.p, n, n, n, c.tG.p, n, c.tPart.p, c.tLam.p));
      }
      const double* Ct = band >= 0 ? c.tCt.p : nullptr;
      DH_TRY(oa_residual_scale(s, c.tYH.p, c.tYS.p, n, n, n, c.tLam.p, c.tPart.p, c.tRsc.p,
                               c.tRinv.p));
      DH_TRY(f16_convert(s, c.tYH.p, c.tYS.p, c.tLam.p, c.tRsc.p, n, n, n, nullptr, 0, c.hR.p,
                         ldh));
      DH_TRY(f16_convert(s, Xt, nullptr, nullptr, nullptr, n, n, n, c.hXt.p, ldh, c.hX.p, ldh));
      float* Ft = reinterpret_cast<float*>(c.tM.p);
      DH_TRY(gemm_f16_tn(s, n, n, n, 1.0f, c.hR.p, ldh, c.hX.p, ldh, Ft, n, c.tRinv.p));
      unsigned* colMax = reinterpret_cast<unsigned*>(c.tRsc.p);   // free once R is converted
      DH_TRY(fill<int>(s, c.tFar.p, 1, 0));
      DH_TRY(oa_correction_f16(s, Ft, n, c.tG.p, n, Ct, n, c.tLam.p, n, band, gband, colMax,
                               c.tClip.p, c.tFar.p));
      int far = 0;
      if (gband >= 0) {
        DH_HIP(hipMemcpyAsync(&far, c.tFar.p, sizeof(int), hipMemcpyDeviceToHost, s));
        DH_HIP(hipStreamSynchronize(s));
      }
      if (far > 0) {
        // a pair outside the first-order range beyond the band: the whole lower triangle of G
        c.lastFar += far;
        gband = -1;
        DH_TRY(gemm<double>(s, false, true, n, n, n, 1.0, Xt, n, c.tYS.p, n, 0.0, c.tG.p, n,
                            true));
        DH_TRY(fill<int>(s, c.tClip.p, 1, 0));
        DH_TRY(oa_correction_f16(s, Ft, n, c.tG.p, n, Ct, n, c.tLam.p, n, band, -1, colMax,
                                 c.tClip.p, nullptr));
      }
      // E per column scaled by a power of two into the FP16 range (the corrections shrink as
      // the cycle converges; a fixed scale would leave the late ones subnormal)
      DH_TRY(oa_correction_f16_write(s, Ft, n, c.tG.p, n, Ct, n, c.tLam.p, n, band, gband,
                                     colMax, c.hE.p, ldh, c.tRinv.p));
      float* D32 = reinterpret_cast<float*>(c.tYH.p);
      DH_TRY(gemm_f16_tn(s, n, n, n, 1.0f, c.hE.p, ldh, c.hXt.p, ldh, D32, n, c.tRinv.p));
      DH_TRY(add_f32(s, Xt, n, D32, n, Xn, n, n, n));
      cur = 1 - cur;
      continue;
    }
    DH_TRY(gemm<double>(s, false, true, n, n, n, 1.0, Xt, n, c.tYH.p, n, 0.0, c.tM.p, n, true));
    DH_TRY(gemm<double>(s, false, true, n, n, n, 1.0, Xt, n, c.tYS.p, n, 0.0, c.tG.p, n, true));
    DH_TRY(symmetrize_lower(s, c.tM.p, n, n));
    DH_TRY(symmetrize_lower(s, c.tG.p, n, n));
    double* Et = c.tYH.p;   // free once M is formed
    DH_TRY(oa_correction_t(s, c.tM.p, c.tG.p, n, c.tLam.p, Et, c.tClip.p));
    if (g_tailUpdate32) {
      // the correction product in FP32: its rounding is a relative 1e-7-level error of the
      // correction, which the next step corrects like any other basis error (the fixed point
      // stays that of the FP64 residuals M and G); FP32 operands in the free YS and M buffers
      float* E32 = reinterpret_cast<float*>(c.tYS.p);
      float* X32 = E32 + (size_t)n * n;
      float* D32 = reinterpret_cast<float*>(c.tM.p);
      DH_TRY((convert2d<double, float>(s, Et, n, E32, n, n, n)));
      DH_TRY((convert2d<double, float>(s, Xt, n, X32, n, n, n)));
      DH_TRY(gemm<float>(s, false, false, n, n, n, 1.0f, E32, n, X32, n, 0.0f, D32, n));
      DH_TRY(add_f32(s, Xt, n, D32, n, Xn, n, n, n));
    } else {
      DH_TRY(copy2d<double>(s, Xt, n, Xn, n, n, n));
      DH_TRY(gemm<double>(s, false, false, n, n, n, 1.0, Et, n, Xt, n, 1.0, Xn, n));
    }
    cur = 1 - cur;
  }
  DH_TRY((transpose2d<double, double>(s, xt[cur].p, n, Xout, n, n, n)));
  lamHost.resize(n);
  DH_HIP(hipMemcpyAsync(lamHost.data(), c.tLam.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(&c.lastClipped, c.tClip.p, sizeof(int), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

static int oa_refine_sparse(Context& c, int steps, std::vector<double>& lamHost) {
  return oa_refine_sparse_on(c, steps, c.tXt, c.tCur, c.X.p, nullptr, lamHost);
}

// Levels beyond the occupied ones kept in the refined density matrix: refined eigenvalues
// move by the FP32 error of the tridiagonalisation, which can reorder levels at the cut.
constexpr int kRefineMargin = 64;

static int ctx_scc_begin(Context& c) {
  for (int t = 1; t < kTCount; ++t) c.times[t] = 0.0;
  SccState& st = c.st;
  st.nel = 0.0;
  for (double q : c.q0atom) st.nel += q;
  // given input charges (dftbhip_ctx_set_input_charges), else a warm start: the input charges
  // of the previous cycle, when it was on the same atoms; else the neutral atoms
  if (!c.qinGiven.empty() && c.qinGiven.size() == c.q0atom.size()) {
    st.qin = c.qinGiven;
    c.qinGiven.clear();
  } else if (!(c.warmStart && st.it > 0 && st.qin.size() == c.q0atom.size())) {
    st.qin = c.q0atom;
  }
  st.mixer.reset(c.nat, c.opt, c.opt.max_scc_iter + 1);
  c.xFull = false;
  c.xFullK = false;
  c.tailOff = false;
  st.it = 0;
  st.conv = false;
  st.err = 0.0;
  st.prevErr = HUGE_VAL;
  st.eElec = 0.0;
  st.eElecPrev = std::numeric_limits<double>::quiet_NaN();
  st.active = true;
  c.sccDone = false;
  return 0;
}

// One SCC iteration (main.F90:1429-1563): potential from the input charges, H, generalized
// eigenproblem, Fermi filling, density matrix, Mulliken charges, energies with the output
// potential, convergence test and Broyden mixing.
static int ctx_scc_step(Context& c) {
  DH_CHECK(c.st.active, "scc_step: call scc_begin first");
  hipStream_t s = c.s;
  const int nat = c.nat, n = c.norbTot;
  const dftbhip_options& o = c.opt;
  SccState& st = c.st;
  st.it += 1;
  const double tIt = now();
  std::vector<double> dqa(nat), qorbH(n), eorbH(n), vout;
  for (int a = 0; a < nat; ++a) dqa[a] = st.qin[a] - c.q0atom[a];
  Events<8> ev;
  DH_TRY(ev.create());
  DH_HIP(hipEventRecord(ev[0], s));
  DH_TRY(gamma_times(c, dqa, nullptr));
  DH_TRY(launch_shift(s, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p, c.h0.p, c.ov.p, c.vorb.p,
                      c.hs.p));
  double evt[3] = {0, 0, 0};
  // mixed precision: below the switch point (charge error of the previous iteration at most
  // refine_below) the iteration is refined in FP64, or with refine_steps = 0 runs in FP64
  const bool below = o.precision == 1 && st.prevErr <= o.refine_below;
  const bool mixed = o.precision == 1 && !(below && o.refine_steps == 0);
  const bool refined = mixed && below;
  // tail refinement: the previous iteration's n refined eigenpairs are refined against the new
  // H by one Ogita-Aishima step, without an eigensolve
  const bool keepAll = o.tail_below > 0.0;
  // the iteration that starts the tail (the first one below refine_below) keeps all n FP64
  // eigenvectors: refined FP32 ones (refine_steps > 0) or those of a plain FP64 iteration
  // (refine_steps = 0, the more accurate seed: the FP32 tridiagonalisation mixes levels closer
  // than its error, which the refinement cannot separate)
  const bool seedAll = keepAll && below && c.nk == 0;
  const bool tail =
      below && keepAll && c.xFull && c.nk == 0 && st.prevErr <= o.tail_below && !c.tailOff;
  bool tailUsed = tail;
  const double prevErrIn = st.prevErr;
  bool refineTimed = refined || tail;   // events 6 and 7 bracket a refinement
  if (c.nk > 0) {
    // k-points: per k H(k) (real n x n at time-reversal-invariant points, else the 2n
    // embedding), eigenpairs and all eigenvectors; then the k-weighted Fermi level and the
    // density matrix k by k (packHS_cmplx_kpts)
    c.xf64 = true;
    DH_HIP(hipEventRecord(ev[1], s));
    std::vector<std::vector<double>> ec(c.nk);
    bool allReal = true;
    for (int k = 0; k < c.nk; ++k) allReal = allReal && kpoint_dim(&c.kpts[3 * k], n) == n;
    // the tail at k-points (every point time-reversal invariant, mixed precision): no
    // eigensolve, Ogita-Aishima steps on each point's previous FP64 eigenpairs against its new
    // sparse H(k) and S(k); the eigenvalues of a point in the order of its vectors
    const bool ktail =
        keepAll && below && allReal && c.xFullK && st.prevErr <= o.tail_below && !c.tailOff;
    tailUsed = ktail;
    std::vector<std::vector<int>> kperm;
    if (ktail) {
      kperm.resize(c.nk);
      DH_HIP(hipEventRecord(ev[6], s));
      for (int k = 0; k < c.nk; ++k) {
        std::vector<double> lam;
        DH_TRY(oa_refine_sparse_on(c, std::max(o.tail_steps, 1), &c.tXtK[2 * (size_t)k],
                                   c.tCurK[k], c.XK[k].p, &c.kpts[3 * k], lam));
        std::vector<int>& perm = kperm[k];
        perm.resize(n);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return lam[x] < lam[y]; });
        ec[k].resize(n);
        for (int j = 0; j < n; ++j) ec[k][j] = lam[perm[j]];
        c.eigK[k] = lam;
      }
      DH_HIP(hipEventRecord(ev[7], s));
      refineTimed = true;
    }
    bool together = !ktail && g_kBatch && c.nk > 1 && !c.eig.twoStage && allReal;
    if (together) {
      // every point real and of size n: the reductions per point, one tridiagonalisation pass
      // over all points (one launch per column kernel for all of them), then per point the
      // divide and conquer and the back-transformation
      const int K = c.nk;
      const size_t nn = (size_t)n * n, wsStride = sytrd_work_bytes(n, 8);
      DH_TRY(c.eig.ensure(n));
      DH_TRY(c.kA.alloc(K * nn));
      DH_TRY(c.kD.alloc((size_t)K * n));
      DH_TRY(c.kE.alloc((size_t)K * n));
      DH_TRY(c.kTau.alloc((size_t)K * n));
      DH_TRY(c.kWs.alloc(K * wsStride));
      Events<3> kev;
      DH_TRY(kev.create());
      DH_HIP(hipEventRecord(kev[0], s));
      for (int k = 0; k < K; ++k) {
        DH_TRY(launch_unpack_kpt(s, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p, c.blkT.p,
                                 &c.kpts[3 * k], c.hs.p, c.Hd.p, n, n));
        DH_TRY(c.eig.reduce(s, c.LinvK[k].p, c.Hd.p, n, c.kA.p + k * nn));
      }
      DH_HIP(hipEventRecord(kev[1], s));
      DH_TRY(sytrd_lower_batched(s, c.kA.p, n, (long)nn, n, c.kD.p, c.kE.p, c.kTau.p, n,
                                 c.kWs.p, wsStride, K));
      DH_HIP(hipEventRecord(kev[2], s));
      DH_HIP(hipEventSynchronize(kev[2]));
      float ms;
      DH_HIP(hipEventElapsedTime(&ms, kev[0], kev[1]));
      evt[0] += ms * 1e-3;
      DH_HIP(hipEventElapsedTime(&ms, kev[1], kev[2]));
      evt[1] += ms * 1e-3;
      // divide and conquer and back-transformation of the points side by side, one host thread
      // and stream each: a D&C is mostly host deflation between short device phases, which the
      // other points' device work fills
      if ((int)c.kdc.size() < K) c.kdc.resize(K);
      // one workspace per point: the D&C's, then the product L^-T (Q Z) and the reflector
      // workspace of the back-transformation
      const size_t dcWork = std::max(stedc_work_size(n), nn + ormtr_work_size(n, n));
      for (int k = 0; k < K; ++k) {
        if (!c.kdc[k]) c.kdc[k].reset(new Context::KPointDc());
        Context::KPointDc& r = *c.kdc[k];
        if (!r.st) DH_HIP(hipStreamCreate(&r.st));
        if (!r.ws) r.ws.reset(stedc_ws_new(), stedc_ws_delete);
        DH_TRY(r.work.alloc(dcWork));
        DH_TRY(r.w.alloc(n));
      }
      int dev = 0;
      DH_HIP(hipGetDevice(&dev));
      std::vector<int> rcK(K, 0), infoK(K, 0);
      std::vector<std::string> msgK(K);
      auto dcPointBody = [&](int k) -> int {
        Context::KPointDc& r = *c.kdc[k];
        DH_HIP(hipSetDevice(dev));
        std::vector<double> dh(n), eh(std::max(n - 1, 1));
        DH_HIP(hipMemcpyAsync(dh.data(), c.kD.p + (size_t)k * n, n * sizeof(double),
                              hipMemcpyDeviceToHost, r.st));
        if (n > 1)
          DH_HIP(hipMemcpyAsync(eh.data(), c.kE.p + (size_t)k * n, (n - 1) * sizeof(double),
                                hipMemcpyDeviceToHost, r.st));
        DH_HIP(hipStreamSynchronize(r.st));
        DH_TRY(stedc(r.st, n, dh.data(), eh.data(), r.w.p, c.XK[k].p, n, r.work.p, dcWork,
                     &infoK[k], r.ws.get()));
        if (infoK[k] != 0) return 0;
        c.eigK[k].resize(n);
        DH_HIP(hipMemcpyAsync(c.eigK[k].data(), r.w.p, n * sizeof(double), hipMemcpyDeviceToHost,
                              r.st));
        // back-transformation on the same stream, the D&C workspace (free now) holding the
        // product L^-T (Q Z) and the reflector workspace: Q Z in place, product, copy back
        double* prod = r.work.p;
        double* owork = r.work.p + nn;
        DH_TRY(ormtr_lower(r.st, c.kA.p + k * nn, n, n, c.kTau.p + (size_t)k * n, c.XK[k].p, n,
                           n, owork, dcWork - nn));
        DH_TRY(gemm<double>(r.st, true, false, n, n, n, 1.0, c.LinvK[k].p, n, c.XK[k].p, n, 0.0,
                            prod, n, false, kKAUpper));
        DH_HIP(hipMemcpyAsync(c.XK[k].p, prod, nn * sizeof(double), hipMemcpyDeviceToDevice,
                              r.st));
        DH_HIP(hipStreamSynchronize(r.st));
        return 0;
      };
      // an error leaves the point's thread only after the work it queued has ended, so that no
      // kernel still writes the point's buffers when the caller reuses them
      auto dcPoint = [&](int k) -> int {
        const int rc = dcPointBody(k);
        if (rc != 0) (void)hipStreamSynchronize(c.kdc[k]->st);
        return rc;
      };
      const double tdc = now();
      {
        std::vector<std::thread> th;
        th.reserve(K);
        for (int k = 0; k < K; ++k)
          th.emplace_back([&, k]() {
            rcK[k] = dcPoint(k);
            if (rcK[k] != 0) msgK[k] = last_error();
          });
        for (auto& t : th) t.join();
      }
      evt[2] += now() - tdc;
      for (int k = 0; k < K; ++k) {
        if (rcK[k] != 0) {
          set_error(msgK[k]);
          return rcK[k];
        }
        if (infoK[k] != 0) return infoK[k];
      }
      for (int k = 0; k < K; ++k) ec[k] = c.eigK[k];
    }
    for (int k = 0; k < c.nk && !together && !ktail; ++k) {
      const int ndk = kpoint_dim(&c.kpts[3 * k], n);
      DH_TRY(launch_unpack_kpt(s, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p, c.blkT.p,
                               &c.kpts[3 * k], c.hs.p, c.Hd.p, ndk, n));
      DH_TRY(c.eig.ensure(ndk));
      DH_TRY(copy2d<double>(s, c.LinvK[k].p, ndk, c.eig.Linv.p, ndk, ndk, ndk));
      int info = 0;
      DH_TRY(c.eig.solve(s, c.Hd.p, ndk, c.eigK[k], &info, evt));
      if (info != 0) return info;
      DH_TRY(c.eig.vectors(s, ndk, c.XK[k].p, ndk));
      // each complex level appears twice in the embedding
      const int stride = ndk / n;
      ec[k].resize(n);
      for (int j = 0; j < n; ++j) ec[k][j] = c.eigK[k][stride * j];
    }
    DH_HIP(hipEventRecord(ev[2], s));
    const double th = now();
    DH_TRY(fermi_filling_k(ec, c.kw, st.nel, o.kT, c.fillK, st.fr));
    if (ktail) {
      // fillings back into the order of the vectors
      for (int k = 0; k < c.nk; ++k) {
        std::vector<double> fk(n);
        for (int j = 0; j < n; ++j) fk[kperm[k][j]] = c.fillK[k][j];
        c.fillK[k].swap(fk);
      }
    }
    c.times[kTHost] += now() - th;
    DH_HIP(hipEventRecord(ev[3], s));
    DH_TRY(fill<double>(s, c.rho.p, c.nnz, 0.0));
    std::vector<double> fe(2 * (size_t)n);
    for (int k = 0; k < c.nk; ++k) {
      const int ndk = kpoint_dim(&c.kpts[3 * k], n), stride = ndk / n;
      c.noccK[k] = occupied_levels(c.fillK[k]);
      for (int j = 0; j < n; ++j)
        for (int t = 0; t < stride; ++t) fe[stride * j + t] = c.fillK[k][j];
      DH_HIP(hipMemcpyAsync(c.f.p, fe.data(), ndk * sizeof(double), hipMemcpyHostToDevice, s));
      DH_TRY(gemm<double>(s, false, true, ndk, ndk, stride * c.noccK[k], 1.0, c.XK[k].p, ndk,
                          c.XK[k].p, ndk, 0.0, c.P.p, ndk, true, kKFull, c.f.p));
      DH_TRY(launch_pack_kpt(s, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p, c.blkT.p,
                             &c.kpts[3 * k], c.kw[k], c.P.p, ndk, n, c.rho.p));
      DH_HIP(hipStreamSynchronize(s));   // fe is reused for the next k-point
    }
    c.eigv.clear();
    DH_HIP(hipEventRecord(ev[4], s));
    // an FP64 iteration below the switch point seeds the k-point tail: each point's n vectors
    // transposed for the block-sparse products
    if (keepAll && below && allReal && !ktail) {
      if (c.tXtK.size() < 2 * (size_t)c.nk) c.tXtK.resize(2 * (size_t)c.nk);
      c.tCurK.assign(c.nk, 0);
      for (int k = 0; k < c.nk; ++k) {
        DH_TRY(c.tXtK[2 * (size_t)k].alloc((size_t)n * n));
        DH_TRY((transpose2d<double, double>(s, c.XK[k].p, n, c.tXtK[2 * (size_t)k].p, n, n, n)));
      }
      c.xFullK = true;
    }
  } else {
    // FP64: the plain path. Mixed: FP32 reduction, tridiagonalisation, back-transformation and
    // density matrix. Mixed + refined: the FP32 eigenvectors of all levels are refined in FP64
    // (Ogita-Aishima) and the density matrix is built in FP64.
    c.xf64 = !mixed || refined;
    float* X32 = reinterpret_cast<float*>(c.X.p);
    float* P32 = reinterpret_cast<float*>(c.P.p);
    float* H32 = refined ? c.eig.h32() : reinterpret_cast<float*>(c.Hd.p);
    if (tail) {
      DH_HIP(hipEventRecord(ev[1], s));
      DH_HIP(hipEventRecord(ev[2], s));
      DH_HIP(hipEventRecord(ev[6], s));
      std::vector<double> lam;
      DH_TRY(oa_refine_sparse(c, std::max(o.tail_steps, 1), lam));
      DH_HIP(hipEventRecord(ev[7], s));
      const double th3 = now();
      std::vector<int> perm(n);
      std::iota(perm.begin(), perm.end(), 0);
      std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return lam[x] < lam[y]; });
      std::vector<double> ls(n), fs;
      for (int k = 0; k < n; ++k) ls[k] = lam[perm[k]];
      DH_TRY(fermi_filling(ls, st.nel, o.kT, fs, st.fr));
      c.fill.assign(n, 0.0);
      for (int k = 0; k < n; ++k) c.fill[perm[k]] = fs[k];
      c.eigv = lam;
      // the density takes the columns up to the last one with a filling of at least eps
      const int nocc = occupied_levels(c.fill);
      DH_HIP(hipMemcpyAsync(c.f.p, c.fill.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
      c.times[kTHost] += now() - th3;
      c.nocc = nocc;
      c.xFull = true;
      DH_HIP(hipEventRecord(ev[3], s));
      double* Y = c.eig.T.p;
      DH_TRY(scale_cols_sqrt<double>(s, c.X.p, n, c.f.p, Y, n, n, nocc));
      DH_TRY(pack_product<double>(c, n, nocc, Y, Y, c.P.p));
      DH_HIP(hipEventRecord(ev[4], s));
      DH_TRY(launch_pack<double>(s, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p, c.P.p, n,
                                 c.rho.p));
    } else if (mixed && !refined) {
      DH_TRY(launch_unpack_rows<float>(s, c.dblk.rows, c.dblk.norb.p, c.dblk.atomStart.p,
                                       c.dblk.start.p, c.hs.p, H32, n,
                                  n));
    } else {
      DH_TRY(launch_unpack_rows<double>(s, c.dblk.rows, c.dblk.norb.p, c.dblk.atomStart.p,
                                        c.dblk.start.p, c.hs.p, c.Hd.p,
                                   n, n));
      if (refined) DH_TRY((convert2d<double, float>(s, c.Hd.p, n, H32, n, n, n)));
    }
    if (!tail) {
    DH_HIP(hipEventRecord(ev[1], s));
    int info = 0;
    if (mixed)
      DH_TRY(c.eig.solve_mixed(s, H32, n, c.eigv, &info, evt));
    else
      DH_TRY(c.eig.solve(s, c.Hd.p, n, c.eigv, &info, evt));
    if (info != 0) return info;
    const double th = now();
    DH_TRY(fermi_filling(c.eigv, st.nel, o.kT, c.fill, st.fr));
    int nocc = occupied_levels(c.fill);
    c.times[kTHost] += now() - th;
    std::vector<float> fill32;
    if (mixed && !refined) {
      fill32.assign(c.fill.begin(), c.fill.end());
      DH_HIP(hipMemcpyAsync(c.f32.p, fill32.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    } else if (!mixed) {
      DH_HIP(hipMemcpyAsync(c.f.p, c.fill.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    }
    DH_HIP(hipEventRecord(ev[2], s));
    if (refined) {
      // all levels form the refinement basis; the leading nocc + margin refined vectors enter
      // the density matrix with fillings recomputed from the refined eigenvalues
      DH_TRY(c.eig.vectors_mixed(s, n, P32, n));
      DH_HIP(hipEventRecord(ev[6], s));
      // with tail refinement every refined vector is kept (the next iteration refines them) and
      // the refinement runs on the sparse H and S with the vectors held transposed
      const int nvec = keepAll ? n : std::min(n, nocc + kRefineMargin);
      std::vector<double> lam;
      if (keepAll) {
        DH_TRY(c.tXt[c.tCur].alloc((size_t)n * n));
        DH_TRY((transpose2d<float, double>(s, P32, n, c.tXt[c.tCur].p, n, n, n)));
        DH_TRY(oa_refine_sparse(c, o.refine_steps, lam));
      } else {
        DH_TRY(c.eig.refine(s, c.Hd.p, c.Sd.p, P32, o.refine_steps, nvec, c.X.p, n, lam));
      }
      DH_HIP(hipEventRecord(ev[7], s));
      const double th3 = now();
      std::vector<int> perm(n);
      std::iota(perm.begin(), perm.end(), 0);
      std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return lam[x] < lam[y]; });
      std::vector<double> ls(n), fs;
      for (int k = 0; k < n; ++k) ls[k] = lam[perm[k]];
      DH_TRY(fermi_filling(ls, st.nel, o.kT, fs, st.fr));
      for (int k = 0; k < n; ++k) c.fill[perm[k]] = fs[k];
      c.eigv = lam;
      nocc = keepAll ? occupied_levels(c.fill) : nvec;
      DH_HIP(hipMemcpyAsync(c.f.p, c.fill.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
      c.times[kTHost] += now() - th3;
    } else if (mixed) {
      DH_TRY(c.eig.vectors_mixed(s, nocc, X32, n));
    } else {
      DH_TRY(c.eig.vectors(s, seedAll ? n : nocc, c.X.p, n));
    }
    c.nocc = nocc;
    DH_HIP(hipEventRecord(ev[3], s));
    if (c.xf64) {
      // P = Y Y^T with Y = X diag(sqrt f) (fillings are >= 0) staged in the eigensolver's
      // scratch matrix, which is free once the eigenvectors are formed
      double* Y = c.eig.T.p;
      DH_TRY(scale_cols_sqrt<double>(s, c.X.p, n, c.f.p, Y, n, n, nocc));
      DH_TRY(pack_product<double>(c, n, nocc, Y, Y, c.P.p));
    } else
      DH_TRY(pack_product<float>(c, n, nocc, X32, X32, P32, c.f32.p));
    DH_HIP(hipEventRecord(ev[4], s));
    if (c.xf64)
      DH_TRY(launch_pack<double>(s, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p, c.P.p, n,
                                 c.rho.p));
    else
      DH_TRY(launch_pack<float>(s, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p, P32, n, c.rho.p));
    c.xFull = seedAll && (refined || !mixed);
    if (c.xFull && !mixed) {
      // the plain FP64 seed: its vectors transposed for the tail's block-sparse products
      DH_TRY(c.tXt[c.tCur].alloc((size_t)n * n));
      DH_TRY((transpose2d<double, double>(s, c.X.p, n, c.tXt[c.tCur].p, n, n, n)));
    }
    }
  }
  DH_TRY(launch_mulliken_rows(s, c.dblk.rows, c.dblk.norb.p, c.dblk.atomStart.p, c.dblk.start.p,
                              c.ov.p, c.h0.p, c.rho.p, c.qorb.p, c.eorb.p));
  DH_HIP(hipMemcpyAsync(qorbH.data(), c.qorb.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(eorbH.data(), c.eorb.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipEventRecord(ev[5], s));
  DH_HIP(hipStreamSynchronize(s));
  float ms;
  DH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  c.times[kTHam] += ms * 1e-3;
  DH_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
  c.times[kTBack] += ms * 1e-3;
  if (refineTimed) {
    DH_HIP(hipEventElapsedTime(&ms, ev[6], ev[7]));
    c.times[kTRefine] += ms * 1e-3;
    if (refined || tail) c.times[kTBack] -= ms * 1e-3;   // inside the back-transform window
  }
  DH_HIP(hipEventElapsedTime(&ms, ev[3], ev[4]));
  c.times[kTDM] += ms * 1e-3;
  DH_HIP(hipEventElapsedTime(&ms, ev[4], ev[5]));
  c.times[kTPackMull] += ms * 1e-3;
  c.times[kTReduce] += evt[0];
  c.times[kTTridiag] += evt[1];
  c.times[kTDC] += evt[2];
  const double th2 = now();
  c.qorbLast = qorbH;
  std::vector<double> qout(nat, 0.0);
  st.eNonScc = 0.0;
  for (int a = 0; a < nat; ++a)
    for (long oi = c.atomStart[a]; oi < c.atomStart[a + 1]; ++oi) qout[a] += qorbH[oi];
  for (int i = 0; i < n; ++i) st.eNonScc += eorbH[i];
  std::vector<double> dqo(nat), diff(nat);
  double err = 0.0;
  for (int a = 0; a < nat; ++a) {
    dqo[a] = qout[a] - c.q0atom[a];
    diff[a] = qout[a] - st.qin[a];
    err = std::max(err, std::fabs(diff[a]));
  }
  c.times[kTHost] += now() - th2;
  if (tailUsed && err > 4.0 * prevErrIn && err > o.tail_below) {
    // the tail's eigenpairs did not follow this change of H (many near-degenerate levels with
    // different fillings, mixed beyond the first-order range): the charges are not used, the
    // cycle goes on without the tail, and this iteration is repeated with an eigensolve
    c.tailOff = true;
    st.it -= 1;
    return ctx_scc_step(c);
  }
  DH_TRY(gamma_times(c, dqo, &vout));
  st.eScc = 0.0;
  for (int a = 0; a < nat; ++a) st.eScc += 0.5 * vout[a] * dqo[a];
  c.qout = qout;
  c.dqout = dqo;
  st.err = err;
  st.prevErr = err;
  st.eElecPrev = st.it > 1 ? st.eElec : std::numeric_limits<double>::quiet_NaN();
  st.eElec = st.eNonScc + st.eScc;
  st.qinSum = 0.0;
  for (int a = 0; a < nat; ++a) st.qinSum += st.qin[a];
  st.conv = err < o.scc_tol && st.it >= o.min_scc_iter;
  const double th3 = now();
  // the iteration that seeds the tail may restart the Broyden history (the earlier FP32
  // iterations' output charges carry the FP32 noise that the tail's converged charges do not)
  if (g_tailMixReset && seedAll && !tail && c.xFull) st.mixer.reset(c.nat, c.opt, o.max_scc_iter + 1);
  if (!st.conv && st.it != o.max_scc_iter) st.mixer.mix(st.qin, diff);
  c.times[kTHost] += now() - th3;
  c.times[kTIter] = now() - tIt;
  c.sccDone = true;   // a density exists (gradients need convergence for physical meaning)
  return 0;
}

static int ctx_scc_end(Context& c, double* energies, double* charges, int* niter, int* converged) {
  DH_CHECK(c.st.it > 0, "scc_end: no SCC iteration was run");
  hipStream_t s = c.s;
  const int nat = c.nat;
  SccState& st = c.st;
  DH_TRY(fill<double>(s, c.eatom.p, nat, 0.0));
  DH_TRY(fill<double>(s, c.grad.p, 3 * (size_t)nat, 0.0));
  DH_TRY(launch_repulsive(s, c.nbdev, c.coordsAll.p, c.sk.dreps.p, c.sk.nsp, c.eatom.p, c.grad.p));
  std::vector<double> ea(nat);
  DH_HIP(hipMemcpyAsync(ea.data(), c.eatom.p, nat * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  double erep = 0.0;
  for (double x : ea) erep += x;
  const double etot = st.eNonScc + st.eScc + erep;
  c.energies[0] = etot - st.fr.ts;
  c.energies[1] = etot;
  c.energies[2] = st.eNonScc;
  c.energies[3] = st.eScc;
  c.energies[4] = erep;
  c.energies[5] = st.fr.ts;
  c.energies[6] = st.fr.eband;
  c.energies[7] = st.fr.ef;
  if (energies) std::copy(c.energies, c.energies + 8, energies);
  if (charges)
    for (int a = 0; a < nat; ++a) charges[a] = c.q0atom[a] - c.qout[a];
  if (niter) *niter = st.it;
  if (converged) *converged = st.conv ? 1 : 0;
  st.active = false;
  ++c.ncalls;
  return 0;
}

// out[0..5]: iteration number, charge error, electronic energy of the last iteration and of
// the one before (NaN after the first), input electrons of the last iteration, pairs outside
// the first-order range in the last Ogita-Aishima step of the sparse refinement
static int ctx_scc_state(Context& c, double* out, int n) {
  const SccState& st = c.st;
  const double v[8] = {(double)st.it, st.err, st.eElec, st.eElecPrev, st.qinSum,
                       (double)c.lastClipped, (double)c.lastFar, c.tailOff ? 1.0 : 0.0};
  for (int k = 0; k < std::min(n, 8); ++k) out[k] = v[k];
  return 0;
}

static int ctx_scc(Context& c, double* energies, double* charges, int* niter, int* converged) {
  DH_TRY(ctx_scc_begin(c));
  while (!c.st.conv && c.st.it < c.opt.max_scc_iter) DH_TRY(ctx_scc_step(c));
  return ctx_scc_end(c, energies, charges, niter, converged);
}

static int ctx_gradients(Context& c, double* grad) {
  DH_CHECK(c.sccDone, "gradients: run the SCC cycle first");
  hipStream_t s = c.s;
  const int nat = c.nat, n = c.norbTot;
  const double t0 = now();
  // energy-weighted density matrix
  std::vector<double> fe(n, 0.0);
  for (int i = 0; i < c.nocc && i < (int)c.eigv.size(); ++i) fe[i] = c.fill[i] * c.eigv[i];
  if (c.nk > 0) {
    std::vector<double> fk(2 * (size_t)n);
    DH_TRY(fill<double>(s, c.erho.p, c.nnz, 0.0));
    for (int k = 0; k < c.nk; ++k) {
      const int ndk = kpoint_dim(&c.kpts[3 * k], n), stride = ndk / n;
      for (int j = 0; j < n; ++j)
        for (int t = 0; t < stride; ++t)
          fk[stride * j + t] = c.fillK[k][j] * c.eigK[k][stride * j];
      DH_HIP(hipMemcpyAsync(c.f.p, fk.data(), ndk * sizeof(double), hipMemcpyHostToDevice, s));
      // W = (X diag(f e)) X^T: the scaled columns staged in the eigensolver's scratch matrix
      const int kk = stride * c.noccK[k];
      DH_TRY(scale_cols<double>(s, c.XK[k].p, ndk, c.f.p, c.eig.T.p, ndk, ndk, kk));
      DH_TRY(gemm<double>(s, false, true, ndk, ndk, kk, 1.0, c.eig.T.p, ndk, c.XK[k].p, ndk,
                          0.0, c.P.p, ndk, true));
      DH_TRY(launch_pack_kpt(s, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p, c.blkT.p,
                             &c.kpts[3 * k], c.kw[k], c.P.p, ndk, n, c.erho.p));
      DH_HIP(hipStreamSynchronize(s));
    }
  } else if (!c.xf64) {
    std::vector<float> fe32(fe.begin(), fe.end());
    float* X32 = reinterpret_cast<float*>(c.X.p);
    float* P32 = reinterpret_cast<float*>(c.P.p);
    DH_HIP(hipMemcpyAsync(c.f32.p, fe32.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    DH_TRY(pack_product<float>(c, n, c.nocc, X32, X32, P32, c.f32.p));
    DH_TRY(launch_pack<float>(s, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p, P32, n, c.erho.p));
    DH_HIP(hipStreamSynchronize(s));   // fe32 leaves scope
  } else {
    DH_HIP(hipMemcpyAsync(c.feps.p, fe.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    // W = (X diag(f e)) X^T: the scaled columns staged in the eigensolver's scratch matrix,
    // free once the eigenvectors are formed (the weights change sign, so no SYRK of square roots)
    DH_TRY(scale_cols<double>(s, c.X.p, n, c.feps.p, c.eig.T.p, n, n, c.nocc));
    DH_TRY(pack_product<double>(c, n, c.nocc, c.eig.T.p, c.X.p, c.P.p));
    DH_TRY(launch_pack<double>(s, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p, c.P.p, n,
                               c.erho.p));
  }
  // output potential per orbital
  std::vector<double> vout;
  DH_TRY(gamma_times(c, c.dqout, &vout));
  // repulsive gradient (energy recomputed, discarded)
  DH_TRY(fill<double>(s, c.eatom.p, nat, 0.0));
  DH_TRY(fill<double>(s, c.grad.p, 3 * (size_t)nat, 0.0));
  DH_TRY(launch_repulsive(s, c.nbdev, c.coordsAll.p, c.sk.dreps.p, c.sk.nsp, c.eatom.p, c.grad.p));
  DH_TRY(launch_pair_forces(s, c.sk.dev, c.dblk.pr, c.dblk.norb.p, c.dblk.atomStart.p,
                            c.coordsAll.p, c.rho.p, c.erho.p, c.vorb.p, c.opt.force_mode,
                            c.opt.fd_delta, c.grad.p));
  const int ng = (int)c.ew.w.size();
  if (c.periodic && ng > 0) {
    // C_G = sum_j dq_j cos(G r_j), S_G = sum_j dq_j sin(G r_j)
    DH_TRY(gemm<double>(s, true, false, 2 * ng, 1, nat, 1.0, c.ewCS.p, nat, c.dq.p, nat, 0.0,
                        c.ewCSq.p, 2 * ng));
  }
  DH_TRY(launch_gamma_grad(s, c.nbdev, c.coordsAll.p, c.sk.dhubbu.p, c.sk.dgcut.p, c.sk.nsp, nat,
                           c.periodic, (int)c.ew.tr.size() / 3, c.ewTr.p, c.ew.alpha, c.ew.maxr,
                           ng, c.ewG.p, c.ewW.p, c.ewCSq.p, c.dq.p, c.grad.p));
  DH_HIP(hipMemcpyAsync(grad, c.grad.p, 3 * (size_t)nat * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  c.times[kTForces] = now() - t0;
  return 0;
}

// ------------------------------------------------------------------------------------------
// Molecule batch: many independent non-periodic molecules solved together. The union of the
// molecules is one block-diagonal system; every molecule keeps its own electron count, Fermi
// level, Broyden mixer and convergence state, as in a loop over single-molecule calculations.
// ------------------------------------------------------------------------------------------
// Pinned host buffer (grown on demand, freed with the owner) for transfers repeated every
// iteration; pageable copies are staged by the runtime at a fraction of the bandwidth.
struct PinnedHost {
  double* p = nullptr;
  size_t n = 0;
  PinnedHost() = default;
  PinnedHost(const PinnedHost&) = delete;
  PinnedHost& operator=(const PinnedHost&) = delete;
  int ensure(size_t count) {
    if (p && count <= n) return 0;
    release();
    DH_HIP(hipHostMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(double),
                         hipHostMallocDefault));
    n = count;
    return 0;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  ~PinnedHost() { release(); }
};

struct BatchContext {
  dftbhip_options opt{};
  SkTables sk;
  hipStream_t s = nullptr;
  int nmol = 0, nat = 0, norbTot = 0;
  std::vector<int> atom0, atomMol, species, norbAtom, nocc, iters, conv;
  std::vector<long> orb0, dOff, gOff, atomStart;
  std::vector<double> coords, q0atom, nel, eigv, fill, qin, qout, dqout;
  std::vector<Broyden> mixers;
  std::vector<FillResult> fr;
  std::vector<double> eNonScc, eScc, erep, err;
  HostBlocks blk, nbp;
  DevBlocks dblk, dnb;
  NeighDev nbdev{};
  long nnz = 0, denseSize = 0, gammaSize = 0;
  DevBuf<double> coordsD, h0, ov, hs, rho, erho, vorb, vatom, dq, qorb, eorb, grad, eatom, gamma;
  DevBuf<double> Sd, Hd, Cv, P, work, w, f;
  DevBuf<int> atom0D, atomMolD, infoD, noccD, activeD, orderD;
  DevBuf<long> orb0D, dOffD, gOffD;
  BatchLayout L{};
  int it = 0, nactive = 0;
  bool sccDone = false;
  // per-iteration transfers: eigenvalues and orbital charges and energies down, fillings up
  PinnedHost pEig, pFill, pQorb, pEorb;
  // streams of the eigensolver's smaller size classes (default priority) and the events that
  // fork them from s and join them back
  static constexpr int kSx = kBatchClasses - 1;
  hipStream_t sx[kSx] = {};
  hipEvent_t fork = nullptr, join[kSx] = {};
  // SCC iteration with the Fermi levels, energies, convergence and Broyden mixing on the device
  // (sccMode 1: launched one by one; 2: the warm iterations replay one captured HIP graph) or
  // on host threads (sccMode 0)
  int sccMode = 2;
  DevBuf<double> q0D, qinD, qoutD, tmpD, nelD, errD, efD, ebD, tsD, e0D, enD, esD;
  DevBuf<double> wwD, ccD, aaD, BD, dFD, uuD, qlastD, dlastD;
  DevBuf<int> convD, itersD, mixItD, cntD;
  BatchSccDev dev{};
  PinnedHost pCnt;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }
  ~BatchContext() {
    drop_graph();
    for (hipStream_t x : sx)
      if (x) (void)hipStreamDestroy(x);
    for (hipEvent_t e : join)
      if (e) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
  }
};

static int batch_geometry(BatchContext& b, int nmol, const int* natMol, const double* coords,
                          const int* species) {
  hipStream_t s = b.s;
  b.drop_graph();
  b.nmol = nmol;
  b.atom0.assign(nmol + 1, 0);
  for (int m = 0; m < nmol; ++m) {
    DH_CHECK(natMol[m] >= 1, "batch: a molecule without atoms");
    b.atom0[m + 1] = b.atom0[m] + natMol[m];
  }
  const int nat = b.atom0[nmol];
  b.nat = nat;
  b.coords.assign(coords, coords + 3 * (size_t)nat);
  b.species.assign(species, species + nat);
  b.atomMol.resize(nat);
  b.norbAtom.resize(nat);
  b.atomStart.assign(nat + 1, 0);
  b.q0atom.resize(nat);
  for (int m = 0; m < nmol; ++m)
    for (int a = b.atom0[m]; a < b.atom0[m + 1]; ++a) b.atomMol[a] = m;
  for (int a = 0; a < nat; ++a) {
    DH_CHECK(species[a] >= 0 && species[a] < b.sk.nsp, "batch: species out of range");
    b.norbAtom[a] = b.sk.norb[species[a]];
    b.atomStart[a + 1] = b.atomStart[a] + b.norbAtom[a];
    b.q0atom[a] = b.sk.q0[species[a]];
  }
  b.norbTot = (int)b.atomStart[nat];
  b.orb0.assign(nmol + 1, 0);
  b.dOff.assign(nmol + 1, 0);
  b.gOff.assign(nmol + 1, 0);
  b.nel.assign(nmol, 0.0);
  for (int m = 0; m < nmol; ++m) {
    b.orb0[m + 1] = b.atomStart[b.atom0[m + 1]];
    const long n = b.orb0[m + 1] - b.orb0[m];
    DH_CHECK(n <= kBatchMaxOrb, "batch: a molecule has more than 512 orbitals");
    b.dOff[m + 1] = b.dOff[m] + n * n;
    const long na = natMol[m];
    b.gOff[m + 1] = b.gOff[m] + na * na;
    for (int a = b.atom0[m]; a < b.atom0[m + 1]; ++a) b.nel[m] += b.q0atom[a];
  }
  b.denseSize = b.dOff[nmol];
  b.gammaSize = b.gOff[nmol];
  // neighbour lists per molecule, concatenated (image index = atom index: no periodicity)
  double cutoff = b.sk.skcut;
  for (double r : b.sk.repcut) cutoff = std::max(cutoff, r);
  for (double g : b.sk.gcut) cutoff = std::max(cutoff, g);
  b.blk = HostBlocks();
  b.nbp = HostBlocks();
  const double sk2 = b.sk.skcut * b.sk.skcut;
  long st = 0;
  for (int m = 0; m < nmol; ++m) {
    const int a0 = b.atom0[m];
    NeighbourList nl;
    DH_TRY(build_neighbours(natMol[m], coords + 3 * (size_t)a0, nullptr, cutoff, nl));
    for (int a = 0; a < natMol[m]; ++a) {
      const auto& nb = nl.nb[a];
      const auto& d2 = nl.d2[a];
      for (size_t k = 0; k < nb.size(); ++k) {
        const int ga = a0 + a, gj = a0 + nl.img2cent[nb[k]];
        b.nbp.at1.push_back(ga);
        b.nbp.img2.push_back(gj);
        b.nbp.at2.push_back(gj);
        b.nbp.start.push_back(0);
        if (d2[k] <= sk2) {
          b.blk.at1.push_back(ga);
          b.blk.img2.push_back(gj);
          b.blk.at2.push_back(gj);
          b.blk.start.push_back(st);
          st += (long)b.norbAtom[ga] * b.norbAtom[gj];
        }
      }
    }
  }
  b.nnz = st;
  DH_TRY(b.dblk.upload(b.blk, b.species, b.norbAtom, b.atomStart, s));
  DH_TRY(b.dblk.build_rows(b.nat, s));   // row view of the blocks (Mulliken)
  DH_TRY(b.dnb.upload(b.nbp, b.species, b.norbAtom, b.atomStart, s));
  b.nbdev.n = b.dnb.pr.nblk;
  b.nbdev.at1 = b.dnb.pr.at1;
  b.nbdev.img2 = b.dnb.pr.img2;
  b.nbdev.at2 = b.dnb.pr.at2;
  b.nbdev.species = b.dnb.pr.species;
  DH_TRY(b.coordsD.upload(coords, 3 * (size_t)nat, s));
  DH_TRY(b.atom0D.upload(b.atom0.data(), b.atom0.size(), s));
  DH_TRY(b.atomMolD.upload(b.atomMol.data(), b.atomMol.size(), s));
  DH_TRY(b.orb0D.upload(b.orb0.data(), b.orb0.size(), s));
  DH_TRY(b.dOffD.upload(b.dOff.data(), b.dOff.size(), s));
  DH_TRY(b.gOffD.upload(b.gOff.data(), b.gOff.size(), s));
  std::vector<int> order(nmol);
  for (int m = 0; m < nmol; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return b.orb0[x + 1] - b.orb0[x] > b.orb0[y + 1] - b.orb0[y];
  });
  DH_TRY(b.orderD.upload(order.data(), order.size(), s));
  b.L.nmol = nmol;
  b.L.order = b.orderD.p;
  for (int& c : b.L.ncls) c = 0;
  for (int m : order) {
    const long nm = b.orb0[m + 1] - b.orb0[m];
    ++b.L.ncls[nm > kBatchLdsMid ? 0 : nm > kBatchLdsC ? 1 : nm > kBatchLdsSmall ? 2 : 3];
  }
  b.L.atom0 = b.atom0D.p;
  b.L.orb0 = b.orb0D.p;
  b.L.dOff = b.dOffD.p;
  b.L.gOff = b.gOffD.p;
  b.L.atomMol = b.atomMolD.p;
  const int n = b.norbTot;
  for (DevBuf<double>* x : {&b.h0, &b.ov, &b.hs, &b.rho, &b.erho}) DH_TRY(x->alloc(std::max(b.nnz, 1L)));
  for (DevBuf<double>* x : {&b.vorb, &b.qorb, &b.eorb, &b.w, &b.f}) DH_TRY(x->alloc(n));
  for (DevBuf<double>* x : {&b.vatom, &b.dq, &b.eatom}) DH_TRY(x->alloc(nat));
  DH_TRY(b.grad.alloc(3 * (size_t)nat));
  for (DevBuf<double>* x : {&b.Sd, &b.Hd, &b.Cv, &b.P}) DH_TRY(x->alloc(b.denseSize));
  DH_TRY(b.work.alloc(3 * b.denseSize));
  DH_TRY(b.gamma.alloc(b.gammaSize));
  DH_TRY(b.infoD.alloc(nmol));
  DH_TRY(b.noccD.alloc(nmol));
  DH_TRY(b.activeD.alloc(nmol));
  DH_TRY(launch_build_h0s(s, b.sk.dev, b.dblk.pr, b.coordsD.p, b.h0.p, b.ov.p));
  DH_TRY(launch_batch_unpack(s, b.dblk.pr, b.dblk.norb.p, b.dblk.atomStart.p, b.L, b.ov.p,
                             b.Sd.p, b.denseSize));
  DH_TRY(launch_batch_gamma(s, b.L, b.coordsD.p, b.dblk.species.p, b.sk.dhubbu.p, b.sk.dgcut.p,
                            b.sk.nsp, b.gamma.p));
  DH_HIP(hipStreamSynchronize(s));
  b.sccDone = false;
  return 0;
}

// atom potentials of the charges dqa (device dq, vatom, vorb); optionally back to the host
static int batch_potential(BatchContext& b, const std::vector<double>& dqa,
                           std::vector<double>* vout) {
  hipStream_t s = b.s;
  DH_HIP(hipMemcpyAsync(b.dq.p, dqa.data(), b.nat * sizeof(double), hipMemcpyHostToDevice, s));
  DH_TRY(launch_batch_gemv(s, b.L, b.nat, b.gamma.p, b.dq.p, b.vatom.p));
  hipLaunchKernelGGL(expand_kernel, dim3(cdiv(b.nat, 256)), dim3(256), 0, s, b.nat,
                     b.dblk.atomStart.p, b.vatom.p, b.vorb.p);
  DH_LAUNCH_CHECK();
  if (vout) {
    vout->resize(b.nat);
    DH_HIP(hipMemcpyAsync(vout->data(), b.vatom.p, b.nat * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  }
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

// The SCC cycle of every molecule (main.F90:1429-1563 per molecule); molecules that have
// converged keep their input charges, so their state stays that of their last iteration.
// The SCC cycle of every molecule (main.F90:1429-1563 per molecule); molecules that have
// converged keep their input charges, so their state stays that of their last iteration.
static int batch_dev_begin(BatchContext& b);
static int batch_dev_step(BatchContext& b);
static int g_batchSccMode();

static int batch_scc_begin(BatchContext& b) {
  const dftbhip_options& o = b.opt;
  const int nmol = b.nmol, nat = b.nat, n = b.norbTot;
  b.sccMode = g_batchSccMode();
  if (b.sccMode > 0) {
    b.it = 0;
    b.nactive = nmol;
    b.sccDone = false;
    return batch_dev_begin(b);
  }
  b.qin = b.q0atom;
  b.mixers.assign(nmol, Broyden());
  for (int m = 0; m < nmol; ++m)
    b.mixers[m].reset(b.atom0[m + 1] - b.atom0[m], o, o.max_scc_iter + 1);
  b.iters.assign(nmol, 0);
  b.conv.assign(nmol, 0);
  b.err.assign(nmol, 0.0);
  b.fr.assign(nmol, FillResult{});
  b.eNonScc.assign(nmol, 0.0);
  b.eScc.assign(nmol, 0.0);
  b.nocc.assign(nmol, 0);
  b.fill.assign(n, 0.0);
  b.qout.assign(nat, 0.0);
  b.dqout.assign(nat, 0.0);
  b.it = 0;
  b.nactive = nmol;
  b.sccDone = false;
  return 0;
}

// f(i) for i < n on up to 16 host threads (serially below minPar items); the first error of a
// thread is raised again on the calling thread (error messages are per thread)
static int parallel_for(int n, int minPar, const std::function<int(int)>& f) {
  const int T = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (n < minPar || T == 1) {
    for (int i = 0; i < n; ++i) DH_TRY(f(i));
    return 0;
  }
  std::vector<int> rc(T, 0);
  std::vector<std::string> msg(T);
  std::vector<std::thread> th;
  th.reserve(T);
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t]() {
      for (int i = t; i < n; i += T) {
        const int r = f(i);
        if (r != 0) {
          rc[t] = r;
          msg[t] = last_error();
          return;
        }
      }
    });
  for (auto& x : th) x.join();
  for (int t = 0; t < T; ++t)
    if (rc[t] != 0) {
      set_error(msg[t]);
      return rc[t];
    }
  return 0;
}

static int batch_scc_step(BatchContext& b) {
  if (b.sccMode > 0) return batch_dev_step(b);
  hipStream_t s = b.s;
  const dftbhip_options& o = b.opt;
  const int nmol = b.nmol, nat = b.nat, n = b.norbTot;
  DH_CHECK(b.nactive > 0 && b.it < o.max_scc_iter, "batch scc_step: nothing left to iterate");
  const int it = ++b.it;
  int& nactive = b.nactive;
  std::vector<double> dqa(nat), vout;
  std::vector<int> info(nmol);
  DH_TRY(b.pEig.ensure(n));
  DH_TRY(b.pFill.ensure(n));
  DH_TRY(b.pQorb.ensure(n));
  DH_TRY(b.pEorb.ensure(n));
  double* qorbH = b.pQorb.p;
  double* eorbH = b.pEorb.p;
  for (int a = 0; a < nat; ++a) dqa[a] = b.qin[a] - b.q0atom[a];
  DH_TRY(batch_potential(b, dqa, nullptr));
  DH_TRY(launch_shift(s, b.dblk.pr, b.dblk.norb.p, b.dblk.atomStart.p, b.h0.p, b.ov.p,
                      b.vorb.p, b.hs.p));
  DH_TRY(launch_batch_unpack(s, b.dblk.pr, b.dblk.norb.p, b.dblk.atomStart.p, b.L, b.hs.p,
                             b.Hd.p, b.denseSize));
  std::vector<int> active(nmol);
  for (int m = 0; m < nmol; ++m) active[m] = b.conv[m] ? 0 : 1;
  DH_HIP(hipMemcpyAsync(b.activeD.p, active.data(), nmol * sizeof(int), hipMemcpyHostToDevice, s));
  // the size classes of the eigensolver run side by side
  DH_HIP(hipEventRecord(b.fork, s));
  for (int q = 0; q < BatchContext::kSx; ++q) DH_HIP(hipStreamWaitEvent(b.sx[q], b.fork, 0));
  hipStream_t st[kBatchClasses] = {s};
  for (int q = 0; q < BatchContext::kSx; ++q) st[q + 1] = b.sx[q];
  DH_TRY(launch_batch_sygv(st, b.L, b.Sd.p, b.Hd.p, b.work.p, b.w.p, b.Cv.p, b.infoD.p,
                           b.activeD.p, it > 1));
  for (int q = 0; q < BatchContext::kSx; ++q) {
    DH_HIP(hipEventRecord(b.join[q], b.sx[q]));
    DH_HIP(hipStreamWaitEvent(s, b.join[q], 0));
  }
  b.eigv.resize(n);
  DH_HIP(hipMemcpyAsync(b.pEig.p, b.w.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(info.data(), b.infoD.p, nmol * sizeof(int), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  std::copy(b.pEig.p, b.pEig.p + n, b.eigv.begin());
  for (int m = 0; m < nmol; ++m)
    if (info[m] != 0) {
      set_error(info[m] == kBatchJacobiNoConv
                    ? "batch: Jacobi eigensolver of molecule " + std::to_string(m) +
                          " did not converge within its sweep limit"
                : info[m] < 0 ? "batch: molecule " + std::to_string(m) + " of " +
                                    std::to_string(-info[m]) + " orbitals in a smaller size class"
                              : "batch: overlap of molecule " + std::to_string(m) +
                                    " not positive definite (minor " + std::to_string(info[m]) +
                                    ")");
      return kErrState;
    }
  // Fermi levels per molecule, molecules on host threads
  DH_TRY(parallel_for(nmol, 64, [&](int m) -> int {
    if (b.conv[m]) return 0;   // fillings of converged molecules stay from their last pass
    const long o0 = b.orb0[m], nm = b.orb0[m + 1] - o0;
    std::vector<double> e(b.eigv.begin() + o0, b.eigv.begin() + o0 + nm), fm;
    DH_TRY(fermi_filling(e, b.nel[m], o.kT, fm, b.fr[m]));
    std::copy(fm.begin(), fm.end(), b.fill.begin() + o0);
    b.nocc[m] = occupied_levels(fm);
    return 0;
  }));
  std::copy(b.fill.begin(), b.fill.end(), b.pFill.p);
  DH_HIP(hipMemcpyAsync(b.f.p, b.pFill.p, n * sizeof(double), hipMemcpyHostToDevice, s));
  DH_HIP(hipMemcpyAsync(b.noccD.p, b.nocc.data(), nmol * sizeof(int), hipMemcpyHostToDevice, s));
  DH_TRY(launch_batch_density(s, b.L, b.Cv.p, b.f.p, b.noccD.p, b.P.p, b.activeD.p));
  DH_TRY(launch_batch_pack(s, b.dblk.pr, b.dblk.norb.p, b.dblk.atomStart.p, b.L, b.P.p,
                           b.rho.p));
  DH_TRY(launch_mulliken_rows(s, b.dblk.rows, b.dblk.norb.p, b.dblk.atomStart.p, b.dblk.start.p,
                              b.ov.p, b.h0.p, b.rho.p, b.qorb.p, b.eorb.p));
  DH_HIP(hipMemcpyAsync(qorbH, b.qorb.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(eorbH, b.eorb.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  std::vector<double> qo(nat, 0.0);
  for (int a = 0; a < nat; ++a)
    for (long oi = b.atomStart[a]; oi < b.atomStart[a + 1]; ++oi) qo[a] += qorbH[oi];
  std::vector<double> dqo(nat);
  for (int a = 0; a < nat; ++a) dqo[a] = qo[a] - b.q0atom[a];
  DH_TRY(batch_potential(b, dqo, &vout));
  // energies, convergence and Broyden mixing per molecule, molecules on host threads
  DH_TRY(parallel_for(nmol, 64, [&](int m) -> int {
    if (b.conv[m]) return 0;
    b.iters[m] = it;
    double en = 0.0, es = 0.0, er = 0.0;
    for (long oi = b.orb0[m]; oi < b.orb0[m + 1]; ++oi) en += eorbH[oi];
    std::vector<double> q(b.qin.begin() + b.atom0[m], b.qin.begin() + b.atom0[m + 1]);
    std::vector<double> diff(q.size());
    for (int a = b.atom0[m]; a < b.atom0[m + 1]; ++a) {
      es += 0.5 * vout[a] * dqo[a];
      diff[a - b.atom0[m]] = qo[a] - b.qin[a];
      er = std::max(er, std::fabs(diff[a - b.atom0[m]]));
      b.qout[a] = qo[a];
      b.dqout[a] = dqo[a];
    }
    b.eNonScc[m] = en;
    b.eScc[m] = es;
    b.err[m] = er;
    if (er < o.scc_tol && it >= o.min_scc_iter) {
      b.conv[m] = 1;
    } else if (it != o.max_scc_iter) {
      b.mixers[m].mix(q, diff);
      std::copy(q.begin(), q.end(), b.qin.begin() + b.atom0[m]);
    }
    return 0;
  }));
  nactive = 0;
  for (int m = 0; m < nmol; ++m) nactive += b.conv[m] ? 0 : 1;
  return 0;
}

// ---- the SCC cycle on the device ---------------------------------------------------------
// molecules selected with set_tuning parameter 4: 0 host threads, 1 device kernels, 2 device
// kernels with the warm iterations replayed from one captured graph (the default)
static int g_batchScc = 2;
static int g_batchSccMode() { return g_batchScc; }

static int batch_dev_begin(BatchContext& b) {
  hipStream_t s = b.s;
  const dftbhip_options& o = b.opt;
  const int nmol = b.nmol, nat = b.nat;
  const int H = std::max(o.max_scc_iter, 1);
  b.drop_graph();
  DH_TRY(b.q0D.upload(b.q0atom.data(), nat, s));
  DH_TRY(b.nelD.upload(b.nel.data(), nmol, s));
  for (DevBuf<double>* x : {&b.qinD, &b.qoutD, &b.tmpD, &b.qlastD, &b.dlastD}) DH_TRY(x->alloc(nat));
  for (DevBuf<double>* x : {&b.errD, &b.efD, &b.ebD, &b.tsD, &b.e0D, &b.enD, &b.esD})
    DH_TRY(x->alloc(nmol));
  DH_TRY(b.wwD.alloc((size_t)nmol * H));
  DH_TRY(b.ccD.alloc((size_t)nmol * H));
  DH_TRY(b.aaD.alloc((size_t)nmol * H * H));
  DH_TRY(b.BD.alloc((size_t)nmol * H * H));
  DH_TRY(b.dFD.alloc((size_t)H * nat));
  DH_TRY(b.uuD.alloc((size_t)H * nat));
  for (DevBuf<int>* x : {&b.convD, &b.itersD, &b.mixItD}) DH_TRY(x->alloc(nmol));
  DH_TRY(b.cntD.alloc(4));
  DH_TRY(b.pCnt.ensure(2));
  BatchSccDev& d = b.dev;
  d.nmol = nmol;
  d.nat = nat;
  d.maxHist = H;
  d.atom0 = b.atom0D.p;
  d.orb0 = b.orb0D.p;
  d.q0 = b.q0D.p;
  d.nel = b.nelD.p;
  d.qin = b.qinD.p;
  d.qout = b.qoutD.p;
  d.dq = b.dq.p;
  d.tmp = b.tmpD.p;
  d.conv = b.convD.p;
  d.active = b.activeD.p;
  d.iters = b.itersD.p;
  d.nocc = b.noccD.p;
  d.mixIt = b.mixItD.p;
  d.counters = b.cntD.p;
  d.err = b.errD.p;
  d.ef = b.efD.p;
  d.eband = b.ebD.p;
  d.ts = b.tsD.p;
  d.e0 = b.e0D.p;
  d.eNonScc = b.enD.p;
  d.eScc = b.esD.p;
  d.ww = b.wwD.p;
  d.aa = b.aaD.p;
  d.B = b.BD.p;
  d.cc = b.ccD.p;
  d.dF = b.dFD.p;
  d.uu = b.uuD.p;
  d.qlast = b.qlastD.p;
  d.dlast = b.dlastD.p;
  d.kT = o.kT;
  d.sccTol = o.scc_tol;
  d.alpha = o.mix_param;
  d.omega0 = o.omega0;
  d.minw = o.min_weight;
  d.maxw = o.max_weight;
  d.wfac = o.weight_fac;
  d.minSccIter = o.min_scc_iter;
  d.maxSccIter = o.max_scc_iter;
  DH_TRY(launch_bscc_reset(s, d));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

// one SCC iteration of every active molecule as launches on b.s (and the size classes' streams,
// forked from and joined to it), ending with the copy of the counters to pinned host memory;
// nothing here waits on the host, so the sequence can be captured
static int batch_dev_enqueue(BatchContext& b, bool warm) {
  hipStream_t s = b.s;
  const BatchSccDev& d = b.dev;
  const int n = b.norbTot;
  DH_TRY(launch_bscc_begin(s, d));
  DH_TRY(launch_batch_gemv(s, b.L, b.nat, b.gamma.p, b.dq.p, b.vatom.p));
  hipLaunchKernelGGL(expand_kernel, dim3(cdiv(b.nat, 256)), dim3(256), 0, s, b.nat,
                     b.dblk.atomStart.p, b.vatom.p, b.vorb.p);
  DH_LAUNCH_CHECK();
  DH_TRY(launch_shift(s, b.dblk.pr, b.dblk.norb.p, b.dblk.atomStart.p, b.h0.p, b.ov.p,
                      b.vorb.p, b.hs.p));
  DH_TRY(launch_batch_unpack(s, b.dblk.pr, b.dblk.norb.p, b.dblk.atomStart.p, b.L, b.hs.p,
                             b.Hd.p, b.denseSize));
  DH_HIP(hipEventRecord(b.fork, s));
  for (int q = 0; q < BatchContext::kSx; ++q) DH_HIP(hipStreamWaitEvent(b.sx[q], b.fork, 0));
  hipStream_t st[kBatchClasses] = {s};
  for (int q = 0; q < BatchContext::kSx; ++q) st[q + 1] = b.sx[q];
  DH_TRY(launch_batch_sygv(st, b.L, b.Sd.p, b.Hd.p, b.work.p, b.w.p, b.Cv.p, b.infoD.p,
                           b.activeD.p, warm));
  for (int q = 0; q < BatchContext::kSx; ++q) {
    DH_HIP(hipEventRecord(b.join[q], b.sx[q]));
    DH_HIP(hipStreamWaitEvent(s, b.join[q], 0));
  }
  DH_TRY(launch_bscc_info(s, d, b.infoD.p));
  DH_TRY(launch_bscc_fermi(s, d, b.w.p, b.f.p));
  DH_TRY(launch_batch_density(s, b.L, b.Cv.p, b.f.p, b.noccD.p, b.P.p, b.activeD.p));
  DH_TRY(launch_batch_pack(s, b.dblk.pr, b.dblk.norb.p, b.dblk.atomStart.p, b.L, b.P.p,
                           b.rho.p));
  DH_TRY(launch_mulliken_rows(s, b.dblk.rows, b.dblk.norb.p, b.dblk.atomStart.p, b.dblk.start.p,
                              b.ov.p, b.h0.p, b.rho.p, b.qorb.p, b.eorb.p));
  DH_TRY(launch_bscc_outputs(s, d, b.dblk.atomStart.p, b.atomMolD.p, b.qorb.p, b.eorb.p));
  DH_TRY(launch_batch_gemv(s, b.L, b.nat, b.gamma.p, b.dq.p, b.vatom.p));
  DH_TRY(launch_bscc_mix(s, d, b.vatom.p));
  DH_HIP(hipMemcpyAsync(b.pCnt.p, b.cntD.p + 1, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
  return 0;
}

static int batch_dev_step(BatchContext& b) {
  hipStream_t s = b.s;
  const dftbhip_options& o = b.opt;
  DH_CHECK(b.nactive > 0 && b.it < o.max_scc_iter, "batch scc_step: nothing left to iterate");
  const int it = ++b.it;
  const bool warm = it > 1;
  if (b.sccMode == 2 && warm) {
    if (!b.gexec) {
      DH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const int rc = batch_dev_enqueue(b, true);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(s, &g);
      if (rc != 0) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      DH_HIP(ec);
      b.graph = g;
      DH_HIP(hipGraphInstantiate(&b.gexec, b.graph, nullptr, nullptr, 0));
    }
    DH_HIP(hipGraphLaunch(b.gexec, s));
  } else {
    DH_TRY(batch_dev_enqueue(b, warm));
  }
  DH_HIP(hipStreamSynchronize(s));
  const int* c = reinterpret_cast<const int*>(b.pCnt.p);
  if (c[1] != 0) {
    const int m = c[2];
    if (c[1] == 1) {
      std::vector<int> info(b.nmol);
      DH_HIP(hipMemcpy(info.data(), b.infoD.p, b.nmol * sizeof(int), hipMemcpyDeviceToHost));
      const int im = (m >= 0 && m < b.nmol) ? info[m] : 0;
      set_error(im == kBatchJacobiNoConv
                    ? "batch: Jacobi eigensolver of molecule " + std::to_string(m) +
                          " did not converge within its sweep limit"
                : im < 0 ? "batch: molecule " + std::to_string(m) + " of " +
                               std::to_string(-im) + " orbitals in a smaller size class"
                         : "batch: overlap of molecule " + std::to_string(m) +
                               " not positive definite (minor " + std::to_string(im) + ")");
    } else if (c[1] == 2) {
      set_error("batch: Fermi level search did not converge for molecule " + std::to_string(m));
    } else {
      set_error("batch: Broyden mixer of molecule " + std::to_string(m) +
                ": Maximal nr. of steps exceeded");
    }
    return kErrState;
  }
  b.nactive = c[0];
  return 0;
}

// the device state the host-side end of the cycle and the gradients read
static int batch_dev_fetch(BatchContext& b) {
  hipStream_t s = b.s;
  const int nmol = b.nmol, nat = b.nat, n = b.norbTot;
  std::vector<double> ef(nmol), eb(nmol), ts(nmol), e0(nmol);
  b.qout.resize(nat);
  b.eNonScc.resize(nmol);
  b.eScc.resize(nmol);
  b.err.resize(nmol);
  b.iters.resize(nmol);
  b.conv.resize(nmol);
  b.nocc.resize(nmol);
  b.fill.resize(n);
  b.eigv.resize(n);
  b.fr.resize(nmol);
  DH_HIP(hipMemcpyAsync(b.qout.data(), b.qoutD.p, nat * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(b.eNonScc.data(), b.enD.p, nmol * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(b.eScc.data(), b.esD.p, nmol * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(b.err.data(), b.errD.p, nmol * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(b.iters.data(), b.itersD.p, nmol * sizeof(int), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(b.conv.data(), b.convD.p, nmol * sizeof(int), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(b.nocc.data(), b.noccD.p, nmol * sizeof(int), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(b.fill.data(), b.f.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(b.eigv.data(), b.w.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(ef.data(), b.efD.p, nmol * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(eb.data(), b.ebD.p, nmol * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(ts.data(), b.tsD.p, nmol * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(e0.data(), b.e0D.p, nmol * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  for (int m = 0; m < nmol; ++m) b.fr[m] = FillResult{ef[m], eb[m], ts[m], e0[m]};
  b.dqout.resize(nat);
  for (int a = 0; a < nat; ++a) b.dqout[a] = b.qout[a] - b.q0atom[a];
  return 0;
}

static int batch_scc_end(BatchContext& b, double* energies, double* charges, int* niter,
                         int* converged) {
  DH_CHECK(b.it > 0, "batch scc_end: no SCC iteration was run");
  if (b.sccMode > 0) DH_TRY(batch_dev_fetch(b));
  hipStream_t s = b.s;
  const int nmol = b.nmol, nat = b.nat;
  // repulsive energies and the final energies per molecule
  DH_TRY(fill<double>(s, b.eatom.p, nat, 0.0));
  DH_TRY(fill<double>(s, b.grad.p, 3 * (size_t)nat, 0.0));
  DH_TRY(launch_repulsive(s, b.nbdev, b.coordsD.p, b.sk.dreps.p, b.sk.nsp, b.eatom.p, b.grad.p));
  std::vector<double> ea(nat);
  DH_HIP(hipMemcpyAsync(ea.data(), b.eatom.p, nat * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  for (int m = 0; m < nmol; ++m) {
    double erep = 0.0;
    for (int a = b.atom0[m]; a < b.atom0[m + 1]; ++a) erep += ea[a];
    const double etot = b.eNonScc[m] + b.eScc[m] + erep;
    double* e = energies ? energies + 8 * (size_t)m : nullptr;
    if (e) {
      e[0] = etot - b.fr[m].ts;
      e[1] = etot;
      e[2] = b.eNonScc[m];
      e[3] = b.eScc[m];
      e[4] = erep;
      e[5] = b.fr[m].ts;
      e[6] = b.fr[m].eband;
      e[7] = b.fr[m].ef;
    }
    if (niter) niter[m] = b.iters[m];
    if (converged) converged[m] = b.conv[m];
  }
  if (charges)
    for (int a = 0; a < nat; ++a) charges[a] = b.q0atom[a] - b.qout[a];
  b.sccDone = true;
  return 0;
}


static int batch_scc(BatchContext& b, double* energies, double* charges, int* niter,
                     int* converged) {
  DH_TRY(batch_scc_begin(b));
  while (b.nactive > 0 && b.it < b.opt.max_scc_iter) DH_TRY(batch_scc_step(b));
  return batch_scc_end(b, energies, charges, niter, converged);
}

static int batch_gradients(BatchContext& b, double* grad) {
  DH_CHECK(b.sccDone, "batch gradients: run the SCC cycle first");
  hipStream_t s = b.s;
  const int nat = b.nat, n = b.norbTot;
  std::vector<double> fe(n, 0.0);
  for (int m = 0; m < b.nmol; ++m)
    for (long k = 0; k < b.nocc[m]; ++k) fe[b.orb0[m] + k] = b.fill[b.orb0[m] + k] * b.eigv[b.orb0[m] + k];
  DH_HIP(hipMemcpyAsync(b.f.p, fe.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
  DH_TRY(launch_batch_density(s, b.L, b.Cv.p, b.f.p, b.noccD.p, b.P.p, nullptr));
  DH_TRY(launch_batch_pack(s, b.dblk.pr, b.dblk.norb.p, b.dblk.atomStart.p, b.L, b.P.p, b.erho.p));
  DH_TRY(batch_potential(b, b.dqout, nullptr));
  DH_TRY(fill<double>(s, b.eatom.p, nat, 0.0));
  DH_TRY(fill<double>(s, b.grad.p, 3 * (size_t)nat, 0.0));
  DH_TRY(launch_repulsive(s, b.nbdev, b.coordsD.p, b.sk.dreps.p, b.sk.nsp, b.eatom.p, b.grad.p));
  DH_TRY(launch_pair_forces(s, b.sk.dev, b.dblk.pr, b.dblk.norb.p, b.dblk.atomStart.p,
                            b.coordsD.p, b.rho.p, b.erho.p, b.vorb.p, b.opt.force_mode,
                            b.opt.fd_delta, b.grad.p));
  DH_TRY(launch_short_gamma_grad(s, b.nbdev, b.coordsD.p, b.sk.dhubbu.p, b.sk.dgcut.p, b.sk.nsp,
                                 b.dq.p, b.grad.p));
  DH_TRY(launch_batch_invr_grad(s, b.L, nat, b.coordsD.p, b.dq.p, b.grad.p));
  DH_HIP(hipMemcpyAsync(grad, b.grad.p, 3 * (size_t)nat * sizeof(double), hipMemcpyDeviceToHost,
                        s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace dftbhip

struct dftbhip_batch {
  dftbhip::BatchContext b;
};

extern "C" int dftbhip_batch_create(dftbhip_batch** bp, const dftbhip_params* p,
                                    const dftbhip_options* opt) {
  DH_CHECK(bp && p && opt, "batch_create: null pointer");
  DH_CHECK(opt->precision == 0, "batch_create: the molecule batch runs in FP64 (precision 0)");
  DH_CHECK(opt->max_scc_iter >= 1, "batch_create: max_scc_iter < 1");
  DH_TRY(dftbhip::ensure_runtime());
  auto* h = new dftbhip_batch();
  h->b.opt = *opt;
  h->b.s = dftbhip::g_rt.stream;
  for (int q = 0; q < dftbhip::BatchContext::kSx; ++q) {
    if (hipStreamCreate(&h->b.sx[q]) != hipSuccess || hipEventCreate(&h->b.join[q]) != hipSuccess) {
      delete h;
      dftbhip::set_error("batch_create: stream or event creation failed");
      return dftbhip::kErrDevice;
    }
  }
  if (hipEventCreate(&h->b.fork) != hipSuccess) {
    delete h;
    dftbhip::set_error("batch_create: event creation failed");
    return dftbhip::kErrDevice;
  }
  const int rc = dftbhip::sk_setup(h->b.sk, p, h->b.s);
  if (rc != 0) {
    delete h;
    return rc;
  }
  *bp = h;
  return 0;
}

extern "C" int dftbhip_batch_destroy(dftbhip_batch* b) {
  delete b;
  return 0;
}

extern "C" int dftbhip_batch_set_geometries(dftbhip_batch* b, int nMol, const int* nAtomMol,
                                            const double* coords, const int* species) {
  DH_CHECK(b && nAtomMol && coords && species && nMol > 0, "batch_set_geometries: bad arguments");
  return dftbhip::batch_geometry(b->b, nMol, nAtomMol, coords, species);
}

extern "C" int dftbhip_batch_scc(dftbhip_batch* b, double* energies, double* charges, int* niter,
                                 int* converged) {
  DH_CHECK(b && b->b.nmol > 0, "batch_scc: no geometries");
  return dftbhip::batch_scc(b->b, energies, charges, niter, converged);
}

extern "C" int dftbhip_batch_scc_begin(dftbhip_batch* b) {
  DH_CHECK(b && b->b.nmol > 0, "batch_scc_begin: no geometries");
  return dftbhip::batch_scc_begin(b->b);
}

extern "C" int dftbhip_batch_scc_step(dftbhip_batch* b, int* nactive) {
  DH_CHECK(b, "batch_scc_step: null pointer");
  DH_TRY(dftbhip::batch_scc_step(b->b));
  if (nactive) *nactive = b->b.nactive;
  return 0;
}

extern "C" int dftbhip_batch_scc_end(dftbhip_batch* b, double* energies, double* charges,
                                     int* niter, int* converged) {
  DH_CHECK(b, "batch_scc_end: null pointer");
  return dftbhip::batch_scc_end(b->b, energies, charges, niter, converged);
}

extern "C" int dftbhip_batch_gradients(dftbhip_batch* b, double* grad) {
  DH_CHECK(b && grad, "batch_gradients: null pointer");
  return dftbhip::batch_gradients(b->b, grad);
}

struct dftbhip_ctx {
  dftbhip::Context c;
};

extern "C" int dftbhip_ctx_create(dftbhip_ctx** ctx, const dftbhip_params* p,
                                  const dftbhip_options* opt) {
  DH_CHECK(ctx && p && opt, "ctx_create: null pointer");
  DH_CHECK(opt->precision == 0 || opt->precision == 1,
           "ctx_create: precision must be 0 (FP64) or 1 (mixed FP32/FP64)");
  DH_CHECK(opt->max_scc_iter >= 1, "ctx_create: max_scc_iter < 1");
  DH_CHECK(opt->tridiag == 0 || opt->tridiag == 1, "ctx_create: tridiag must be 0 or 1");
  DH_TRY(ensure_runtime());
  auto* h = new dftbhip_ctx();
  h->c.opt = *opt;
  h->c.s = g_rt.stream;
  const int rc = sk_setup(h->c.sk, p, h->c.s);
  if (rc != 0) {
    delete h;
    return rc;
  }
  *ctx = h;
  return 0;
}

extern "C" int dftbhip_ctx_destroy(dftbhip_ctx* ctx) {
  delete ctx;
  return 0;
}

extern "C" int dftbhip_ctx_set_kpoints(dftbhip_ctx* ctx, int nK, const double* kpts,
                                       const double* weights) {
  DH_CHECK(ctx && nK >= 0 && (nK == 0 || kpts), "set_kpoints: bad arguments");
  dftbhip::Context& c = ctx->c;
  c.nk = nK;
  c.kpts.assign(kpts, kpts + 3 * (size_t)nK);
  c.kw.assign(nK, 1.0);
  if (weights) c.kw.assign(weights, weights + nK);
  double sum = 0.0;
  for (double w : c.kw) {
    DH_CHECK(w >= 0.0, "set_kpoints: negative weight");
    sum += w;
  }
  DH_CHECK(nK == 0 || sum > 0.0, "set_kpoints: weights sum to zero");
  for (double& w : c.kw) w /= sum;
  if (c.nat > 0)   // re-derive the per-k-point overlaps of the current geometry
    return dftbhip::ctx_geometry(c, c.nat, c.coords.data(), c.species.data(),
                                 c.periodic ? c.lat : nullptr);
  return 0;
}

extern "C" int dftbhip_ctx_set_geometry(dftbhip_ctx* ctx, int nAtom, const double* coords,
                                        const int* species, const double* latvecs) {
  DH_CHECK(ctx && coords && species && nAtom > 0, "set_geometry: bad arguments");
  return ctx_geometry(ctx->c, nAtom, coords, species, latvecs);
}

extern "C" int dftbhip_ctx_scc(dftbhip_ctx* ctx, double* energies, double* charges, int* niter,
                               int* converged) {
  DH_CHECK(ctx && ctx->c.nat > 0, "scc: no geometry");
  return ctx_scc(ctx->c, energies, charges, niter, converged);
}

extern "C" int dftbhip_ctx_set_input_charges(dftbhip_ctx* ctx, int nAtom, const double* qin,
                                             int checkCount) {
  DH_CHECK(ctx && qin, "set_input_charges: null pointer");
  dftbhip::Context& c = ctx->c;
  DH_CHECK(c.nat > 0 && nAtom == c.nat, "set_input_charges: atom count differs from the geometry");
  double nel = 0.0, sum = 0.0;
  for (int a = 0; a < nAtom; ++a) {
    nel += c.q0atom[a];
    sum += qin[a];
  }
  // the electron-count check of initQFromFile (sccinit.F90:397-402), skipped as the reference's
  // SkipChargeTest does
  if (checkCount && std::fabs(nel - sum) >= 1e-3) {
    char buf[160];
    std::snprintf(buf, sizeof buf,
                  "External file of charges has a total charge:%18.6f, instead of %18.6f", sum,
                  nel);
    dftbhip::set_error(buf);
    return dftbhip::kErrArg;
  }
  c.qinGiven.assign(qin, qin + nAtom);
  return 0;
}

extern "C" int dftbhip_ctx_get_input_charges(dftbhip_ctx* ctx, int nAtom, double* qin) {
  DH_CHECK(ctx && qin, "get_input_charges: null pointer");
  const dftbhip::Context& c = ctx->c;
  DH_CHECK(nAtom == c.nat && c.st.qin.size() == (size_t)c.nat,
           "get_input_charges: no SCC cycle on this geometry, or the atom count differs");
  std::copy(c.st.qin.begin(), c.st.qin.end(), qin);
  return 0;
}

extern "C" int dftbhip_ctx_set_charge_restart(dftbhip_ctx* ctx, int warm) {
  DH_CHECK(ctx, "set_charge_restart: null context");
  ctx->c.warmStart = warm != 0;
  return 0;
}

extern "C" int dftbhip_ctx_scc_begin(dftbhip_ctx* ctx) {
  DH_CHECK(ctx && ctx->c.nat > 0, "scc_begin: no geometry");
  return ctx_scc_begin(ctx->c);
}

extern "C" int dftbhip_ctx_scc_step(dftbhip_ctx* ctx, int* converged, double* qerr) {
  DH_CHECK(ctx && ctx->c.nat > 0, "scc_step: no geometry");
  DH_TRY(ctx_scc_step(ctx->c));
  if (converged) *converged = ctx->c.st.conv ? 1 : 0;
  if (qerr) *qerr = ctx->c.st.err;
  return 0;
}

extern "C" int dftbhip_ctx_scc_end(dftbhip_ctx* ctx, double* energies, double* charges,
                                   int* niter, int* converged) {
  DH_CHECK(ctx, "scc_end: null context");
  return ctx_scc_end(ctx->c, energies, charges, niter, converged);
}

extern "C" int dftbhip_ctx_scc_state(dftbhip_ctx* ctx, double* out, int n) {
  DH_CHECK(ctx && out && n >= 0, "scc_state: bad arguments");
  return ctx_scc_state(ctx->c, out, n);
}

extern "C" int dftbhip_ctx_gradients(dftbhip_ctx* ctx, double* grad) {
  DH_CHECK(ctx && grad, "gradients: null pointer");
  return ctx_gradients(ctx->c, grad);
}

extern "C" int dftbhip_ctx_eigenvalues(dftbhip_ctx* ctx, double* eig, double* fill, int n) {
  DH_CHECK(ctx && ctx->c.sccDone, "eigenvalues: run the SCC cycle first");
  DH_CHECK(n == ctx->c.norbTot, "eigenvalues: size mismatch");
  if (eig) std::copy(ctx->c.eigv.begin(), ctx->c.eigv.end(), eig);
  if (fill) std::copy(ctx->c.fill.begin(), ctx->c.fill.end(), fill);
  return 0;
}

extern "C" int dftbhip_ctx_eigenvalues_k(dftbhip_ctx* ctx, int k, double* eig, double* fill,
                                         int n) {
  DH_CHECK(ctx && ctx->c.sccDone, "eigenvalues_k: run the SCC cycle first");
  const dftbhip::Context& c = ctx->c;
  DH_CHECK(n == c.norbTot, "eigenvalues_k: size mismatch");
  if (c.nk == 0) {
    DH_CHECK(k == 0, "eigenvalues_k: k out of range");
    if (eig) std::copy(c.eigv.begin(), c.eigv.end(), eig);
    if (fill) std::copy(c.fill.begin(), c.fill.end(), fill);
    return 0;
  }
  DH_CHECK(k >= 0 && k < c.nk, "eigenvalues_k: k out of range");
  const int stride = dftbhip::kpoint_dim(&c.kpts[3 * k], n) / n;
  for (int j = 0; j < n; ++j) {
    if (eig) eig[j] = c.eigK[k][(size_t)stride * j];
    if (fill) fill[j] = c.fillK[k][j];
  }
  return 0;
}

extern "C" int dftbhip_ctx_orbital_populations(dftbhip_ctx* ctx, double* q, int n) {
  DH_CHECK(ctx && ctx->c.sccDone && q, "orbital_populations: run the SCC cycle first");
  DH_CHECK(n == ctx->c.norbTot && (int)ctx->c.qorbLast.size() == n,
           "orbital_populations: size mismatch");
  std::copy(ctx->c.qorbLast.begin(), ctx->c.qorbLast.end(), q);
  return 0;
}

extern "C" int dftbhip_ctx_nr_orbitals(dftbhip_ctx* ctx) { return ctx ? ctx->c.norbTot : -1; }

extern "C" int dftbhip_ctx_timings(dftbhip_ctx* ctx, double* t, int n) {
  DH_CHECK(ctx && t, "timings: null pointer");
  for (int i = 0; i < n && i < kTCount; ++i) t[i] = ctx->c.times[i];
  return ctx->c.ncalls;
}

// ------------------------------------------------------------------------------------------
// Host-side pieces of the SCC cycle with the reference's interfaces: Fermi filling and the
// Broyden mixer (no device work; callable without a GPU)
// ------------------------------------------------------------------------------------------
extern "C" int dftbhip_efilling(int nLevels, int nKPoints, const double* eigenvals,
                                const double* kWeight, double nElectrons, double kT,
                                double* filling, double* Ef, double* Ebs, double* TS,
                                double* E0) {
  DH_CHECK(nLevels >= 0 && nKPoints >= 1, "efilling: bad dimensions");
  DH_CHECK(eigenvals && kWeight && filling && Ef && Ebs && TS && E0, "efilling: null pointer");
  DH_CHECK(nElectrons >= 0.0 && kT > 0.0, "efilling: negative electron count or temperature");
  DH_CHECK(std::ceil(nElectrons) <= 2.0 * nLevels, "efilling: too many electrons for the levels");
  double wsum = 0.0;
  for (int k = 0; k < nKPoints; ++k) {
    DH_CHECK(kWeight[k] >= 0.0, "efilling: negative k-point weight");
    wsum += kWeight[k];
  }
  DH_CHECK(wsum > 0.0, "efilling: k-point weights sum to zero");
  for (int k = 0; k < nKPoints; ++k)
    for (int j = 1; j < nLevels; ++j)
      DH_CHECK(eigenvals[j + (size_t)k * nLevels] >= eigenvals[j - 1 + (size_t)k * nLevels],
               "efilling: eigenvalues of a k-point must be ascending");
  // the SCC cycle's fillings carry the factor 2 of the spin-unpolarised case; the
  // reference's Efilling fills one spin channel with nElectrons
  FillResult r{};
  if (nKPoints == 1 && kWeight[0] == 1.0) {
    std::vector<double> e(eigenvals, eigenvals + nLevels), f;
    DH_TRY(fermi_filling(e, 2.0 * nElectrons, kT, f, r));
    for (int j = 0; j < nLevels; ++j) filling[j] = 0.5 * f[j];
  } else {
    std::vector<std::vector<double>> e(nKPoints), f;
    std::vector<double> kw(nKPoints);
    for (int k = 0; k < nKPoints; ++k) {
      e[k].assign(eigenvals + (size_t)k * nLevels, eigenvals + (size_t)(k + 1) * nLevels);
      kw[k] = kWeight[k];
    }
    DH_TRY(fermi_filling_k(e, kw, 2.0 * nElectrons, kT, f, r));
    for (int k = 0; k < nKPoints; ++k)
      for (int j = 0; j < nLevels; ++j) filling[j + (size_t)k * nLevels] = 0.5 * f[k][j];
  }
  *Ef = r.ef;
  *Ebs = 0.5 * r.eband;
  *TS = 0.5 * r.ts;
  *E0 = 0.5 * r.e0;
  return 0;
}

struct dftbhip_mixer {
  dftbhip::Broyden b;
  dftbhip_options o{};
  int maxIter = 0;
  int nElem = 0;
};

extern "C" int dftbhip_mixer_create(dftbhip_mixer** m, int maxSccIter, double mixParam,
                                    double omega0, double minWeight, double maxWeight,
                                    double weightFac) {
  DH_CHECK(m, "mixer_create: null pointer");
  DH_CHECK(maxSccIter > 0 && mixParam > 0.0 && omega0 > 0.0, "mixer_create: bad parameters");
  auto* x = new dftbhip_mixer();
  x->maxIter = maxSccIter;
  x->o.mix_param = mixParam;
  x->o.omega0 = omega0;
  x->o.min_weight = minWeight;
  x->o.max_weight = maxWeight;
  x->o.weight_fac = weightFac;
  *m = x;
  return 0;
}

extern "C" int dftbhip_mixer_reset(dftbhip_mixer* m, int nElem) {
  DH_CHECK(m && nElem > 0, "mixer_reset: bad arguments");
  m->nElem = nElem;
  m->b.reset(nElem, m->o, m->maxIter);
  return 0;
}

extern "C" int dftbhip_mixer_mix(dftbhip_mixer* m, double* qInpResult, const double* qDiff) {
  DH_CHECK(m && qInpResult && qDiff, "mixer_mix: null pointer");
  DH_CHECK(m->nElem > 0, "mixer_mix: call mixer_reset first");
  if (m->b.it + 1 > m->maxIter) {
    set_error("Broyden mixer: Maximal nr. of steps exceeded");
    return kErrState;
  }
  std::vector<double> q(qInpResult, qInpResult + m->nElem), d(qDiff, qDiff + m->nElem);
  m->b.mix(q, d);
  std::copy(q.begin(), q.end(), qInpResult);
  return 0;
}

extern "C" int dftbhip_mixer_destroy(dftbhip_mixer* m) {
  delete m;
  return 0;
}

// ------------------------------------------------------------------------------------------
// BLAS/LAPACK-style drop-ins on host arrays
// ------------------------------------------------------------------------------------------
extern "C" int dftbhip_dgemm(char transa, char transb, int m, int n, int k, double alpha,
                             const double* a, int lda, const double* b, int ldb, double beta,
                             double* c, int ldc) {
  const bool ta = (transa == 'T' || transa == 't'), tb = (transb == 'T' || transb == 't');
  DH_CHECK(m >= 0 && n >= 0 && k >= 0, "dgemm: negative size");
  DH_CHECK(a && b && c, "dgemm: null pointer");
  const int ra = ta ? k : m, ca = ta ? m : k, rb = tb ? n : k, cb = tb ? k : n;
  DH_CHECK(lda >= std::max(1, ra) && ldb >= std::max(1, rb) && ldc >= std::max(1, m),
           "dgemm: bad leading dimension");
  if (m == 0 || n == 0) return 0;
  DH_TRY(ensure_runtime());
  hipStream_t s = g_rt.stream;
  DevBuf<double> A, B, Cm;
  DH_TRY(A.alloc((size_t)std::max(ra, 1) * std::max(ca, 1)));
  DH_TRY(B.alloc((size_t)std::max(rb, 1) * std::max(cb, 1)));
  DH_TRY(Cm.alloc((size_t)m * n));
  if (ra > 0 && ca > 0)
    DH_HIP(hipMemcpy2DAsync(A.p, ra * sizeof(double), a, lda * sizeof(double), ra * sizeof(double),
                            ca, hipMemcpyHostToDevice, s));
  if (rb > 0 && cb > 0)
    DH_HIP(hipMemcpy2DAsync(B.p, rb * sizeof(double), b, ldb * sizeof(double), rb * sizeof(double),
                            cb, hipMemcpyHostToDevice, s));
  DH_HIP(hipMemcpy2DAsync(Cm.p, m * sizeof(double), c, ldc * sizeof(double), m * sizeof(double), n,
                          hipMemcpyHostToDevice, s));
  DH_TRY(gemm<double>(s, ta, tb, m, n, k, alpha, A.p, std::max(ra, 1), B.p, std::max(rb, 1), beta,
                      Cm.p, m));
  DH_HIP(hipMemcpy2DAsync(c, ldc * sizeof(double), Cm.p, m * sizeof(double), m * sizeof(double), n,
                          hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

extern "C" int dftbhip_dsytrd(char uplo, int n, double* a, int lda, double* d, double* e,
                              double* tau) {
  DH_CHECK(uplo == 'L' || uplo == 'l', "dsytrd: only uplo = 'L'");
  DH_CHECK(n >= 0 && a && d && (n < 2 || (e && tau)) && lda >= std::max(n, 1),
           "dsytrd: bad arguments");
  if (n == 0) return 0;
  DH_TRY(dftbhip::ensure_runtime());
  hipStream_t s = dftbhip::g_rt.stream;
  dftbhip::DevBuf<double> A, D, E, TAU, W;
  DH_TRY(A.alloc((size_t)n * n));
  DH_TRY(D.alloc(n));
  DH_TRY(E.alloc(std::max(n, 2)));
  DH_TRY(TAU.alloc(std::max(n, 1)));
  const size_t ws = dftbhip::sytrd_work_size(n);
  DH_TRY(W.alloc(ws));
  DH_HIP(hipMemcpy2DAsync(A.p, n * sizeof(double), a, (size_t)lda * sizeof(double),
                          n * sizeof(double), n, hipMemcpyHostToDevice, s));
  DH_TRY(dftbhip::sytrd_lower(s, A.p, n, n, D.p, E.p, TAU.p, W.p, ws));
  DH_HIP(hipMemcpy2DAsync(a, (size_t)lda * sizeof(double), A.p, n * sizeof(double),
                          n * sizeof(double), n, hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(d, D.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (n > 1) {
    DH_HIP(hipMemcpyAsync(e, E.p, (n - 1) * sizeof(double), hipMemcpyDeviceToHost, s));
    DH_HIP(hipMemcpyAsync(tau, TAU.p, (n - 1) * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

// Tridiagonalisation of device-resident arrays with the workspace in the fixed device array
// (no allocation): the same kernels as dftbhip_dsytrd, on the library stream, synchronised
// before returning. Calls that overlap on one GPU share the array and race.
extern "C" int dftbhip_dsytrd_noalloc(int n, double* a, int lda, double* d, double* e,
                                      double* tau) {
  DH_CHECK(n >= 0 && a && d && (n < 2 || (e && tau)) && lda >= std::max(n, 1),
           "dsytrd_noalloc: bad arguments");
  if (n == 0) return 0;
  if (!dftbhip::g_rt.stream) {
    dftbhip::set_error("dsytrd_noalloc: call dftbhip_init first");
    return dftbhip::kErrState;
  }
  const size_t need = dftbhip::sytrd_work_bytes(n, 8);
  if (need > dftbhip::kScratchBytes) {
    dftbhip::set_error("dsytrd_noalloc: n = " + std::to_string(n) + " needs " +
                       std::to_string(need) + " bytes of workspace, more than the fixed " +
                       std::to_string(dftbhip::kScratchBytes));
    return dftbhip::kErrArg;
  }
  void* ws = nullptr;
  DH_TRY(dftbhip::scratch_ptr(&ws));
  hipStream_t s = dftbhip::g_rt.stream;
  DH_TRY(dftbhip::sytrd_lower(s, a, lda, n, d, e, tau, static_cast<double*>(ws), need / 8));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

extern "C" int dftbhip_dlacpy(char uplo, int m, int n, const double* a, int lda, double* b,
                              int ldb) {
  DH_CHECK(uplo == 'A' || uplo == 'a', "dlacpy: only uplo = 'A'");
  DH_CHECK(m >= 0 && n >= 0 && a && b && lda >= std::max(m, 1) && ldb >= std::max(m, 1),
           "dlacpy: bad arguments");
  if (m == 0 || n == 0) return 0;
  DH_TRY(dftbhip::ensure_runtime());
  hipStream_t s = dftbhip::g_rt.stream;
  dftbhip::DevBuf<double> A, B;
  DH_TRY(A.alloc((size_t)m * n));
  DH_TRY(B.alloc((size_t)m * n));
  DH_HIP(hipMemcpy2DAsync(A.p, m * sizeof(double), a, (size_t)lda * sizeof(double),
                          m * sizeof(double), n, hipMemcpyHostToDevice, s));
  DH_TRY(dftbhip::copy2d<double>(s, A.p, m, B.p, m, m, n));
  DH_HIP(hipMemcpy2DAsync(b, (size_t)ldb * sizeof(double), B.p, m * sizeof(double),
                          m * sizeof(double), n, hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

extern "C" int dftbhip_set_tuning(int param, int value) {
  switch (param) {
    case 0:
      return dftbhip::set_symv_tiles(value);
    case 1:
      DH_CHECK(value == 0, "set_tuning: parameter 1 takes 0 (drop the autotuned tables)");
      return dftbhip::clear_symv_tables();
    case 2:
      DH_CHECK(value == 0 || value == 1, "set_tuning: parameter 2 takes 0 or 1");
      dftbhip::g_kBatch = value;
      return 0;
    case 3:
      DH_CHECK(value == 0 || value == 1, "set_tuning: parameter 3 takes 0 or 1");
      dftbhip::g_devNeighbours = value;
      return 0;
    case 4:
      DH_CHECK(value >= 0 && value <= 2, "set_tuning: parameter 4 takes 0, 1 or 2");
      dftbhip::g_batchScc = value;
      return 0;
    case 5:
      return dftbhip::set_gemm_order(value);
    case 6:
      return dftbhip::set_oa_mode(value);
    case 7:
      DH_CHECK(value == 0 || value == 1, "set_tuning: parameter 7 takes 0 or 1");
      dftbhip::g_tailMixReset = value;
      return 0;
    case 8:
      DH_CHECK(value >= 0 && value <= 2, "set_tuning: parameter 8 takes 0, 1 or 2");
      dftbhip::g_tailUpdate32 = value;
      return 0;
    case 9:
      return dftbhip::set_spmm_cols(value);
    case 10:
      DH_CHECK(value >= -1 && value <= 8, "set_tuning: parameter 10 takes -1 .. 8");
      dftbhip::g_tailBand = value;
      return 0;
    case 12:
      DH_CHECK(value == 0 || value == 1, "set_tuning: parameter 12 takes 0 or 1");
      return dftbhip::set_sytrd_graphs(value == 1);
    case 13:
      DH_CHECK(value == 0 || value == 1, "set_tuning: parameter 13 takes 0 or 1");
      return dftbhip::set_sb2st_single(value == 1);
    case 11:
      DH_CHECK(value >= 0 && value <= 15, "set_tuning: parameter 11 takes 0 .. 15");
      dftbhip::g_tailFineBelow = value == 0 ? 0.0 : std::pow(10.0, -value);
      return 0;
    default:
      dftbhip::set_error("set_tuning: unknown parameter " + std::to_string(param));
      return dftbhip::kErrArg;
  }
}

extern "C" int dftbhip_neighbour_list(int nAtom, const double* coords, const double* latVecs,
                                      double cutoff, int* nAll, int* mNeighbour,
                                      double* coordsAll, int* img2CentCell, int* iNeighbour,
                                      int* nNeighbour, double* neighDist2) {
  DH_CHECK(nAtom >= 1 && coords && cutoff > 0.0 && nAll && mNeighbour,
           "neighbour_list: bad arguments");
  const bool fillOut = coordsAll != nullptr;
  DH_CHECK(!fillOut || (img2CentCell && iNeighbour && nNeighbour),
           "neighbour_list: output arrays missing");
  const int capAll = *nAll, capNb = *mNeighbour;
  DH_TRY(ensure_runtime());
  hipStream_t s = g_rt.stream;
  DevNeighbourList nl;
  std::vector<int> ones(nAtom, 1);
  DH_TRY(device_neighbours(s, nAtom, coords, latVecs, cutoff, cutoff, ones.data(), nl));
  *nAll = nl.nAll;
  *mNeighbour = nl.maxNeighbour;
  if (!fillOut) return 0;
  DH_CHECK(capAll >= nl.nAll && capNb >= nl.maxNeighbour,
           "neighbour_list: output arrays smaller than nAll = " + std::to_string(nl.nAll) +
               ", mNeighbour = " + std::to_string(nl.maxNeighbour));
  std::vector<int> img(nl.nPairs), cnt(nAtom), i2c(nl.nAll);
  std::vector<long> start(nAtom + 1);
  std::vector<double> d2(nl.nPairs);
  DH_HIP(hipMemcpyAsync(coordsAll, nl.coordsAll.p, 3 * (size_t)nl.nAll * sizeof(double),
                        hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(i2c.data(), nl.img2cent.p, nl.nAll * sizeof(int), hipMemcpyDeviceToHost,
                        s));
  DH_HIP(hipMemcpyAsync(img.data(), nl.pImg2.p, nl.nPairs * sizeof(int), hipMemcpyDeviceToHost,
                        s));
  DH_HIP(hipMemcpyAsync(d2.data(), nl.pD2.p, nl.nPairs * sizeof(double), hipMemcpyDeviceToHost,
                        s));
  DH_HIP(hipMemcpyAsync(cnt.data(), nl.pCount.p, nAtom * sizeof(int), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpyAsync(start.data(), nl.pStart.p, (nAtom + 1) * sizeof(long),
                        hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < nl.nAll; ++i) img2CentCell[i] = i2c[i] + 1;
  const long ld = (long)capNb + 1;
  for (int a = 0; a < nAtom; ++a) {
    nNeighbour[a] = cnt[a];
    for (long k = 0; k <= capNb; ++k) {
      const bool in = k <= cnt[a];
      iNeighbour[k + a * ld] = in ? img[start[a] + k] + 1 : 0;
      if (neighDist2) neighDist2[k + a * ld] = in ? d2[start[a] + k] : 0.0;
    }
  }
  return 0;
}

extern "C" int dftbhip_autotune(int n, int precision, int* table, int maxRows, int* nRows,
                                double* detail, int detailLen) {
  DH_CHECK(maxRows >= 0 && (maxRows == 0 || table) && nRows, "autotune: bad arguments");
  DH_TRY(ensure_runtime());
  std::vector<std::pair<int, int>> rows;
  std::vector<double> det;
  DH_TRY(symv_autotune(g_rt.stream, n, precision, rows, detail ? &det : nullptr));
  *nRows = (int)rows.size();
  for (int r = 0; r < std::min(maxRows, (int)rows.size()); ++r) {
    table[2 * r] = rows[r].first;
    table[2 * r + 1] = rows[r].second;
  }
  if (detail)
    for (int k = 0; k < std::min(detailLen, (int)det.size()); ++k) detail[k] = det[k];
  return 0;
}

extern "C" int dftbhip_dstedc(int n, double* d, double* e, double* z, int ldz) {
  DH_CHECK(n >= 0 && d && (n <= 1 || e) && z && ldz >= std::max(1, n), "dstedc: bad arguments");
  if (n == 0) return 0;
  DH_TRY(ensure_runtime());
  hipStream_t s = g_rt.stream;
  DevBuf<double> w, Z, work;
  DH_TRY(w.alloc(n));
  DH_TRY(Z.alloc((size_t)n * n));
  const size_t ws = stedc_work_size(n);
  DH_TRY(work.alloc(ws));
  int info = 0;
  DH_TRY(stedc(s, n, d, e, w.p, Z.p, n, work.p, ws, &info));
  if (info != 0) return info;
  DH_HIP(hipMemcpyAsync(d, w.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipMemcpy2DAsync(z, ldz * sizeof(double), Z.p, n * sizeof(double), n * sizeof(double), n,
                          hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

extern "C" int dftbhip_kernel_profile(int enable) {
  symv_profiling(enable != 0);
  return 0;
}

extern "C" int dftbhip_kernel_stats(double* out) {
  DH_CHECK(out, "kernel_stats: null pointer");
  symv_profile_stats(out);
  return 0;
}
