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

int EigenSolver::solve(hipStream_t s, const double* H, long ldh, std::vector<double>& wHost,
                       int* info, double* times) {
  DH_CHECK(factored, "eigensolver: overlap not factorised");
  *info = 0;
  hipEvent_t ev[5];
  for (auto& x : ev) DH_HIP(hipEventCreate(&x));
  DH_HIP(hipEventRecord(ev[0], s));
  // A = Linv H Linv^T
  DH_TRY(gemm<double>(s, false, false, n, n, n, 1.0, Linv.p, n, H, ldh, 0.0, T.p, n, false,
                      kKALower));
  DH_TRY(gemm<double>(s, false, true, n, n, n, 1.0, T.p, n, Linv.p, n, 0.0, C.p, n, true,
                      kKBUpper));
  DH_HIP(hipEventRecord(ev[1], s));
  DH_TRY(sytrd_lower(s, C.p, n, n, d.p, e.p, tau.p, work.p, workSize));
  DH_HIP(hipEventRecord(ev[2], s));
  std::vector<double> dh(n), eh(std::max(n - 1, 1));
  DH_HIP(hipMemcpyAsync(dh.data(), d.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (n > 1)
    DH_HIP(hipMemcpyAsync(eh.data(), e.p, (n - 1) * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  DH_TRY(stedc(s, n, dh.data(), eh.data(), w.p, Zd.p, n, work.p, workSize, info));
  DH_HIP(hipEventRecord(ev[3], s));
  wHost.resize(n);
  DH_HIP(hipMemcpyAsync(wHost.data(), w.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  if (times) {
    float ms;
    DH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    times[0] += ms * 1e-3;
    DH_HIP(hipEventElapsedTime(&ms, ev[1], ev[2]));
    times[1] += ms * 1e-3;
    DH_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
    times[2] += ms * 1e-3;
  }
  for (auto& x : ev) (void)hipEventDestroy(x);
  return 0;
}

int EigenSolver::vectors(hipStream_t s, int nvec, double* X, long ldx) {
  if (nvec <= 0) return 0;
  DH_TRY(ormtr_lower(s, C.p, n, n, tau.p, Zd.p, n, nvec, work.p, workSize));
  DH_TRY(gemm<double>(s, true, false, n, nvec, n, 1.0, Linv.p, n, Zd.p, n, 0.0, X, ldx, false,
                      kKAUpper));
  return 0;
}

}  // namespace dftbhip
