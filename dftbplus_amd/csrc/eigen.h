// Dense generalized eigensolver state (eigen.hip).
#pragma once

#include <memory>
#include <vector>

#include "common.h"

namespace dftbhip {

// Ogita-Aishima correction of n eigenpairs, transposed: lam_i = M_ii / G_ii and
// Et(j, i) = E(i, j) (E as in EigenSolver::refine) from the symmetric M = X^T H X and
// G = X^T S X (both triangles stored, leading dimension n); nClipped (device, may be null)
// counts the pairs outside the first-order range
int oa_correction_t(hipStream_t s, const double* M, const double* G, int n, double* lam,
                    double* Et, int* nClipped = nullptr);
// pairs outside the first-order range: 0 re-orthonormalised only (the default), 1 rotated by
// the angle of the exact 2x2 problem
int set_oa_mode(int mode);

// Residual form of the same step for nv vectors held transposed (Xt, YH^T = (H X)^T and
// YS^T = (S X)^T, nv x n, leading dimension ld), with G = X^T S X (FP64; its diagonal and, for
// the correction, its lower triangle, or only the band of 128 x 128 tiles of gemm_band):
//  * oa_rayleigh: lam_j = x_j^T H x_j / G_jj (FP64 row dot products; part: 64 nv doubles);
//  * oa_residual_scale: per vector j a power of two rsc_j with max_k |r_j(k)| rsc_j in
//    [2^14, 2^15), r_j = (H - lam_j S) x_j, and rinv_j = 1 / rsc_j (FP32);
//  * oa_correction_f16: from Ft(j, i) = x_i^T r_j (FP32, leading dimension ldf) the correction
//    E(i, j) (the off-diagonal first-order terms x_i^T r_j / (lam_j - lam_i); pairs outside the
//    first-order range and the diagonal from G, as in oa_correction_t, which gives the same E
//    up to the rounding of Ft): colMax[j] = max_i |E(i, j)| (as FP32 bits) and the pair
//    counters; then oa_correction_f16_write: E16[i + j lde] = E(i, j) 2^q_j in FP16, the
//    largest scaled value of column j in [2^14, 2^15), and rinv[j] = 2^-q_j. gband >= 0: G
//    holds the tile band of gemm_band only; pairs outside the first-order range beyond it are
//    counted in nFar and left uncorrected (repeat with the whole lower triangle, gband -1).
int oa_rayleigh(hipStream_t s, const double* Xt, const double* YH, long ld, int nv, int n,
                const double* G, long ldg, double* part, double* lam);
int oa_residual_scale(hipStream_t s, const double* YH, const double* YS, long ld, int nv, int n,
                      const double* lam, double* part, double* rsc, float* rinv);
// Ct (may be null): the couplings of oa_band_coupling, used instead of Ft for the pairs whose
// tiles are at most `band` apart; gband: the tile band G holds (-1: its whole lower triangle).
int oa_correction_f16(hipStream_t s, const float* Ft, long ldf, const double* G, long ldg,
                      const double* Ct, long ldc, const double* lam, int n, int band, int gband,
                      unsigned* colMax, int* nClipped, int* nFar);
int oa_correction_f16_write(hipStream_t s, const float* Ft, long ldf, const double* G, long ldg,
                            const double* Ct, long ldc, const double* lam, int n, int band,
                            int gband, const unsigned* colMax, _Float16* E16, long lde,
                            float* rinv);
// lam_j = M_jj / G_jj and the in-band couplings Ct[j + i ldc] = M_ij - lam_j G_ij (FP64, 0 on
// the diagonal) for the pairs whose 128 x 128 tiles are at most band apart, from the band tiles
// of M = X^T H X and G = X^T S X (lower triangle, leading dimension ldb, both n x n)
int oa_band_coupling(hipStream_t s, const double* Mb, const double* Gb, long ldb, int n, int band,
                     double* lam, double* Ct, long ldc);

struct EigenSolver {
  int n = 0;
  bool factored = false;
  DevBuf<double> L, Linv, C, Zd, T, d, e, tau, w, work;
  DevBuf<int> dinfo;
  DevBuf<float> Linv32, d32, e32, tau32;
  DevBuf<double> band;          // two-stage tridiagonalisation workspace (reflectors of stage 2)
  size_t workSize = 0, bandBytes = 0;
  bool twoStage = true;         // two-stage (band) tridiagonalisation, else one-stage
  std::shared_ptr<StedcWs> sws;   // divide-and-conquer device buffers, kept between solves
  StedcWs* dc_ws() {
    if (!sws) sws.reset(stedc_ws_new(), stedc_ws_delete);
    return sws.get();
  }

  int ensure(int n_);
  void release() {
    for (DevBuf<double>* b : {&L, &Linv, &C, &Zd, &T, &d, &e, &tau, &w, &work, &band})
      b->release();
    for (DevBuf<float>* b : {&Linv32, &d32, &e32, &tau32}) b->release();
    dinfo.release();
    sws.reset();
    n = 0;
    factored = false;
  }
  // Cholesky of S (lower part read, leading dimension lds) and the inverse factor.
  int factor_overlap(hipStream_t s, const double* S, long lds, int* info);
  // Eigenvalues (ascending, to wHost) of the full symmetric H in the S metric; keeps the
  // tridiagonal eigenvectors for vectors(). times[0..2] += reduce, tridiagonalise, D&C (s).
  int solve(hipStream_t s, const double* H, long ldh, std::vector<double>& wHost, int* info,
            double* times = nullptr);
  // The first nvec generalized eigenvectors into X (n x nvec).
  int vectors(hipStream_t s, int nvec, double* X, long ldx);

  // Pieces of solve/vectors for problems that keep their own inverse factor and reflectors
  // (k-points tridiagonalised together): A = Linv H Linv^T (lower triangle, n x n, T as the
  // intermediate); the eigenvalues (to wHost) and tridiagonal eigenvectors (kept for
  // vectors_from) of the device tridiagonal (dDev, eDev); X = Linv^T Q Z for the reflectors
  // (A, tauDev) of the one-stage tridiagonalisation.
  int reduce(hipStream_t s, const double* Li, const double* H, long ldh, double* A);
  int tridiag_eigen(hipStream_t s, const double* dDev, const double* eDev,
                    std::vector<double>& wHost, int* info);
  int vectors_from(hipStream_t s, const double* Li, const double* A, const double* tauDev,
                   int nvec, double* X, long ldx);

  // Mixed precision: the reduction, tridiagonalisation and back-transformation run in FP32
  // (FP32 MFMA, FP32 accumulation); the Cholesky factor comes from the FP64 factor_overlap and
  // the tridiagonal eigenproblem is solved in FP64. prepare_mixed rounds Linv once per geometry.
  int prepare_mixed(hipStream_t s);
  // FP32 staging area for H in refined iterations (second half of the FP64 buffer T)
  float* h32() { return reinterpret_cast<float*>(T.p) + (size_t)n * n; }
  int solve_mixed(hipStream_t s, const float* H, long ldh, std::vector<double>& wHost, int* info,
                  double* times = nullptr);
  int vectors_mixed(hipStream_t s, int nvec, float* X, long ldx);

  // Ogita-Aishima refinement in FP64 of all n approximate generalized eigenpairs X0 (FP32,
  // n x n, from vectors_mixed) against the FP64 matrices H and S (both triangles stored):
  // steps refinement steps; the first nvec refined vectors go to X (FP64), the refined
  // eigenvalues (Rayleigh quotients of the last input basis, unsorted) to lamHost.
  int refine(hipStream_t s, const double* H, const double* S, const float* X0, int steps,
             int nvec, double* X, long ldx, std::vector<double>& lamHost);
  // the same from an FP64 basis X0 (n x n, leading dimension n); X0 may be X itself
  int refine(hipStream_t s, const double* H, const double* S, const double* X0, int steps,
             int nvec, double* X, long ldx, std::vector<double>& lamHost);
  int refine_from_basis(hipStream_t s, const double* H, const double* S, int steps, int nvec,
                        double* X, long ldx, std::vector<double>& lamHost);
};

}  // namespace dftbhip
