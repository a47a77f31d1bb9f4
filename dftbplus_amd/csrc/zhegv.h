// Complex Hermitian generalized eigensolver on split planes (zhegv.hip).
#pragma once

#include <memory>
#include <vector>

#include "common.h"

namespace dftbhip {

// A complex matrix as two real planes with a common leading dimension.
struct ZcPlanes {
  const double* r;
  const double* i;
  long ld;
};
struct ZPlanes {
  double* r;
  double* i;
  long ld;
  ZcPlanes c() const { return ZcPlanes{r, i, ld}; }
};

// C = alpha op(A) op(B) + beta C with op 'N' or 'C' (conjugate transpose); real alpha, beta
int zgemm(hipStream_t s, char opA, char opB, int M, int N, int K, double alpha, ZcPlanes A,
          ZcPlanes B, double beta, ZPlanes C, bool lowerOut, GemmKLimit klim);
// re/im of a Hermitian matrix completed from the lower triangle (imaginary diagonal zeroed)
int hermitian_from_lower(hipStream_t s, double* re, double* im, long ld, int n);

struct ZhegvSolver {
  int n = 0;
  int nClusters = 0;   // degenerate clusters of the last solve (Gram-Schmidt)
  // planes: H (then the reduced A, then the eigenvectors), L (the factor), V (L^-1), A (the
  // product Linv H, then the standard-form vectors)
  DevBuf<double> Hr, Hi, Lr, Li, Vr, Vi, Ar, Ai;
  DevBuf<double> E2, Z2, d2, e2, tau2, w2, work;   // the real 2n embedding
  DevBuf<int> dinfo, clusterDev;
  size_t workSize = 0;
  std::shared_ptr<StedcWs> sws;
  int ensure(int n_);
  void release();
  // aDev, bDev: interleaved complex n x n (leading dimensions in complex elements), lower
  // triangles read. On return: wHost the ascending eigenvalues; with vectors the S-orthonormal
  // eigenvectors in (Hr, Hi); the Cholesky factor in the lower triangles of (Lr, Li). info:
  // 0, n + minor for an overlap that is not positive definite, or the divide and conquer's.
  int solve(hipStream_t s, const double* aDev, long lda, const double* bDev, long ldb,
            bool vectors, std::vector<double>& wHost, int* info);
};

}  // namespace dftbhip
