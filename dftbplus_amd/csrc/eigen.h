// Dense generalized eigensolver state (eigen.hip).
#pragma once

#include <vector>

#include "common.h"

namespace dftbhip {

struct EigenSolver {
  int n = 0;
  bool factored = false;
  DevBuf<double> L, Linv, C, Zd, T, d, e, tau, w, work;
  DevBuf<int> dinfo;
  size_t workSize = 0;

  int ensure(int n_);
  void release() {
    for (DevBuf<double>* b : {&L, &Linv, &C, &Zd, &T, &d, &e, &tau, &w, &work}) b->release();
    dinfo.release();
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
};

}  // namespace dftbhip
