// Per-molecule layout of a molecule batch and the batched kernels (batch.hip).
#pragma once

#include "sparse.h"

namespace dftbhip {

constexpr int kBatchMaxOrb = 512;   // orbitals per molecule in a batch
// size classes of the batched eigensolver, by orbital count: up to kBatchLdsSmall and
// kBatchLdsC the reduced matrix A and the eigenvectors C stay in LDS (49 / 153 KB per
// workgroup), up to kBatchLdsMid A alone (148 KB, C in HBM), larger molecules keep both in HBM
constexpr int kBatchLdsSmall = 64;
constexpr int kBatchLdsC = 112;
constexpr int kBatchLdsMid = 190;
constexpr int kBatchClasses = 4;
// info of a molecule whose Jacobi sweeps did not converge within their limit (distinct from the
// leading minor of a non-positive-definite overlap and from the negative size-class error)
constexpr int kBatchJacobiNoConv = 1 << 24;

struct BatchLayout {
  int nmol;
  const int* atom0;     // nmol + 1: first atom of each molecule
  const long* orb0;     // nmol + 1: first orbital
  const long* dOff;     // nmol + 1: offsets of the packed n_m x n_m dense matrices
  const long* gOff;     // nmol + 1: offsets of the packed nat_m x nat_m gamma matrices
  const int* atomMol;   // molecule of each atom
  const int* order;     // nmol: molecules by decreasing orbital count (launch order)
  int ncls[kBatchClasses];   // molecules per size class, largest class first, in order
};

int launch_batch_unpack(hipStream_t s, const PairsDev& pr, const int* norb,
                        const long* iAtomStart, const BatchLayout& L, const double* x,
                        double* dense, long denseSize);
int launch_batch_pack(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                      const BatchLayout& L, const double* dense, double* x);
// eigenvalues w (ascending per molecule, at orb0), eigenvectors C (packed), info per molecule
// (0, or j+1 when the leading minor j+1 of S is not positive definite); work 3 x packed size,
// kept between calls. Molecules with active[m] == 0 are skipped (active may be null: all).
// warm: S is the one of the previous call on the same work array and the Jacobi sweeps start
// from that call's eigenvectors (a cold call factors S).
// st[0 .. kBatchClasses-1]: the streams of the size classes (largest first; may coincide)
int launch_batch_sygv(const hipStream_t* st, const BatchLayout& L, const double* S,
                      const double* H, double* work, double* w, double* C, int* info,
                      const int* active, bool warm);
int launch_batch_density(hipStream_t s, const BatchLayout& L, const double* C, const double* f,
                         const int* nocc, double* P, const int* active);
int launch_batch_gamma(hipStream_t s, const BatchLayout& L, const double* coords,
                       const int* species, const double* hubbu, const double* gcut, int nsp,
                       double* g);
int launch_batch_gemv(hipStream_t s, const BatchLayout& L, int nat, const double* g,
                      const double* dq, double* v);
int launch_batch_invr_grad(hipStream_t s, const BatchLayout& L, int nat, const double* coords,
                           const double* dq, double* grad);

// Device state of a molecule batch's SCC cycle (batch_scc.hip): per atom (nat) and per molecule
// (nmol) arrays, the Broyden histories (maxHist steps: ww, cc per molecule; aa, B maxHist^2 per
// molecule; dF, uu one vector of all atoms per step) and four counters: iteration, molecules
// still iterating, failure code (1 eigensolver, 2 Fermi level, 3 mixer history) and the lowest
// failing molecule.
struct BatchSccDev {
  int nmol, nat, maxHist;
  const int* atom0;
  const long* orb0;
  const double* q0;
  const double* nel;
  double *qin, *qout, *dq, *tmp;
  int *conv, *active, *iters, *nocc, *mixIt, *counters;
  double *err, *ef, *eband, *ts, *e0, *eNonScc, *eScc;
  double *ww, *aa, *B, *cc, *dF, *uu, *qlast, *dlast;
  double kT, sccTol, alpha, omega0, minw, maxw, wfac;
  int minSccIter, maxSccIter;
};
int launch_bscc_reset(hipStream_t s, const BatchSccDev& d);
int launch_bscc_begin(hipStream_t s, const BatchSccDev& d);
int launch_bscc_info(hipStream_t s, const BatchSccDev& d, const int* info);
int launch_bscc_fermi(hipStream_t s, const BatchSccDev& d, const double* w, double* fill);
int launch_bscc_outputs(hipStream_t s, const BatchSccDev& d, const long* atomStart,
                        const int* atomMol, const double* qorb, const double* eorb);
int launch_bscc_mix(hipStream_t s, const BatchSccDev& d, const double* vout);

}  // namespace dftbhip
