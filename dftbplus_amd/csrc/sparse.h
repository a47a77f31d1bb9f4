// Device-side descriptions of the Slater-Koster tables, neighbour blocks and repulsives, and the
// launch wrappers of the sparse kernels (sparse.hip).
#pragma once

#include "common.h"

namespace dftbhip {

constexpr int kMaxSpecies = 8;

struct SpeciesDev {
  int nshell;
  int ang[2];
  int norb;
  double onsite[4];   // per orbital
};

struct SkDev {
  int nsp;
  int nrows;
  double dist;
  const double* htab;   // concatenated tables, row-major (row, integral)
  const double* stab;
  long tabOff[kMaxSpecies * kMaxSpecies];
  int nint[kMaxSpecies * kMaxSpecies];
  SpeciesDev sp[kMaxSpecies];
};

// Sparse blocks (atom, neighbour) within the SK cutoff.
struct PairsDev {
  int nblk;
  const int* at1;       // central atom of the block column
  const int* img2;      // image index of the neighbour (coordinates index)
  const int* at2;       // central atom of the neighbour
  const long* start;    // block offset in the sparse arrays
  const int* species;   // species of central atoms
};

// Neighbour pairs (any cutoff) for gamma and repulsive.
struct NeighDev {
  int n;
  const int* at1;
  const int* img2;
  const int* at2;
  const int* species;
};

struct RepDev {
  int nint;
  double cutoff;
  double expc[3];
  const double* xs;     // nint starts
  const double* coef;   // (nint-1) x 4
  const double* last;   // 6
};

int launch_build_h0s(hipStream_t s, const SkDev& sk, const PairsDev& pr, const double* coords,
                     double* h0, double* ov);
template <typename T>
int launch_unpack(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                  const double* x, T* dense, long ld, int n);
template <typename T>
int launch_pack(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                const T* dense, long ld, double* x);
// k-point matrices (blkT: fractional cell translation of the image of each block, kf:
// fractional k-point): at a time-reversal-invariant point (2 kf integral, kpoint_is_real) the
// phases are +-1 and the matrix is the real n x n one; otherwise its real 2n x 2n embedding.
// pack_kpt accumulates into x.
bool kpoint_is_real(const double* kf);
int kpoint_dim(const double* kf, int n);   // n or 2n
int launch_unpack_kpt(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                      const double* blkT, const double* kf, const double* x, double* dense,
                      long ld, int n);
int launch_pack_kpt(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                    const double* blkT, const double* kf, double w, const double* dense, long ld,
                    int n, double* x);
int launch_shift(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                 const double* h0, const double* ov, const double* vorb, double* h);
int launch_mulliken(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                    const double* ov, const double* h0, const double* rho, double* qorb,
                    double* eorb, int nOrb);
int launch_pair_forces(hipStream_t s, const SkDev& sk, const PairsDev& pr, const int* norb,
                       const long* iAtomStart, const double* coords, const double* rho,
                       const double* erho, const double* vorb, int mode, double delta,
                       double* grad);
int launch_invr_cluster(hipStream_t s, int nat, const double* coords, double* g);
int launch_ewald_real(hipStream_t s, int nat, const double* coords, int ntr, const double* tr,
                      double alpha, double maxr, double cst, double self, double* g);
int launch_strucfac(hipStream_t s, int nat, const double* coords, int ng, const double* G,
                    double* cs);
int launch_short_gamma(hipStream_t s, const NeighDev& nb, const double* coords,
                       const double* hubbu, const double* gcut, int nsp, int nat, double* g);
int launch_short_gamma_grad(hipStream_t s, const NeighDev& nb, const double* coords,
                            const double* hubbu, const double* gcut, int nsp, const double* dq,
                            double* grad);
int launch_gamma_grad(hipStream_t s, const NeighDev& nb, const double* coords,
                      const double* hubbu, const double* gcut, int nsp, int nat, bool periodic,
                      int ntr, const double* tr, double alpha, double maxr, int ng,
                      const double* G, const double* w, const double* CSq, const double* dq,
                      double* grad);
// Row view of the dense matrix that unpack builds from the sparse blocks: row atom A holds the
// entries (2 k + t, B) of rowPtr[A] .. rowPtr[A + 1] - 1: block k of the pair (at1 = B,
// at2 = A) with t = 0 and, for at1 != at2, block k of the pair (at1 = A, at2 = B) transposed
// (t = 1). Image blocks of one atom pair are separate entries (they are summed by the product).
struct RowBlocksDev {
  int nat;
  const int* rowPtr;
  const int2* ent;    // in block order within a row
  // the same entries with each row sorted by column atom; the entries of one column atom (its
  // image blocks) form a group, grp[g] = (first, one past the last entry of entS), the groups
  // of row A at gPtr[A] .. gPtr[A + 1] - 1
  const int2* entS;
  const int* gPtr;
  const int2* grp;
};
// dense = the matrix of the sparse x (both triangles, image blocks summed; zero elsewhere) from
// the row view: every dense element written once, without atomics
template <typename T>
int launch_unpack_rows(hipStream_t s, const RowBlocksDev& rb, const int* norb,
                       const long* iAtomStart, const long* blkStart, const double* x, T* dense,
                       long ld, int n);
// Dense 4 x 4 blocks of H and S per group of the row view (row atom, column atom: image blocks
// summed, at a time-reversal-invariant k-point kf (fractional; blkT the fractional image
// translations) with their phases +-1), fold[32 g .. 32 g + 31] = H(r, c) at 4 r + c, then S,
// zero past the atoms' orbitals; colOff[g] = first orbital of the column atom. fold holds 32
// doubles and colOff one int per group of rb (gPtr[nat] groups).
int launch_spmm_fold(hipStream_t s, const RowBlocksDev& rb, const int* norb,
                     const long* iAtomStart, const long* blkStart, const double* h,
                     const double* ov, const double* blkT, const double* kf, double* fold,
                     int* colOff);
// YH^T = (H X)^T and YS^T = (S X)^T with Xt = X^T (nvec x n, leading dimension ldx) and the
// outputs nvec x n (leading dimension ldy), H and S from the folded blocks
int launch_spmm_hs_t(hipStream_t s, const RowBlocksDev& rb, const int* norb,
                     const long* iAtomStart, int n, const double* fold, const int* colOff,
                     const double* Xt, long ldx, int nvec, double* yh, double* ys, long ldy);
// vectors per lane of the block-sparse products (1, the default, 2 or 4)
int set_spmm_cols(int vpl);
// Mulliken populations qorb and band-energy contributions eorb (may be null) per orbital from
// the row view (no atomics; every orbital of the rb.nat atoms written)
int launch_mulliken_rows(hipStream_t s, const RowBlocksDev& rb, const int* norb,
                         const long* iAtomStart, const long* blkStart, const double* ov,
                         const double* h0, const double* rho, double* qorb, double* eorb);
int launch_repulsive(hipStream_t s, const NeighDev& nb, const double* coords, const RepDev* reps,
                     int nsp, double* eatom, double* grad);

}  // namespace dftbhip
