// Batched SCC kernels for many independent molecules in one launch (batch.hip).
//
// A batch is one block-diagonal system: atoms and orbitals of all molecules are numbered
// consecutively and the neighbour list holds only intra-molecule pairs, so the sparse kernels
// (H0/S assembly, shift, pack, Mulliken, pair forces, repulsive) run unchanged over the union.
// What is per molecule lives here: the dense matrices (packed one after the other, n_m x n_m
// column-major), the generalized eigensolver (one workgroup per molecule: Cholesky and L^-T once
// per geometry, then per call a congruence with workgroup GEMMs and parallel cyclic Jacobi), the
// density matrices and the cluster gamma matrix with its products and gradient. The eigensolver
// and density kernels visit the molecules largest first (BatchLayout::order).
#include "batch.h"
#include "gamma_dev.h"

namespace dftbhip {

// The two kernels below give each SK block 16 lanes, lane e holding element e = r + c n2 of the
// column-major n2 x n1 block (lanes past n1 n2 idle): a wave's blocks are consecutive in the
// sparse arrays, so their loads and stores are contiguous.
constexpr int kBatchBlkLanes = 16;

// sparse -> per-molecule dense (both triangles); the dense array is zeroed beforehand
__global__ void batch_unpack_kernel(PairsDev pr, const int* norbAtom, const long* iAtomStart,
                                    BatchLayout L, const double* x, double* dense) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int k = (int)(t / kBatchBlkLanes), e = (int)(t % kBatchBlkLanes);
  if (k >= pr.nblk) return;
  const int a1 = pr.at1[k], a2 = pr.at2[k];
  const int n1 = norbAtom[a1], n2 = norbAtom[a2];
  if (e >= n1 * n2) return;
  const int r = e % n2, c = e / n2;
  const int m = L.atomMol[a1];
  const long o0 = L.orb0[m];
  const int n = (int)(L.orb0[m + 1] - o0);
  double* d = dense + L.dOff[m];
  const long o1 = iAtomStart[a1] - o0, o2 = iAtomStart[a2] - o0;
  const double v = x[pr.start[k] + e];
  atomicAdd(&d[(o2 + r) + (o1 + c) * n], v);
  if (a1 != a2) atomicAdd(&d[(o1 + c) + (o2 + r) * n], v);
}

// per-molecule dense (lower triangle read for on-site blocks) -> sparse
__global__ void batch_pack_kernel(PairsDev pr, const int* norbAtom, const long* iAtomStart,
                                  BatchLayout L, const double* dense, double* x) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int k = (int)(t / kBatchBlkLanes), e = (int)(t % kBatchBlkLanes);
  if (k >= pr.nblk) return;
  const int a1 = pr.at1[k], a2 = pr.at2[k];
  const int n1 = norbAtom[a1], n2 = norbAtom[a2];
  if (e >= n1 * n2) return;
  const int r = e % n2, c = e / n2;
  const int m = L.atomMol[a1];
  const long o0 = L.orb0[m];
  const int n = (int)(L.orb0[m + 1] - o0);
  const double* d = dense + L.dOff[m];
  const long o1 = iAtomStart[a1] - o0, o2 = iAtomStart[a2] - o0;
  long rr = r, cc = c;
  if (a1 == a2 && r < c) {
    rr = c;
    cc = r;
  }
  x[pr.start[k] + e] = d[(o2 + rr) + (o1 + cc) * n];
}

// ------------------------------------------------------------------------------------------
// Generalized eigensolver H C = S C E per molecule, one 1024-thread workgroup each.
//
// S is fixed for a geometry, so the reduction to standard form is done once: a cold call
// factors S = L L^T and starts from C = L^-T. Every call forms A = C^T H C with two workgroup
// GEMMs and diagonalizes A by parallel cyclic Jacobi, applying each round's rotations to A from
// both sides (one pass over its 2 x 2 blocks) and to the columns of C. Rotations keep
// C^T S C = I, so the rotated C holds the generalized eigenvectors. A warm call starts from the
// C of the previous call, for which A is already nearly diagonal when H changed little.
// Work: 3 n^2 doubles per molecule at 3 dOff[m]: B0 (L, then A), B1 (H C) and Cw (the C above,
// columns in Jacobi order, kept between calls).
// ------------------------------------------------------------------------------------------
constexpr int kJacobiMaxSweeps = 30;

// Rotation test of a Jacobi pair: the off-diagonal element is above the relative level of the
// smaller diagonal entry AND above the rounding level of the matrix (floor = 4 eps max|a_ii| of
// the matrix the sweeps start from). Without the floor, a pair touching a diagonal entry near 0
// would rotate forever on the rounding noise that other rotations leave in its element.
__device__ inline bool jacobi_rotates(double apq, double app, double aqq, double floor) {
  const double eps = 2.220446049250313e-16;
  const double a = fabs(apq);
  return a > 1e-300 && a > floor && 100.0 * a > eps * fmin(fabs(app), fabs(aqq));
}
constexpr int kSygvThreads = 1024;
constexpr int kGT = 64;   // workgroup GEMM output tile
constexpr int kGK = 16;   // workgroup GEMM k chunk

__device__ __forceinline__ void jacobi_pair(int N2, int r, int k, int& p, int& q) {
  p = (k == 0) ? 0 : 1 + (r + k - 1) % (N2 - 1);
  q = 1 + (r + N2 - 2 - k) % (N2 - 1);
}

// Unit u of the unordered pairs (k1, k2), k1 <= k2, numbered k2 (k2 + 1) / 2 + k1.
__device__ __forceinline__ void unit_pairs(int u, int& k1, int& k2) {
  int b = (int)((sqrtf(8.0f * (float)u + 1.0f) - 1.0f) * 0.5f);
  while ((b + 1) * (b + 2) / 2 <= u) ++b;
  while (b * (b + 1) / 2 > u) --b;
  k2 = b;
  k1 = u - b * (b + 1) / 2;
}

// sum_{k < len} x[k * sx] * y[k * sy] with four independent partial sums
__device__ __forceinline__ double dot4(const double* __restrict__ x, long sx,
                                       const double* __restrict__ y, long sy, int len) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int k = 0;
  for (; k + 4 <= len; k += 4) {
    s0 += x[k * sx] * y[k * sy];
    s1 += x[(k + 1) * sx] * y[(k + 1) * sy];
    s2 += x[(k + 2) * sx] * y[(k + 2) * sy];
    s3 += x[(k + 3) * sx] * y[(k + 3) * sy];
  }
  for (; k < len; ++k) s0 += x[k * sx] * y[k * sy];
  return (s0 + s1) + (s2 + s3);
}

// Applies one Jacobi round's rotations (pairs rp/rq with cosines rc and sines rs, s = 0: none)
// to the columns of the n x n column-major C (leading dimension n, global memory or LDS): thread
// t takes row i = t mod n and the pairs g, g + G, ... with g = t / n, G = threads / n, all of
// its kB loads issued before the stores. Needs n <= the workgroup size.
template <int kB>
__device__ __forceinline__ void rotate_columns(double* C, int n, int np, const int* rp,
                                               const int* rq, const double* rc,
                                               const double* rs) {
  const int G = kSygvThreads / n;
  const int g = threadIdx.x / n, i = threadIdx.x - g * n;
  if (g >= G) return;
  for (int k0 = g; k0 < np; k0 += kB * G) {
    double cp[kB], cq[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int k = k0 + u * G;
      cp[u] = 0.0;
      cq[u] = 0.0;
      if (k < np && rs[k] != 0.0) {
        cp[u] = C[i + rp[k] * n];
        cq[u] = C[i + rq[k] * n];
      }
    }
    // the pair data are read again from LDS below rather than held in registers
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int k = k0 + u * G;
      if (k < np && rs[k] != 0.0) {
        const double c = rc[k], sn = rs[k];
        C[i + rp[k] * n] = c * cp[u] - sn * cq[u];
        C[i + rq[k] * n] = sn * cp[u] + c * cq[u];
      }
    }
  }
}


// Workgroup GEMM on n x n column-major matrices (leading dimension n):
//   Z = op(X)(:, 0:kd) diag(wk) op(Y)(0:kd, :),  op(X) = X or X^T, op(Y) = Y or Y^T,
// wk optional. 1024 threads; 64 x 64 output tiles, each thread rows tx, tx + 32 by columns
// ty, ty + 32; k in chunks of 16 staged in LDS, the next chunk's loads issued before the
// current chunk's products. Z must not overlap X or Y.
template <bool TX, bool TY>
__device__ void wg_gemm(int n, int kd, const double* __restrict__ X, const double* __restrict__ Y,
                        const double* __restrict__ wk, double* __restrict__ Z,
                        double (*Xs)[kGT + 1], double (*Ys)[kGT + 1]) {
  const int tid = threadIdx.x;
  const int tx = tid % 32, ty = tid / 32;
  // the element each thread stages per chunk: op(X)(xr, xk) and op(Y)(yk, yc)
  const int xr = TX ? tid / kGK : tid % kGT, xk = TX ? tid % kGK : tid / kGT;
  const int yk = TY ? tid / kGT : tid % kGK, yc = TY ? tid % kGT : tid / kGK;
  const int nt = (n + kGT - 1) / kGT;
  for (int tile = 0; tile < nt * nt; ++tile) {
    const int r0 = (tile % nt) * kGT, c0 = (tile / nt) * kGT;
    auto ldx = [&](int k0) -> double {
      const int r = r0 + xr, k = k0 + xk;
      if (r >= n || k >= kd) return 0.0;
      const double v = TX ? X[k + (long)r * n] : X[r + (long)k * n];
      return wk ? v * wk[k] : v;
    };
    auto ldy = [&](int k0) -> double {
      const int k = k0 + yk, c = c0 + yc;
      if (c >= n || k >= kd) return 0.0;
      return TY ? Y[c + (long)k * n] : Y[k + (long)c * n];
    };
    double a00 = 0.0, a10 = 0.0, a01 = 0.0, a11 = 0.0;
    double xv = ldx(0), yv = ldy(0);
    for (int k0 = 0; k0 < kd; k0 += kGK) {
      __syncthreads();
      Xs[xk][xr] = xv;
      Ys[yk][yc] = yv;
      __syncthreads();
      if (k0 + kGK < kd) {
        xv = ldx(k0 + kGK);
        yv = ldy(k0 + kGK);
      }
#pragma unroll
      for (int k = 0; k < kGK; ++k) {
        const double x0 = Xs[k][tx], x1 = Xs[k][tx + 32];
        const double y0 = Ys[k][ty], y1 = Ys[k][ty + 32];
        a00 += x0 * y0;
        a10 += x1 * y0;
        a01 += x0 * y1;
        a11 += x1 * y1;
      }
    }
    const int ra = r0 + tx, rb = ra + 32, ca = c0 + ty, cb = ca + 32;
    if (ra < n && ca < n) Z[ra + (long)ca * n] = a00;
    if (rb < n && ca < n) Z[rb + (long)ca * n] = a10;
    if (ra < n && cb < n) Z[ra + (long)cb * n] = a01;
    if (rb < n && cb < n) Z[rb + (long)cb * n] = a11;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kSygvThreads) void batch_sygv_kernel(BatchLayout L, const int* list,
                                                                  const double* S,
                                                                  const double* H, double* work,
                                                                  double* w, double* C,
                                                                  int* info, const int* active,
                                                                  int warm) {
  constexpr int NT = kSygvThreads;
  const int m = list[blockIdx.x];
  if (active && !active[m]) return;   // converged: eigenpairs stay from the last pass
  __shared__ double Xs[kGK][kGT + 1], Ys[kGK][kGT + 1];
  __shared__ double rc[kBatchMaxOrb / 2], rs[kBatchMaxOrb / 2];
  __shared__ int rp[kBatchMaxOrb / 2], rq[kBatchMaxOrb / 2];
  __shared__ int perm[kBatchMaxOrb];
  __shared__ int flag, nrot, need;
  __shared__ unsigned long long dmaxBits;
  const int tid = threadIdx.x;
  const long base = L.dOff[m];
  const int n = (int)(L.orb0[m + 1] - L.orb0[m]);
  const long nn = (long)n * n;
  const double* Hm = H + base;
  double* B0 = work + 3 * base;
  double* B1 = B0 + nn;
  double* Cw = B1 + nn;
  double* Cm = C + base;
  double* wm = w + L.orb0[m];
  if (!warm) {
    const double* Sm = S + base;
    if (tid == 0) flag = 0;
    for (long e = tid; e < nn; e += NT) B0[e] = Sm[e];
    __syncthreads();
    // Cholesky S = L L^T (lower, right-looking)
    for (int j = 0; j < n; ++j) {
      if (tid == 0) {
        const double d = B0[j + (long)j * n];
        if (!(d > 0.0)) {
          flag = j + 1;
        } else {
          B0[j + (long)j * n] = sqrt(d);
        }
      }
      __syncthreads();
      if (flag != 0) {
        if (tid == 0) info[m] = flag;
        return;
      }
      const double djj = B0[j + (long)j * n];
      for (int i = j + 1 + tid; i < n; i += NT) B0[i + (long)j * n] /= djj;
      __syncthreads();
      const int len = n - j - 1;
      for (long e = tid; e < (long)len * len; e += NT) {
        const int i = j + 1 + (int)(e % len), k = j + 1 + (int)(e / len);
        if (i >= k) B0[i + (long)k * n] -= B0[i + (long)j * n] * B0[k + (long)j * n];
      }
      __syncthreads();
    }
    // L^T into the upper triangle: row i of L is then contiguous (column i of the upper part)
    for (long e = tid; e < nn; e += NT) {
      const int i = (int)(e % n), k = (int)(e / n);
      if (i > k) B0[k + (long)i * n] = B0[i + (long)k * n];
    }
    __syncthreads();
    // Cw = L^-T: thread c forms column c of L^-1 by forward substitution and stores it as row c
    // of Cw, so that the lanes of a wave read and write consecutive addresses
    for (int c = tid; c < n; c += NT)
      for (int i = 0; i < n; ++i) {
        double v = 0.0;
        if (i >= c)
          v = ((i == c ? 1.0 : 0.0) - dot4(B0 + (long)i * n, 1, Cw + c, n, i)) /
              B0[i + (long)i * n];
        Cw[c + (long)i * n] = v;
      }
    __syncthreads();
  }
  // A = Cw^T H Cw (symmetrized)
  wg_gemm<false, false>(n, n, Hm, Cw, nullptr, B1, Xs, Ys);
  wg_gemm<true, false>(n, n, Cw, B1, nullptr, B0, Xs, Ys);
  double* A = B0;
  for (long e = tid; e < nn; e += NT) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i > j) {
      const double v = 0.5 * (A[i + (long)j * n] + A[j + (long)i * n]);
      A[i + (long)j * n] = v;
      A[j + (long)i * n] = v;
    }
  }
  __syncthreads();
  // parallel cyclic Jacobi, round-robin pairing (n odd: index n is a dummy that never rotates)
  const int N2 = n + (n & 1);
  const int np = N2 / 2;
  const int nunit = np * (np + 1) / 2;
  constexpr int kAU = 4;   // A-update units per thread and pass
  const int npair = n * (n - 1) / 2;
  // rounding level of the matrix: 4 eps max |a_ii| (bits of non-negative doubles order as
  // unsigned integers, so the maximum is an integer atomic)
  if (tid == 0) dmaxBits = 0ull;
  __syncthreads();
  {
    double dm = 0.0;
    for (int i = tid; i < n; i += NT) dm = fmax(dm, fabs(A[i + (long)i * n]));
    atomicMax(&dmaxBits, (unsigned long long)__double_as_longlong(dm));
  }
  __syncthreads();
  const double rotFloor = 4.0 * 2.220446049250313e-16 * __longlong_as_double((long long)dmaxBits);
  bool done = n <= 1;
  for (int sweep = 0; sweep < kJacobiMaxSweeps && n > 1; ++sweep) {
    // a sweep in which no pair passes the rotation test changes nothing: test all pairs first
    if (tid == 0) {
      nrot = 0;
      need = 0;
    }
    __syncthreads();
    for (int u = tid; u < npair; u += NT) {
      int k1, k2;
      unit_pairs(u, k1, k2);
      const int p = k1, q = k2 + 1;
      const double apq = A[p + (long)q * n];
      const double app = A[p + (long)p * n], aqq = A[q + (long)q * n];
      if (jacobi_rotates(apq, app, aqq, rotFloor)) need = 1;
    }
    __syncthreads();
    if (need == 0) {
      done = true;
      break;
    }
    for (int r = 0; r < N2 - 1; ++r) {
      for (int k = tid; k < np; k += NT) {
        int p, q;
        jacobi_pair(N2, r, k, p, q);
        double c = 1.0, s = 0.0;
        if (p < n && q < n) {
          const double apq = A[p + (long)q * n];
          const double app = A[p + (long)p * n], aqq = A[q + (long)q * n];
          if (jacobi_rotates(apq, app, aqq, rotFloor)) {
            const double th = (aqq - app) / (2.0 * apq);
            const double t = copysign(1.0, th) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
            atomicAdd(&nrot, 1);
          }
        }
        rp[k] = p;
        rq[k] = q;
        rc[k] = c;
        rs[k] = s;
      }
      __syncthreads();
      // A <- J^T A J over the unordered 2 x 2 blocks (rows of pair k1, columns of pair k2,
      // k1 <= k2; the block and its transpose take the same values), kAU per thread with all
      // loads ahead of the stores; only the loaded values stay in registers in between
      for (int u0 = tid; u0 < nunit; u0 += kAU * NT) {
        double x[kAU][4];
#pragma unroll
        for (int a = 0; a < kAU; ++a) {
          const int u = u0 + a * NT;
          if (u >= nunit) continue;
          int k1, k2;
          unit_pairs(u, k1, k2);
          if (rs[k1] == 0.0 && rs[k2] == 0.0) continue;
          const int row[2] = {rp[k1], rq[k1]}, col[2] = {rp[k2], rq[k2]};
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int ri = row[e & 1], ci = col[e >> 1];
            x[a][e] = ri < n && ci < n ? A[ri + ci * n] : 0.0;   // n <= kBatchMaxOrb
          }
        }
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int a = 0; a < kAU; ++a) {
          const int u = u0 + a * NT;
          if (u >= nunit) continue;
          int k1, k2;
          unit_pairs(u, k1, k2);
          const double c1 = rc[k1], s1 = rs[k1], c2 = rc[k2], s2 = rs[k2];
          if (s1 == 0.0 && s2 == 0.0) continue;
          // x[e]: e = 0 (p1, p2), 1 (q1, p2), 2 (p1, q2), 3 (q1, q2); columns first
          const double y0 = c2 * x[a][0] - s2 * x[a][2];
          const double y1 = c2 * x[a][1] - s2 * x[a][3];
          const double y2 = s2 * x[a][0] + c2 * x[a][2];
          const double y3 = s2 * x[a][1] + c2 * x[a][3];
          double z[4] = {c1 * y0 - s1 * y1, s1 * y0 + c1 * y1, c1 * y2 - s1 * y3,
                         s1 * y2 + c1 * y3};
          if (k1 == k2) z[1] = z[2];   // (q, p) and (p, q) alike
          const int row[2] = {rp[k1], rq[k1]}, col[2] = {rp[k2], rq[k2]};
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int ri = row[e & 1], ci = col[e >> 1];
            if (ri < n && ci < n) {
              A[ri + ci * n] = z[e];
              if (k1 != k2) A[ci + ri * n] = z[e];
            }
          }
        }
      }
      rotate_columns<8>(Cw, n, np, rp, rq, rc, rs);
      __syncthreads();
    }
    if (nrot == 0) {
      done = true;
      break;
    }
    __syncthreads();
  }
  // ascending eigenvalues (stable ranks), C(:, rank) = Cw(:, i)
  for (int i = tid; i < n; i += NT) {
    const double di = A[i + (long)i * n];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = A[j + (long)j * n];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    perm[rank] = i;
    wm[rank] = di;
  }
  __syncthreads();
  for (long e = tid; e < nn; e += NT) {
    const int i = (int)(e % n), c = (int)(e / n);
    Cm[e] = Cw[i + (long)perm[c] * n];
  }
  // a molecule whose last allowed sweep still rotated is reported (info > 0), not taken as
  // converged
  if (tid == 0) info[m] = done ? 0 : kBatchJacobiNoConv;
}

// The same eigensolver for molecules of at most kMaxN orbitals with A resident in LDS (upper
// triangle packed by columns: (i, j), i <= j, at i + j(j+1)/2) and, for kCLds, C as well: a
// Jacobi round then waits on LDS alone. Otherwise C stays in HBM and each round ends with a
// full barrier. The LDS block of A doubles as the GEMM staging before the sweeps.
__device__ __forceinline__ int sym_idx(int i, int j) {
  return i <= j ? i + j * (j + 1) / 2 : j + i * (i + 1) / 2;
}

// The smallest class (49 KB of LDS per workgroup) is held to 64 VGPRs so that two 1024-thread
// workgroups share a CU; the larger classes fill a CU's LDS with one.
template <int kMaxN, bool kCLds>
__global__ __launch_bounds__(kSygvThreads)
__attribute__((amdgpu_waves_per_eu(kMaxN <= 64 ? 8 : 1))) void batch_sygv_lds_kernel(
    BatchLayout L, const int* list, const double* S, const double* H, double* work, double* w,
    double* C, int* info, const int* active, int warm) {
  constexpr int NT = kSygvThreads;
  constexpr int kPacked = kMaxN * (kMaxN + 1) / 2;
  constexpr int kStage = 2 * kGK * (kGT + 1);
  __shared__ double As[kPacked > kStage ? kPacked : kStage];
  __shared__ double Cs[kCLds ? kMaxN * kMaxN : 1];
  __shared__ double rc[kMaxN / 2 + 1], rs[kMaxN / 2 + 1];
  __shared__ int rp[kMaxN / 2 + 1], rq[kMaxN / 2 + 1];
  __shared__ int perm[kMaxN];
  __shared__ int flag, nrot, need;
  __shared__ unsigned long long dmaxBits;
  const int m = list[blockIdx.x];
  if (active && !active[m]) return;   // converged: eigenpairs stay from the last pass
  const int tid = threadIdx.x;
  const long base = L.dOff[m];
  const int n = (int)(L.orb0[m + 1] - L.orb0[m]);
  if (n > kMaxN) {   // a molecule in the wrong size class: reported, nothing touched
    if (tid == 0) info[m] = -n;
    return;
  }
  const long nn = (long)n * n;
  double* B0 = work + 3 * base;
  double* B1 = B0 + nn;
  double* Cw = B1 + nn;
  if (!warm) {
    if (tid == 0) flag = 0;
    for (long e = tid; e < nn; e += NT) B0[e] = S[base + e];
    __syncthreads();
    for (int j = 0; j < n; ++j) {
      if (tid == 0) {
        const double d = B0[j + (long)j * n];
        if (!(d > 0.0)) {
          flag = j + 1;
        } else {
          B0[j + (long)j * n] = sqrt(d);
        }
      }
      __syncthreads();
      if (flag != 0) {
        if (tid == 0) info[m] = flag;
        return;
      }
      const double djj = B0[j + (long)j * n];
      for (int i = j + 1 + tid; i < n; i += NT) B0[i + (long)j * n] /= djj;
      __syncthreads();
      const int len = n - j - 1;
      for (long e = tid; e < (long)len * len; e += NT) {
        const int i = j + 1 + (int)(e % len), k = j + 1 + (int)(e / len);
        if (i >= k) B0[i + (long)k * n] -= B0[i + (long)j * n] * B0[k + (long)j * n];
      }
      __syncthreads();
    }
    for (long e = tid; e < nn; e += NT) {
      const int i = (int)(e % n), k = (int)(e / n);
      if (i > k) B0[k + (long)i * n] = B0[i + (long)k * n];
    }
    __syncthreads();
    for (int c = tid; c < n; c += NT)
      for (int i = 0; i < n; ++i) {
        double v = 0.0;
        if (i >= c)
          v = ((i == c ? 1.0 : 0.0) - dot4(B0 + (long)i * n, 1, Cw + c, n, i)) /
              B0[i + (long)i * n];
        Cw[c + (long)i * n] = v;
      }
    __syncthreads();
  }
  double (*Xs)[kGT + 1] = reinterpret_cast<double (*)[kGT + 1]>(As);
  double (*Ys)[kGT + 1] = reinterpret_cast<double (*)[kGT + 1]>(As + kGK * (kGT + 1));
  wg_gemm<false, false>(n, n, H + base, Cw, nullptr, B1, Xs, Ys);
  wg_gemm<true, false>(n, n, Cw, B1, nullptr, B0, Xs, Ys);
  // A (symmetrized) into LDS
  for (long e = tid; e < nn; e += NT) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i <= j) As[i + j * (j + 1) / 2] = 0.5 * (B0[i + (long)j * n] + B0[j + (long)i * n]);
    if (kCLds) Cs[e] = Cw[e];
  }
  __syncthreads();
  double* Cr = kCLds ? Cs : Cw;   // the C the sweeps rotate
  auto lds_barrier = [&]() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0); vmcnt, expcnt untouched (gfx9 encoding)
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  };
  const int N2 = n + (n & 1);
  const int np = N2 / 2;
  const int nunit = np * (np + 1) / 2;
  const int npair = n * (n - 1) / 2;
  // rounding level of the matrix: 4 eps max |a_ii| (bits of non-negative doubles order as
  // unsigned integers, so the maximum is an integer atomic)
  if (tid == 0) dmaxBits = 0ull;
  __syncthreads();
  {
    double dm = 0.0;
    for (int i = tid; i < n; i += NT) dm = fmax(dm, fabs(As[sym_idx(i, i)]));
    atomicMax(&dmaxBits, (unsigned long long)__double_as_longlong(dm));
  }
  __syncthreads();
  const double rotFloor = 4.0 * 2.220446049250313e-16 * __longlong_as_double((long long)dmaxBits);
  bool done = n <= 1;
  for (int sweep = 0; sweep < kJacobiMaxSweeps && n > 1; ++sweep) {
    // a sweep in which no pair passes the rotation test changes nothing: test all pairs first
    if (tid == 0) {
      nrot = 0;
      need = 0;
    }
    lds_barrier();
    for (int u = tid; u < npair; u += NT) {
      int k1, k2;
      unit_pairs(u, k1, k2);
      const int p = k1, q = k2 + 1;
      const double apq = As[sym_idx(p, q)];
      const double app = As[sym_idx(p, p)], aqq = As[sym_idx(q, q)];
      if (jacobi_rotates(apq, app, aqq, rotFloor)) need = 1;
    }
    lds_barrier();
    if (need == 0) {
      done = true;
      break;
    }
    for (int r = 0; r < N2 - 1; ++r) {
      for (int k = tid; k < np; k += NT) {
        int p, q;
        jacobi_pair(N2, r, k, p, q);
        double c = 1.0, s = 0.0;
        if (p < n && q < n) {
          const double apq = As[sym_idx(p, q)];
          const double app = As[sym_idx(p, p)], aqq = As[sym_idx(q, q)];
          if (jacobi_rotates(apq, app, aqq, rotFloor)) {
            const double th = (aqq - app) / (2.0 * apq);
            const double t = copysign(1.0, th) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
            atomicAdd(&nrot, 1);
          }
        }
        rp[k] = p;
        rq[k] = q;
        rc[k] = c;
        rs[k] = s;
      }
      lds_barrier();
      // A <- J^T A J over the unordered 2 x 2 blocks (pairs k1 <= k2)
      for (int u = tid; u < nunit; u += NT) {
        int k1, k2;
        unit_pairs(u, k1, k2);
        const double s1 = rs[k1], s2 = rs[k2];
        if (s1 == 0.0 && s2 == 0.0) continue;
        const double c1 = rc[k1], c2 = rc[k2];
        const int row[2] = {rp[k1], rq[k1]}, col[2] = {rp[k2], rq[k2]};
        double x[4];
        int o[4];
        bool v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int ri = row[e & 1], ci = col[e >> 1];
          v[e] = ri < n && ci < n;
          o[e] = v[e] ? sym_idx(ri, ci) : 0;
          x[e] = v[e] ? As[o[e]] : 0.0;
        }
        // e = 0 (p1, p2), 1 (q1, p2), 2 (p1, q2), 3 (q1, q2); columns first
        const double y0 = c2 * x[0] - s2 * x[2];
        const double y1 = c2 * x[1] - s2 * x[3];
        const double y2 = s2 * x[0] + c2 * x[2];
        const double y3 = s2 * x[1] + c2 * x[3];
        const double z[4] = {c1 * y0 - s1 * y1, s1 * y0 + c1 * y1, c1 * y2 - s1 * y3,
                             s1 * y2 + c1 * y3};
        if (k1 == k2) {
          // (p, p), (q, q) and the shared off-diagonal element
          if (v[0]) As[o[0]] = z[0];
          if (v[3]) As[o[3]] = z[3];
          if (v[2]) As[o[2]] = z[2];
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (v[e]) As[o[e]] = z[e];
        }
      }
      // columns of C: LDS-resident C needs only the LDS barrier below; C in HBM is shared
      // between the threads of a row across rounds, so its stores complete first
      if (kCLds) {
        rotate_columns<8>(Cr, n, np, rp, rq, rc, rs);
      } else {
        rotate_columns<10>(Cr, n, np, rp, rq, rc, rs);
        __syncthreads();
      }
      lds_barrier();
    }
    const int nr = nrot;
    lds_barrier();   // every thread has read nrot before the next sweep clears it
    if (nr == 0) {
      done = true;
      break;
    }
  }
  __syncthreads();   // every column of C written before other threads read it
  double* wm = w + L.orb0[m];
  for (int i = tid; i < n; i += NT) {
    const double di = As[sym_idx(i, i)];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = As[sym_idx(j, j)];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    perm[rank] = i;
    wm[rank] = di;
  }
  __syncthreads();
  double* Cm = C + base;
  for (long e = tid; e < nn; e += NT) {
    const int i = (int)(e % n), c = (int)(e / n);
    Cm[e] = Cr[i + (long)perm[c] * n];
    if (kCLds) Cw[e] = Cs[e];   // the next warm call starts from the rotated C
  }
  // a molecule whose last allowed sweep still rotated is reported (info > 0), not taken as
  // converged
  if (tid == 0) info[m] = done ? 0 : kBatchJacobiNoConv;
}

// P_m = C_m diag(f) C_m^T over the first nocc[m] columns (lower and upper)
__global__ __launch_bounds__(kSygvThreads) void batch_density_kernel(BatchLayout L,
                                                                     const double* C,
                                                                     const double* f,
                                                                     const int* nocc, double* P,
                                                                     const int* active) {
  const int m = L.order[blockIdx.x];
  if (active && !active[m]) return;
  __shared__ double Xs[kGK][kGT + 1], Ys[kGK][kGT + 1];
  const long base = L.dOff[m];
  const int n = (int)(L.orb0[m + 1] - L.orb0[m]);
  wg_gemm<false, true>(n, nocc[m], C + base, C + base, f + L.orb0[m], P + base, Xs, Ys);
}

// gamma_m: 1/r off the diagonal minus the short-range expGamma part within its cutoff (on the
// diagonal the r -> 0 limit), the same sum the single-system path builds from the half list
__global__ __launch_bounds__(256) void batch_gamma_kernel(BatchLayout L, const double* coords,
                                                          const int* species,
                                                          const double* hubbu,
                                                          const double* gcut, int nsp,
                                                          double* g) {
  const int m = blockIdx.x;
  const int a0 = L.atom0[m];
  const int na = L.atom0[m + 1] - a0;
  double* gm = g + L.gOff[m];
  for (int e = threadIdx.x; e < na * na; e += 256) {
    const int i = e % na, j = e / na;
    const int ai = a0 + i, aj = a0 + j;
    const double dx = coords[3 * ai] - coords[3 * aj], dy = coords[3 * ai + 1] - coords[3 * aj + 1],
                 dz = coords[3 * ai + 2] - coords[3 * aj + 2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    double v = (i != j) ? 1.0 / r : 0.0;
    const int hi = ai > aj ? ai : aj, lo = ai > aj ? aj : ai;
    const int sh = species[hi], sl = species[lo];
    if (i == j || r <= gcut[sh * nsp + sl]) v += -expgamma(r, hubbu[sh], hubbu[sl]);
    gm[e] = v;
  }
}

// v_a = sum_b gamma(b, a) dq_b within the molecule of a
__global__ void batch_gemv_kernel(BatchLayout L, int nat, const double* g, const double* dq,
                                  double* v) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nat) return;
  const int m = L.atomMol[a];
  const int a0 = L.atom0[m];
  const int na = L.atom0[m + 1] - a0;
  const double* col = g + L.gOff[m] + (long)(a - a0) * na;
  double acc = 0.0;
  for (int b = 0; b < na; ++b) acc += col[b] * dq[a0 + b];
  v[a] = acc;
}

// cluster 1/R gradient within each molecule
__global__ void batch_invr_grad_kernel(BatchLayout L, int nat, const double* coords,
                                       const double* dq, double* grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nat) return;
  const int m = L.atomMol[i];
  const int a0 = L.atom0[m], a1 = L.atom0[m + 1];
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;
  for (int j = a0; j < a1; ++j) {
    if (j == i) continue;
    const double d0 = coords[3 * i] - coords[3 * j], d1 = coords[3 * i + 1] - coords[3 * j + 1],
                 d2 = coords[3 * i + 2] - coords[3 * j + 2];
    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    const double f = -dq[i] * dq[j] / (r2 * sqrt(r2));
    g0 += f * d0;
    g1 += f * d1;
    g2 += f * d2;
  }
  grad[3 * i] += g0;
  grad[3 * i + 1] += g1;
  grad[3 * i + 2] += g2;
}

int launch_batch_unpack(hipStream_t s, const PairsDev& pr, const int* norb,
                        const long* iAtomStart, const BatchLayout& L, const double* x,
                        double* dense, long denseSize) {
  DH_TRY(fill<double>(s, dense, denseSize, 0.0));
  if (pr.nblk == 0) return 0;
  hipLaunchKernelGGL(batch_unpack_kernel, dim3(cdiv((long)pr.nblk * kBatchBlkLanes, 256)),
                     dim3(256), 0, s, pr, norb,
                     iAtomStart, L, x, dense);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_batch_pack(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                      const BatchLayout& L, const double* dense, double* x) {
  if (pr.nblk == 0) return 0;
  hipLaunchKernelGGL(batch_pack_kernel, dim3(cdiv((long)pr.nblk * kBatchBlkLanes, 256)),
                     dim3(256), 0, s, pr, norb,
                     iAtomStart, L, dense, x);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_batch_sygv(const hipStream_t* st, const BatchLayout& L, const double* S,
                      const double* H, double* work, double* w, double* C, int* info,
                      const int* active, bool warm) {
  DH_CHECK(L.ncls[0] + L.ncls[1] + L.ncls[2] + L.ncls[3] == L.nmol,
           "batch_sygv: size classes do not cover the batch");
  const int wm = warm ? 1 : 0;
  const int* list = L.order;   // molecules by decreasing size: the classes in turn
  if (L.ncls[0] > 0)
    hipLaunchKernelGGL(batch_sygv_kernel, dim3(L.ncls[0]), dim3(kSygvThreads), 0, st[0], L, list,
                       S, H, work, w, C, info, active, wm);
  list += L.ncls[0];
  if (L.ncls[1] > 0)
    hipLaunchKernelGGL((batch_sygv_lds_kernel<kBatchLdsMid, false>), dim3(L.ncls[1]),
                       dim3(kSygvThreads), 0, st[1], L, list, S, H, work, w, C, info, active, wm);
  list += L.ncls[1];
  if (L.ncls[2] > 0)
    hipLaunchKernelGGL((batch_sygv_lds_kernel<kBatchLdsC, true>), dim3(L.ncls[2]),
                       dim3(kSygvThreads), 0, st[2], L, list, S, H, work, w, C, info, active, wm);
  list += L.ncls[2];
  if (L.ncls[3] > 0)
    hipLaunchKernelGGL((batch_sygv_lds_kernel<kBatchLdsSmall, true>), dim3(L.ncls[3]),
                       dim3(kSygvThreads), 0, st[3], L, list, S, H, work, w, C, info, active, wm);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_batch_density(hipStream_t s, const BatchLayout& L, const double* C, const double* f,
                         const int* nocc, double* P, const int* active) {
  if (L.nmol == 0) return 0;
  hipLaunchKernelGGL(batch_density_kernel, dim3(L.nmol), dim3(kSygvThreads), 0, s, L, C, f, nocc,
                     P, active);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_batch_gamma(hipStream_t s, const BatchLayout& L, const double* coords,
                       const int* species, const double* hubbu, const double* gcut, int nsp,
                       double* g) {
  if (L.nmol == 0) return 0;
  hipLaunchKernelGGL(batch_gamma_kernel, dim3(L.nmol), dim3(256), 0, s, L, coords, species, hubbu,
                     gcut, nsp, g);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_batch_gemv(hipStream_t s, const BatchLayout& L, int nat, const double* g,
                      const double* dq, double* v) {
  if (nat == 0) return 0;
  hipLaunchKernelGGL(batch_gemv_kernel, dim3(cdiv(nat, 256)), dim3(256), 0, s, L, nat, g, dq, v);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_batch_invr_grad(hipStream_t s, const BatchLayout& L, int nat, const double* coords,
                           const double* dq, double* grad) {
  if (nat == 0) return 0;
  hipLaunchKernelGGL(batch_invr_grad_kernel, dim3(cdiv(nat, 256)), dim3(256), 0, s, L, nat,
                     coords, dq, grad);
  DH_LAUNCH_CHECK();
  return 0;
}

}  // namespace dftbhip
