// Batched SCC kernels for many independent molecules in one launch (batch.hip).
//
// A batch is one block-diagonal system: atoms and orbitals of all molecules are numbered
// consecutively and the neighbour list holds only intra-molecule pairs, so the sparse kernels
// (H0/S assembly, shift, pack, Mulliken, pair forces, repulsive) run unchanged over the union.
// What is per molecule lives here: the dense matrices (packed one after the other, n_m x n_m
// column-major), the generalized eigensolver (one workgroup per molecule: Cholesky, two
// triangular substitutions, parallel cyclic Jacobi, back-substitution), the density matrices and
// the cluster gamma matrix with its products and gradient.
#include "batch.h"
#include "gamma_dev.h"

namespace dftbhip {

// sparse -> per-molecule dense (both triangles); the dense array is zeroed beforehand
__global__ void batch_unpack_kernel(PairsDev pr, const int* norbAtom, const long* iAtomStart,
                                    BatchLayout L, const double* x, double* dense) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= pr.nblk) return;
  const int a1 = pr.at1[k], a2 = pr.at2[k];
  const int m = L.atomMol[a1];
  const long o0 = L.orb0[m];
  const int n = (int)(L.orb0[m + 1] - o0);
  double* d = dense + L.dOff[m];
  const int n1 = norbAtom[a1], n2 = norbAtom[a2];
  const long o1 = iAtomStart[a1] - o0, o2 = iAtomStart[a2] - o0;
  const long st = pr.start[k];
  for (int c = 0; c < n1; ++c)
    for (int r = 0; r < n2; ++r) {
      const double v = x[st + r + c * n2];
      atomicAdd(&d[(o2 + r) + (o1 + c) * n], v);
      if (a1 != a2) atomicAdd(&d[(o1 + c) + (o2 + r) * n], v);
    }
}

// per-molecule dense (lower triangle read for on-site blocks) -> sparse
__global__ void batch_pack_kernel(PairsDev pr, const int* norbAtom, const long* iAtomStart,
                                  BatchLayout L, const double* dense, double* x) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= pr.nblk) return;
  const int a1 = pr.at1[k], a2 = pr.at2[k];
  const int m = L.atomMol[a1];
  const long o0 = L.orb0[m];
  const int n = (int)(L.orb0[m + 1] - o0);
  const double* d = dense + L.dOff[m];
  const int n1 = norbAtom[a1], n2 = norbAtom[a2];
  const long o1 = iAtomStart[a1] - o0, o2 = iAtomStart[a2] - o0;
  const long st = pr.start[k];
  for (int c = 0; c < n1; ++c)
    for (int r = 0; r < n2; ++r) {
      long rr = r, cc = c;
      if (a1 == a2 && r < c) {
        rr = c;
        cc = r;
      }
      x[st + r + c * n2] = d[(o2 + rr) + (o1 + cc) * n];
    }
}

// ------------------------------------------------------------------------------------------
// Generalized eigensolver H C = S C E per molecule, one 256-thread workgroup each.
// Work: 3 n^2 doubles per molecule at 3 dOff[m] (L, A, V).
// ------------------------------------------------------------------------------------------
constexpr int kJacobiMaxSweeps = 30;
constexpr int kSygvThreads = 1024;

__device__ __forceinline__ void jacobi_pair(int N2, int r, int k, int& p, int& q) {
  p = (k == 0) ? 0 : 1 + (r + k - 1) % (N2 - 1);
  q = 1 + (r + N2 - 2 - k) % (N2 - 1);
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

__global__ __launch_bounds__(kSygvThreads) void batch_sygv_kernel(BatchLayout L, const double* S,
                                                                  const double* H, double* work,
                                                                  double* w, double* C,
                                                                  int* info, const int* active,
                                                                  int warm) {
  constexpr int NT = kSygvThreads;
  if (active && !active[blockIdx.x]) return;   // converged: eigenpairs stay from the last pass
  __shared__ double rc[kBatchMaxOrb / 2], rs[kBatchMaxOrb / 2];
  __shared__ int rp[kBatchMaxOrb / 2], rq[kBatchMaxOrb / 2];
  __shared__ int perm[kBatchMaxOrb];
  __shared__ int flag, nrot;
  const int m = blockIdx.x;
  const int tid = threadIdx.x;
  const long base = L.dOff[m];
  const int n = (int)(L.orb0[m + 1] - L.orb0[m]);
  const double* Sm = S + base;
  const double* Hm = H + base;
  double* Lm = work + 3 * base;    // L in the lower triangle, L^T mirrored into the upper one
  double* Am = Lm + (long)n * n;
  double* Vm = Am + (long)n * n;
  double* Cm = C + base;
  double* wm = w + L.orb0[m];
  const long nn = (long)n * n;
  if (tid == 0) flag = 0;
  for (long e = tid; e < nn; e += NT) Lm[e] = Sm[e];
  __syncthreads();
  // Cholesky S = L L^T (lower, right-looking)
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      const double d = Lm[j + (long)j * n];
      if (!(d > 0.0)) {
        flag = j + 1;
      } else {
        Lm[j + (long)j * n] = sqrt(d);
      }
    }
    __syncthreads();
    if (flag != 0) {
      if (tid == 0) info[m] = flag;
      return;
    }
    const double djj = Lm[j + (long)j * n];
    for (int i = j + 1 + tid; i < n; i += NT) Lm[i + (long)j * n] /= djj;
    __syncthreads();
    const int len = n - j - 1;
    for (long e = tid; e < (long)len * len; e += NT) {
      const int i = j + 1 + (int)(e % len), k = j + 1 + (int)(e / len);
      if (i >= k) Lm[i + (long)k * n] -= Lm[i + (long)j * n] * Lm[k + (long)j * n];
    }
    __syncthreads();
  }
  // L^T into the upper triangle: row i of L is then contiguous (column i of the upper part)
  for (long e = tid; e < nn; e += NT) {
    const int i = (int)(e % n), k = (int)(e / n);
    if (i > k) Lm[k + (long)i * n] = Lm[i + (long)k * n];
  }
  __syncthreads();
  // Y = L^-1 H (into C), one column per thread, row i from L^T's column i
  for (int c = tid; c < n; c += NT)
    for (int i = 0; i < n; ++i) {
      const double acc = Hm[i + (long)c * n] - dot4(Lm + (long)i * n, 1, Cm + (long)c * n, 1, i);
      Cm[i + (long)c * n] = acc / Lm[i + (long)i * n];
    }
  __syncthreads();
  // A = L^-1 Y^T
  for (int c = tid; c < n; c += NT)
    for (int i = 0; i < n; ++i) {
      const double acc = Cm[c + (long)i * n] - dot4(Lm + (long)i * n, 1, Am + (long)c * n, 1, i);
      Am[i + (long)c * n] = acc / Lm[i + (long)i * n];
    }
  __syncthreads();
  if (warm) {
    // warm start: V holds the eigenvectors of the previous pass; A <- V^T A V is nearly
    // diagonal when H changed little. T = A V into C, then A = V^T T.
    for (long e = tid; e < nn; e += NT) {
      const int i = (int)(e % n), j = (int)(e / n);
      Cm[e] = dot4(Am + i, n, Vm + (long)j * n, 1, n);
    }
    __syncthreads();
    for (long e = tid; e < nn; e += NT) {
      const int i = (int)(e % n), j = (int)(e / n);
      Am[e] = dot4(Vm + (long)i * n, 1, Cm + (long)j * n, 1, n);
    }
    __syncthreads();
  }
  for (long e = tid; e < nn; e += NT) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i > j) {
      const double v = 0.5 * (Am[i + (long)j * n] + Am[j + (long)i * n]);
      Am[i + (long)j * n] = v;
      Am[j + (long)i * n] = v;
    }
  }
  if (!warm)
    for (long e = tid; e < nn; e += NT) Vm[e] = ((e % n) == (e / n)) ? 1.0 : 0.0;
  __syncthreads();
  // parallel cyclic Jacobi, round-robin pairing (n odd: one dummy index). The rotation of a
  // round touch disjoint column (row) pairs; each thread updates four elements per pass with
  // all loads ahead of the stores.
  const int N2 = n + (n & 1);
  const int np = N2 / 2;
  const long work4 = ((long)np * n + 3) / 4;
  const double eps = 2.220446049250313e-16;
  for (int sweep = 0; sweep < kJacobiMaxSweeps && n > 1; ++sweep) {
    if (tid == 0) nrot = 0;
    __syncthreads();
    for (int r = 0; r < N2 - 1; ++r) {
      for (int k = tid; k < np; k += NT) {
        int p, q;
        jacobi_pair(N2, r, k, p, q);
        double c = 1.0, s = 0.0;
        if (p < n && q < n) {
          const double apq = Am[p + (long)q * n];
          const double app = Am[p + (long)p * n], aqq = Am[q + (long)q * n];
          if (fabs(apq) > 1e-300 && 100.0 * fabs(apq) > eps * fmin(fabs(app), fabs(aqq))) {
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
      // columns p, q of A and V
      for (long g = tid; g < work4; g += NT) {
        double ap[4], aq[4], vp[4], vq[4];
        long ip[4], iq[4];
        double cc[4], ss[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const long e = g * 4 + u;
          ss[u] = 0.0;
          if (e < (long)np * n) {
            const int k = (int)(e / n), i = (int)(e % n);
            ss[u] = rs[k];
            cc[u] = rc[k];
            ip[u] = i + (long)rp[k] * n;
            iq[u] = i + (long)rq[k] * n;
            if (ss[u] != 0.0) {
              ap[u] = Am[ip[u]];
              aq[u] = Am[iq[u]];
              vp[u] = Vm[ip[u]];
              vq[u] = Vm[iq[u]];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (ss[u] == 0.0) continue;
          Am[ip[u]] = cc[u] * ap[u] - ss[u] * aq[u];
          Am[iq[u]] = ss[u] * ap[u] + cc[u] * aq[u];
          Vm[ip[u]] = cc[u] * vp[u] - ss[u] * vq[u];
          Vm[iq[u]] = ss[u] * vp[u] + cc[u] * vq[u];
        }
      }
      __syncthreads();
      // rows p, q of A
      for (long g = tid; g < work4; g += NT) {
        double ap[4], aq[4];
        long ip[4], iq[4];
        double cc[4], ss[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const long e = g * 4 + u;
          ss[u] = 0.0;
          if (e < (long)np * n) {
            const int k = (int)(e / n), j = (int)(e % n);
            ss[u] = rs[k];
            cc[u] = rc[k];
            ip[u] = rp[k] + (long)j * n;
            iq[u] = rq[k] + (long)j * n;
            if (ss[u] != 0.0) {
              ap[u] = Am[ip[u]];
              aq[u] = Am[iq[u]];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (ss[u] == 0.0) continue;
          Am[ip[u]] = cc[u] * ap[u] - ss[u] * aq[u];
          Am[iq[u]] = ss[u] * ap[u] + cc[u] * aq[u];
        }
      }
      __syncthreads();
    }
    if (nrot == 0) break;
    __syncthreads();
  }
  // ascending eigenvalues (stable ranks), C = L^-T V(:, perm)
  for (int i = tid; i < n; i += NT) {
    const double di = Am[i + (long)i * n];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = Am[j + (long)j * n];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    perm[rank] = i;
    wm[rank] = di;
  }
  __syncthreads();
  for (int c = tid; c < n; c += NT) {
    const int pc = perm[c];
    for (int i = n - 1; i >= 0; --i) {
      const int len = n - 1 - i;
      const double acc = Vm[i + (long)pc * n] -
                         dot4(Lm + (i + 1) + (long)i * n, 1, Cm + (i + 1) + (long)c * n, 1, len);
      Cm[i + (long)c * n] = acc / Lm[i + (long)i * n];
    }
  }
  if (tid == 0) info[m] = 0;
}

// P_m = C_m diag(f) C_m^T over the first nocc[m] columns (lower and upper)
__global__ __launch_bounds__(256) void batch_density_kernel(BatchLayout L, const double* C,
                                                            const double* f, const int* nocc,
                                                            double* P, const int* active) {
  const int m = blockIdx.x;
  if (active && !active[m]) return;
  const long base = L.dOff[m];
  const int n = (int)(L.orb0[m + 1] - L.orb0[m]);
  const double* Cm = C + base;
  const double* fm = f + L.orb0[m];
  double* Pm = P + base;
  const int k = nocc[m];
  for (long e = threadIdx.x; e < (long)n * n; e += 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    double acc = 0.0;
    for (int l = 0; l < k; ++l) acc += Cm[i + (long)l * n] * fm[l] * Cm[j + (long)l * n];
    Pm[e] = acc;
  }
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
  hipLaunchKernelGGL(batch_unpack_kernel, dim3(cdiv(pr.nblk, 128)), dim3(128), 0, s, pr, norb,
                     iAtomStart, L, x, dense);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_batch_pack(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                      const BatchLayout& L, const double* dense, double* x) {
  if (pr.nblk == 0) return 0;
  hipLaunchKernelGGL(batch_pack_kernel, dim3(cdiv(pr.nblk, 128)), dim3(128), 0, s, pr, norb,
                     iAtomStart, L, dense, x);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_batch_sygv(hipStream_t s, const BatchLayout& L, const double* S, const double* H,
                      double* work, double* w, double* C, int* info, const int* active,
                      bool warm) {
  if (L.nmol == 0) return 0;
  hipLaunchKernelGGL(batch_sygv_kernel, dim3(L.nmol), dim3(kSygvThreads), 0, s, L, S, H, work, w,
                     C, info, active, warm ? 1 : 0);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_batch_density(hipStream_t s, const BatchLayout& L, const double* C, const double* f,
                         const int* nocc, double* P, const int* active) {
  if (L.nmol == 0) return 0;
  hipLaunchKernelGGL(batch_density_kernel, dim3(L.nmol), dim3(256), 0, s, L, C, f, nocc, P,
                     active);
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
