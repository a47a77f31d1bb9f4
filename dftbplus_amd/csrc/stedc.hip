// Symmetric tridiagonal eigensolver by divide and conquer (Cuppen's tearing with Gu-Eisenstat
// eigenvectors), producing all eigenvalues and eigenvectors.
//
//  * The tridiagonal is torn recursively into halves down to <= 64 rows; each tear subtracts |e|
//    from the two neighbouring diagonal entries and is undone by a rank-one merge
//    D + rho z z^T, rho = 2|e| after normalising z.
//  * Leaves: implicit QL with eigenvector accumulation in LDS, one workgroup per leaf.
//  * Merges (all merges of one tree level at once):
//      - z is gathered from the children's boundary rows on the device;
//      - deflation (small |z| or close eigenvalues, LAPACK dlaed2 rules) is sequential
//        bookkeeping done on the host for the whole level;
//      - secular roots on the device, one thread per root: bisection over the IEEE bit patterns
//        of tau = lambda - d_origin, where the origin is the pole nearer to the root, so
//        d_i - lambda = (d_i - d_origin) - tau is formed without cancellation;
//      - z-hat from the roots (Loewner formula), eigenvectors u_j = zhat / (d - lambda_j);
//      - Q_new = Q_old(:, nondeflated) * U as two batched MFMA GEMMs (top rows use the columns
//        living in the first child plus the mixed ones, bottom rows the second child's).
#include <algorithm>
#include <cmath>

#include <memory>

#include "common.h"

namespace dftbhip {

constexpr int kLeaf = 64;

struct MergeDev {
  int s, n1, n2, k, k1, k2, k3, nrot, ndefl, pad;
  double rho, sigma;
  long dlOff;    // offset into the per-root arrays (DL, W, roots...)
  long qgOff;    // offset of the gathered columns (n x k)
  long uOff;     // offset of U (k x k)
  long rotOff;   // offset into rotation arrays
  long deflOff;  // offset into deflated column list
};

// Device buffers of one divide and conquer, kept by the caller between calls: they only grow,
// so repeated solves of one size allocate (and free, which synchronises the device) nothing.
struct StedcWs {
  DevBuf<double> dd, de, dD, dz, dDL, dW, dtau, dlam, dzhat, dinv, drc, drs;
  DevBuf<int> dls, dlz, dinfo, dorg, drow, dgcol, ddcol, drp, drn, dord;
  DevBuf<MergeDev> dmd;
  DevBuf<GemmDesc<double>> dgd;
};

StedcWs* stedc_ws_new() { return new StedcWs(); }
void stedc_ws_delete(StedcWs* w) { delete w; }


// ------------------------------------------------------------------------------------------
// Leaves: implicit QL (tql2 scheme) with eigenvectors, eigenvalues sorted ascending.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void leaf_kernel(const double* dIn, const double* eIn,
                                                  const int* lstart, const int* lsize, double* Q,
                                                  long ldq, double* D, int* info) {
  __shared__ double z[kLeaf][kLeaf + 1];   // z[row][col]
  __shared__ double d[kLeaf], e[kLeaf];
  __shared__ double rc[kLeaf], rs[kLeaf];
  __shared__ int sh_m, sh_lo, sh_stop, sh_fail;
  __shared__ int perm[kLeaf];
  const int s = lstart[blockIdx.x], m = lsize[blockIdx.x];
  const int t = threadIdx.x;
  if (t < m) {
    d[t] = dIn[s + t];
    e[t] = (t < m - 1) ? eIn[s + t] : 0.0;
    for (int c = 0; c < m; ++c) z[t][c] = (c == t) ? 1.0 : 0.0;
  }
  if (t == 0) sh_fail = 0;
  __syncthreads();
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < m; ++l) {
    int iter = 0;
    while (true) {
      if (t == 0) {
        int mm;
        for (mm = l; mm < m - 1; ++mm) {
          const double dd = fabs(d[mm]) + fabs(d[mm + 1]);
          if (fabs(e[mm]) <= eps * dd) break;
        }
        sh_m = mm;
        if (mm != l) {
          if (++iter > 60) {
            sh_fail = 1;
            sh_m = l;
          }
        }
        if (sh_m != l) {
          double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
          double r = hypot(g, 1.0);
          g = d[mm] - d[l] + e[l] / (g + copysign(r, g));
          double sn = 1.0, cs = 1.0, p = 0.0;
          int i;
          bool under = false;
          for (i = mm - 1; i >= l; --i) {
            const double f = sn * e[i], b = cs * e[i];
            r = hypot(f, g);
            e[i + 1] = r;
            if (r == 0.0) {
              d[i + 1] -= p;
              e[mm] = 0.0;
              under = true;
              break;
            }
            sn = f / r;
            cs = g / r;
            g = d[i + 1] - p;
            r = (d[i] - g) * sn + 2.0 * cs * b;
            p = sn * r;
            d[i + 1] = g + p;
            g = cs * r - b;
            rc[i] = cs;
            rs[i] = sn;
          }
          sh_lo = under ? i + 1 : l;
          if (!under) {
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
          }
        }
      }
      __syncthreads();
      const int mm = sh_m;
      if (mm == l) break;
      const int lo = sh_lo;
      if (t < m) {
        for (int i = mm - 1; i >= lo; --i) {
          const double f = z[t][i + 1];
          z[t][i + 1] = rs[i] * z[t][i] + rc[i] * f;
          z[t][i] = rc[i] * z[t][i] - rs[i] * f;
        }
      }
      __syncthreads();
    }
  }
  if (t == 0) {
    for (int i = 0; i < m; ++i) perm[i] = i;
    for (int i = 0; i < m; ++i) {
      int b = i;
      for (int j = i + 1; j < m; ++j)
        if (d[perm[j]] < d[perm[b]]) b = j;
      const int tmp = perm[i];
      perm[i] = perm[b];
      perm[b] = tmp;
    }
    if (sh_fail) atomicMax(info, s + 1);
  }
  __syncthreads();
  if (t < m) {
    D[s + t] = d[perm[t]];
    for (int c = 0; c < m; ++c) Q[(s + t) + (long)(s + c) * ldq] = z[t][perm[c]];
  }
}

// z for each merge: last row of child 1 (child-1 columns), sigma * first row of child 2.
__global__ void zgather_kernel(const MergeDev* md, const double* Q, long ldq, double* zbuf) {
  const MergeDev m = md[blockIdx.y];
  const int n = m.n1 + m.n2;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const int col = m.s + c;
    zbuf[col] = (c < m.n1) ? Q[(m.s + m.n1 - 1) + (long)col * ldq]
                           : m.sigma * Q[(m.s + m.n1) + (long)col * ldq];
  }
}

// Deflation rotations (in order) on the full rows of the merged block: one thread per row.
__global__ void rotate_kernel(const MergeDev* md, const int* rp, const int* rn, const double* rcs,
                              const double* rsn, double* Q, long ldq) {
  const MergeDev m = md[blockIdx.y];
  const int n = m.n1 + m.n2;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || m.nrot == 0) return;
  const long row = m.s + r;
  for (int q = 0; q < m.nrot; ++q) {
    const long o = m.rotOff + q;
    const double c = rcs[o], s = rsn[o];
    double* x = Q + row + (long)rp[o] * ldq;
    double* y = Q + row + (long)rn[o] * ldq;
    const double xv = *x, yv = *y;
    *x = c * xv + s * yv;
    *y = c * yv - s * xv;
  }
}

// Sums over the k terms of a merge are split over a group of kSecG lanes (lane sub takes
// i = sub, sub + kSecG, ...) and combined by an xor butterfly, which leaves the same total in
// every lane of the group (IEEE addition is commutative), so the group's decisions agree.
constexpr int kSecG = 16;

__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int s = kSecG / 2; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
  return v;
}

__device__ __forceinline__ double secular_f_group(const double* DL, const double* W, int k,
                                                  double rho, double dorg, double tau, int sub) {
  double acc = 0.0;
  for (int i = sub; i < k; i += kSecG) {
    const double den = (DL[i] - dorg) - tau;
    acc += W[i] * W[i] / den;
  }
  return 1.0 + rho * group_sum(acc);
}

// Roots of 1 + rho sum_i W_i^2 / (DL_i - lambda) = 0, one lane group per root.
__global__ void secular_kernel(const MergeDev* md, const double* DLall, const double* Wall,
                               int* orgOut, double* tauOut, double* lamOut) {
  const MergeDev m = md[blockIdx.y];
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = gid / kSecG, sub = gid % kSecG;
  if (j >= m.k) return;   // the whole group: j is the same in its lanes
  const double* DL = DLall + m.dlOff;
  const double* W = Wall + m.dlOff;
  const int k = m.k;
  const double rho = m.rho;
  int org;
  double hi;        // |tau| bound
  bool positive;    // tau > 0
  if (j < k - 1) {
    const double gap = DL[j + 1] - DL[j];
    const double fm = secular_f_group(DL, W, k, rho, DL[j], 0.5 * gap, sub);
    if (fm >= 0.0) {
      org = j;
      positive = true;
    } else {
      org = j + 1;
      positive = false;
    }
    hi = 0.5 * gap;
  } else {
    double nz = 0.0;
    for (int i = sub; i < k; i += kSecG) nz += W[i] * W[i];
    nz = group_sum(nz);
    org = k - 1;
    positive = true;
    hi = rho * nz;
  }
  const double dorg = DL[org];
  // bisection over bit patterns of |tau| in (0, hi]
  unsigned long long lo_b = 0ull, hi_b = (unsigned long long)__double_as_longlong(hi);
  while (hi_b - lo_b > 1ull) {
    const unsigned long long mid = lo_b + (hi_b - lo_b) / 2;
    const double t = __longlong_as_double((long long)mid);
    const double f = secular_f_group(DL, W, k, rho, dorg, positive ? t : -t, sub);
    // positive: f increases with t, root where f >= 0 first; negative: f(-t) decreases with t
    const bool right = positive ? (f >= 0.0) : (f < 0.0);
    if (right) hi_b = mid;
    else lo_b = mid;
  }
  double t = __longlong_as_double((long long)hi_b);
  if (lo_b > 0ull) {
    const double tl = __longlong_as_double((long long)lo_b);
    const double fh = fabs(secular_f_group(DL, W, k, rho, dorg, positive ? t : -t, sub));
    const double fl = fabs(secular_f_group(DL, W, k, rho, dorg, positive ? tl : -tl, sub));
    if (fl < fh) t = tl;
  }
  const double tau = positive ? t : -t;
  if (sub == 0) {
    orgOut[m.dlOff + j] = org;
    tauOut[m.dlOff + j] = tau;
    lamOut[m.dlOff + j] = dorg + tau;
  }
}

// zhat_i = sign(W_i) sqrt( -(DL_i - lam_i)/rho * prod_{j != i} (DL_i - lam_j)/(DL_i - DL_j) )
__global__ void zhat_kernel(const MergeDev* md, const double* DLall, const double* Wall,
                            const int* orgAll, const double* tauAll, double* zhat) {
  const MergeDev m = md[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.k) return;
  const double* DL = DLall + m.dlOff;
  const int* org = orgAll + m.dlOff;
  const double* tau = tauAll + m.dlOff;
  const double di = DL[i];
  double prod = -((di - DL[org[i]]) - tau[i]) / m.rho;
  // interleave factors below and above i to keep the running product near one
  int lo = i - 1, up = i + 1;
  while (lo >= 0 || up < m.k) {
    if (lo >= 0) {
      prod *= ((di - DL[org[lo]]) - tau[lo]) / (di - DL[lo]);
      --lo;
    }
    if (up < m.k) {
      prod *= ((di - DL[org[up]]) - tau[up]) / (di - DL[up]);
      ++up;
    }
  }
  const double w = Wall[m.dlOff + i];
  const double a = sqrt(fmax(prod, 0.0));
  zhat[m.dlOff + i] = w >= 0.0 ? a : -a;
}

// 1 / || zhat / (DL - lam_j) || for each root j, one lane group per root
__global__ void colnorm_kernel(const MergeDev* md, const double* DLall, const int* orgAll,
                               const double* tauAll, const double* zhat, double* inv) {
  const MergeDev m = md[blockIdx.y];
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = gid / kSecG, sub = gid % kSecG;
  if (j >= m.k) return;
  const double* DL = DLall + m.dlOff;
  const double dorg = DL[orgAll[m.dlOff + j]];
  const double tj = tauAll[m.dlOff + j];
  double acc = 0.0;
  for (int i = sub; i < m.k; i += kSecG) {
    const double v = zhat[m.dlOff + i] / ((DL[i] - dorg) - tj);
    acc += v * v;
  }
  acc = group_sum(acc);
  if (sub == 0) inv[m.dlOff + j] = 1.0 / sqrt(acc);
}

// U(row(i), j) = zhat_i / (DL_i - lam_j) * inv_j, rows in column-group order
__global__ void buildu_kernel(const MergeDev* md, const double* DLall, const int* orgAll,
                              const double* tauAll, const double* zhat, const double* inv,
                              const int* rowOf, double* U) {
  const MergeDev m = md[blockIdx.z];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= m.k || j >= m.k) return;
  const double* DL = DLall + m.dlOff;
  const double dorg = DL[orgAll[m.dlOff + j]];
  const double v = zhat[m.dlOff + i] / ((DL[i] - dorg) - tauAll[m.dlOff + j]) * inv[m.dlOff + j];
  U[m.uOff + rowOf[m.dlOff + i] + (long)j * m.k] = v;
}

// Qg(:, q) = Q(block rows, gcol[q]) for the k nondeflated columns in group order
__global__ void gatherq_kernel(const MergeDev* md, const int* gcol, const double* Q, long ldq,
                               double* Qg) {
  const MergeDev m = md[blockIdx.z];
  const int q = blockIdx.y;
  if (q >= m.k) return;
  const int n = m.n1 + m.n2;
  const long col = gcol[m.dlOff + q];
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    Qg[m.qgOff + r + (long)q * n] = Q[(m.s + r) + col * ldq];
}

// deflated columns go after the k roots of the merged block
__global__ void copydefl_kernel(const MergeDev* md, const int* dcol, const double* Qold,
                                double* Qnew, long ldq) {
  const MergeDev m = md[blockIdx.z];
  const int q = blockIdx.y;
  if (q >= m.ndefl) return;
  const int n = m.n1 + m.n2;
  const long src = dcol[m.deflOff + q];
  const long dst = m.s + m.k + q;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    Qnew[(m.s + r) + dst * ldq] = Qold[(m.s + r) + src * ldq];
}

__global__ void gather_cols_kernel(const double* Q, long ldq, int n, const int* order, double* z,
                                   long ldz) {
  const int c = blockIdx.y;
  const long src = order[c];
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    z[r + (long)c * ldz] = Q[r + src * ldq];
}

// ------------------------------------------------------------------------------------------
struct Merge {
  int s, n1, n2;
};

// Every branch is split to the same depth, so each tree level covers the whole matrix and the
// two Q buffers alternate consistently between levels.
static void build_tree(int s, int n, int depth, int leafDepth, std::vector<Merge>* levels,
                       int& maxDepth, std::vector<int>& ls, std::vector<int>& lz) {
  if (depth == leafDepth) {
    ls.push_back(s);
    lz.push_back(n);
    return;
  }
  const int n1 = n / 2;
  build_tree(s, n1, depth + 1, leafDepth, levels, maxDepth, ls, lz);
  build_tree(s + n1, n - n1, depth + 1, leafDepth, levels, maxDepth, ls, lz);
  levels[depth].push_back({s, n1, n - n1});
  maxDepth = std::max(maxDepth, depth);
}

size_t stedc_work_size(int n) {
  // two Q buffers, gathered columns and U: 4 n^2
  return 4 * (size_t)n * n + 64;
}

int stedc(hipStream_t s, int n, const double* dHost, const double* eHost, double* w, double* z,
          long ldz, double* work, size_t workSize, int* info, StedcWs* ws) {
  *info = 0;
  if (n <= 0) return 0;
  std::unique_ptr<StedcWs> own;
  if (!ws) {
    own.reset(new StedcWs());
    ws = own.get();
  }
  DH_CHECK(workSize >= stedc_work_size(n), "stedc: workspace too small");
  double* Qa = work;
  double* Qb = work + (size_t)n * n;
  double* Qg = Qb + (size_t)n * n;
  double* Ub = Qg + (size_t)n * n;
  const long ldq = n;
  std::vector<double> d(dHost, dHost + n);
  std::vector<double> e(std::max(n - 1, 1), 0.0);
  if (n > 1) std::copy(eHost, eHost + n - 1, e.begin());
  std::vector<Merge> levels[64];
  int maxDepth = -1;
  std::vector<int> ls, lz;
  int leafDepth = 0;
  while (((n + (1 << leafDepth) - 1) >> leafDepth) > kLeaf) ++leafDepth;
  build_tree(0, n, 0, leafDepth, levels, maxDepth, ls, lz);
  // tearing
  for (int L = 0; L <= maxDepth; ++L)
    for (const Merge& mg : levels[L]) {
      const double b = fabs(e[mg.s + mg.n1 - 1]);
      d[mg.s + mg.n1 - 1] -= b;
      d[mg.s + mg.n1] -= b;
    }
  DevBuf<double>&dd = ws->dd, &de = ws->de, &dD = ws->dD, &dz = ws->dz;
  DevBuf<int>&dls = ws->dls, &dlz = ws->dlz, &dinfo = ws->dinfo;
  DH_TRY(dd.upload(d.data(), n, s));
  DH_TRY(de.upload(e.data(), std::max(n - 1, 1), s));
  DH_TRY(dls.upload(ls.data(), ls.size(), s));
  DH_TRY(dlz.upload(lz.data(), lz.size(), s));
  DH_TRY(dD.alloc(n));
  DH_TRY(dz.alloc(n));
  DH_TRY(dinfo.alloc(1));
  DH_TRY(fill<int>(s, dinfo.p, 1, 0));
  DH_TRY(fill<double>(s, Qa, (size_t)n * n, 0.0));
  DH_TRY(fill<double>(s, Qb, (size_t)n * n, 0.0));
  hipLaunchKernelGGL(leaf_kernel, dim3((unsigned)ls.size()), dim3(64), 0, s, dd.p, de.p, dls.p,
                     dlz.p, Qa, ldq, dD.p, dinfo.p);
  DH_LAUNCH_CHECK();
  int hinfo = 0;
  DH_HIP(hipMemcpyAsync(&hinfo, dinfo.p, sizeof(int), hipMemcpyDeviceToHost, s));
  std::vector<double> D(n);
  DH_HIP(hipMemcpyAsync(D.data(), dD.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  if (hinfo != 0) {
    *info = hinfo;
    set_error("stedc: QL iteration did not converge in a leaf");
    return 0;
  }
  // per-block ascending order of columns (global column indices)
  std::vector<int> order(n);
  for (size_t q = 0; q < ls.size(); ++q)
    for (int i = 0; i < lz[q]; ++i) order[ls[q] + i] = ls[q] + i;
  const double eps = 2.220446049250313e-16;
  std::vector<double> zh(n);
  DevBuf<MergeDev>& dmd = ws->dmd;
  DevBuf<double>&dDL = ws->dDL, &dW = ws->dW, &dtau = ws->dtau, &dlam = ws->dlam,
         &dzhat = ws->dzhat, &dinv = ws->dinv, &drc = ws->drc, &drs = ws->drs;
  DevBuf<int>&dorg = ws->dorg, &drow = ws->drow, &dgcol = ws->dgcol, &ddcol = ws->ddcol,
      &drp = ws->drp, &drn = ws->drn;
  for (int L = maxDepth; L >= 0; --L) {
    const std::vector<Merge>& mv = levels[L];
    if (mv.empty()) continue;
    const int nm = (int)mv.size();
    std::vector<MergeDev> md(nm);
    int maxn = 0;
    for (int q = 0; q < nm; ++q) {
      md[q] = {};
      md[q].s = mv[q].s;
      md[q].n1 = mv[q].n1;
      md[q].n2 = mv[q].n2;
      const double b = e[mv[q].s + mv[q].n1 - 1];
      md[q].rho = fabs(b);
      md[q].sigma = b >= 0.0 ? 1.0 : -1.0;
      maxn = std::max(maxn, mv[q].n1 + mv[q].n2);
    }
    DH_TRY(dmd.upload(md.data(), nm, s));
    hipLaunchKernelGGL(zgather_kernel, dim3(std::min(cdiv(maxn, 256), 64u), nm), dim3(256), 0, s,
                       dmd.p, Qa, ldq, dz.p);
    DH_LAUNCH_CHECK();
    DH_HIP(hipMemcpyAsync(zh.data(), dz.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    DH_HIP(hipStreamSynchronize(s));
    // host deflation for every merge of the level (LAPACK dlaed2 rules)
    std::vector<double> DL, Wv;
    std::vector<int> rowOf, gcol, dcol, rp, rn;
    std::vector<double> rc, rs;
    std::vector<std::vector<std::pair<double, int>>> deflList(nm);
    long qgOff = 0, uOff = 0;
    for (int q = 0; q < nm; ++q) {
      MergeDev& m = md[q];
      const int nb = m.n1 + m.n2;
      const int s0 = m.s;
      // merge the two ascending child orders
      std::vector<int> idx(nb);
      std::merge(order.begin() + s0, order.begin() + s0 + m.n1, order.begin() + s0 + m.n1,
                 order.begin() + s0 + nb, idx.begin(),
                 [&](int a, int b) { return D[a] < D[b]; });
      const double rho = 2.0 * m.rho;
      const double isq2 = 1.0 / std::sqrt(2.0);
      std::vector<double> zz(nb);
      std::vector<int> type(nb);
      double zmax = 0.0, dmax = 0.0;
      for (int c = 0; c < nb; ++c) {
        const int col = s0 + c;
        zz[c] = zh[col] * isq2;
        type[c] = (c < m.n1) ? 1 : 3;
        zmax = std::max(zmax, std::fabs(zz[c]));
        dmax = std::max(dmax, std::fabs(D[col]));
      }
      const double tol = 8.0 * eps * std::max(dmax, zmax);
      auto zc = [&](int col) -> double& { return zz[col - s0]; };
      auto ty = [&](int col) -> int& { return type[col - s0]; };
      std::vector<int> nd;   // nondeflated columns (ascending)
      std::vector<double> ndD;
      auto& defl = deflList[q];
      m.rotOff = (long)rp.size();
      int pj = -1;
      for (int j = 0; j < nb; ++j) {
        const int nj = idx[j];
        if (rho * std::fabs(zc(nj)) <= tol) {
          defl.push_back({D[nj], nj});
          continue;
        }
        if (pj < 0) {
          pj = nj;
          continue;
        }
        double sv = zc(pj), cv = zc(nj);
        const double tau = std::hypot(cv, sv);
        const double t = D[nj] - D[pj];
        cv /= tau;
        sv = -sv / tau;
        if (std::fabs(t * cv * sv) <= tol) {
          zc(nj) = tau;
          zc(pj) = 0.0;
          if (ty(nj) != ty(pj)) ty(nj) = 2;
          ty(pj) = 4;
          rp.push_back(pj);
          rn.push_back(nj);
          rc.push_back(cv);
          rs.push_back(sv);
          const double tt = D[pj] * cv * cv + D[nj] * sv * sv;
          D[nj] = D[pj] * sv * sv + D[nj] * cv * cv;
          D[pj] = tt;
          defl.push_back({D[pj], pj});
          pj = nj;
        } else {
          nd.push_back(pj);
          ndD.push_back(D[pj]);
          pj = nj;
        }
      }
      if (pj >= 0) {
        nd.push_back(pj);
        ndD.push_back(D[pj]);
      }
      m.nrot = (int)(rp.size() - m.rotOff);
      std::sort(defl.begin(), defl.end());
      m.k = (int)nd.size();
      m.ndefl = (int)defl.size();
      m.dlOff = (long)DL.size();
      m.deflOff = (long)dcol.size();
      m.rho = rho;
      // column groups: 1 (child-1 rows only), 2 (both), 3 (child-2 rows only)
      int k1 = 0, k2 = 0, k3 = 0;
      for (int c : nd) {
        const int t = ty(c);
        if (t == 1) ++k1;
        else if (t == 2) ++k2;
        else ++k3;
      }
      m.k1 = k1;
      m.k2 = k2;
      m.k3 = k3;
      std::vector<int> g(m.k);
      int p1 = 0, p2 = k1, p3 = k1 + k2;
      std::vector<int> gc(m.k);
      for (int i = 0; i < m.k; ++i) {
        const int t = ty(nd[i]);
        const int pos = (t == 1) ? p1++ : (t == 2 ? p2++ : p3++);
        g[i] = pos;
        gc[pos] = nd[i];
      }
      for (int i = 0; i < m.k; ++i) {
        DL.push_back(ndD[i]);
        Wv.push_back(zc(nd[i]));
        rowOf.push_back(g[i]);
        gcol.push_back(gc[i]);
      }
      for (auto& pr : defl) dcol.push_back(pr.second);
      m.qgOff = qgOff;
      m.uOff = uOff;
      qgOff += (long)nb * m.k;
      uOff += (long)m.k * m.k;
    }
    const long totK = (long)DL.size();
    if (totK > 0) {
      DH_TRY(dDL.upload(DL.data(), totK, s));
      DH_TRY(dW.upload(Wv.data(), totK, s));
      DH_TRY(drow.upload(rowOf.data(), totK, s));
      DH_TRY(dgcol.upload(gcol.data(), totK, s));
      DH_TRY(dorg.alloc(totK));
      DH_TRY(dtau.alloc(totK));
      DH_TRY(dlam.alloc(totK));
      DH_TRY(dzhat.alloc(totK));
      DH_TRY(dinv.alloc(totK));
    }
    if (!dcol.empty()) DH_TRY(ddcol.upload(dcol.data(), dcol.size(), s));
    if (!rp.empty()) {
      DH_TRY(drp.upload(rp.data(), rp.size(), s));
      DH_TRY(drn.upload(rn.data(), rn.size(), s));
      DH_TRY(drc.upload(rc.data(), rc.size(), s));
      DH_TRY(drs.upload(rs.data(), rs.size(), s));
    }
    DH_TRY(dmd.upload(md.data(), nm, s));
    int maxk = 0, maxdefl = 0;
    for (auto& m : md) {
      maxk = std::max(maxk, m.k);
      maxdefl = std::max(maxdefl, m.ndefl);
    }
    if (!rp.empty()) {
      hipLaunchKernelGGL(rotate_kernel, dim3(cdiv(maxn, 256), nm), dim3(256), 0, s, dmd.p, drp.p,
                         drn.p, drc.p, drs.p, Qa, ldq);
      DH_LAUNCH_CHECK();
    }
    if (maxk > 0) {
      const dim3 g1(cdiv(maxk, 64), nm), gg(cdiv((long)maxk * kSecG, 64), nm);
      hipLaunchKernelGGL(secular_kernel, gg, dim3(64), 0, s, dmd.p, dDL.p, dW.p, dorg.p, dtau.p,
                         dlam.p);
      DH_LAUNCH_CHECK();
      hipLaunchKernelGGL(zhat_kernel, g1, dim3(64), 0, s, dmd.p, dDL.p, dW.p, dorg.p, dtau.p,
                         dzhat.p);
      DH_LAUNCH_CHECK();
      hipLaunchKernelGGL(colnorm_kernel, gg, dim3(64), 0, s, dmd.p, dDL.p, dorg.p, dtau.p,
                         dzhat.p, dinv.p);
      DH_LAUNCH_CHECK();
      hipLaunchKernelGGL(buildu_kernel, dim3(cdiv(maxk, 256), maxk, nm), dim3(256), 0, s, dmd.p,
                         dDL.p, dorg.p, dtau.p, dzhat.p, dinv.p, drow.p, Ub);
      DH_LAUNCH_CHECK();
      hipLaunchKernelGGL(gatherq_kernel, dim3(std::min(cdiv(maxn, 256), 16u), maxk, nm), dim3(256),
                         0, s, dmd.p, dgcol.p, Qa, ldq, Qg);
      DH_LAUNCH_CHECK();
      std::vector<GemmDesc<double>> gd;
      int gm = 0, gn = 0;
      for (auto& m : md) {
        if (m.k == 0) continue;
        const int nb = m.n1 + m.n2;
        const double* qg = Qg + m.qgOff;
        const double* u = Ub + m.uOff;
        double* qn = Qb + m.s + (long)m.s * ldq;
        if (m.k1 + m.k2 > 0)
          gd.push_back({m.n1, m.k, m.k1 + m.k2, qg, (long)nb, u, (long)m.k, qn, ldq});
        if (m.k2 + m.k3 > 0)
          gd.push_back({m.n2, m.k, m.k2 + m.k3, qg + m.n1 + (long)m.k1 * nb, (long)nb, u + m.k1,
                        (long)m.k, qn + m.n1, ldq});
        if (m.k1 + m.k2 == 0) gd.push_back({m.n1, m.k, 0, qg, (long)nb, u, (long)m.k, qn, ldq});
        if (m.k2 + m.k3 == 0)
          gd.push_back({m.n2, m.k, 0, qg, (long)nb, u, (long)m.k, qn + m.n1, ldq});
        gm = std::max(gm, std::max(m.n1, m.n2));
        gn = std::max(gn, m.k);
      }
      DevBuf<GemmDesc<double>>& dgd = ws->dgd;
      DH_TRY(dgd.upload(gd.data(), gd.size(), s));
      DH_TRY(gemm_batched<double>(s, false, false, dgd.p, (int)gd.size(), gm, gn, 1.0, 0.0));
      DH_HIP(hipStreamSynchronize(s));
    }
    if (maxdefl > 0) {
      hipLaunchKernelGGL(copydefl_kernel, dim3(std::min(cdiv(maxn, 256), 16u), maxdefl, nm),
                         dim3(256), 0, s, dmd.p, ddcol.p, Qa, Qb, ldq);
      DH_LAUNCH_CHECK();
    }
    std::vector<double> lam(totK);
    if (totK > 0)
      DH_HIP(hipMemcpyAsync(lam.data(), dlam.p, totK * sizeof(double), hipMemcpyDeviceToHost, s));
    DH_HIP(hipStreamSynchronize(s));
    // new eigenvalues and ascending orders of the merged blocks
    for (int q = 0; q < nm; ++q) {
      const MergeDev& m = md[q];
      const int nb = m.n1 + m.n2;
      std::vector<std::pair<double, int>> all;
      all.reserve(nb);
      for (int j = 0; j < m.k; ++j) {
        D[m.s + j] = lam[m.dlOff + j];
        all.push_back({D[m.s + j], m.s + j});
      }
      for (int t = 0; t < m.ndefl; ++t) {
        D[m.s + m.k + t] = deflList[q][t].first;
        all.push_back({deflList[q][t].first, m.s + m.k + t});
      }
      std::stable_sort(all.begin(), all.end(),
                       [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                         return a.first < b.first;
                       });
      for (int i = 0; i < nb; ++i) order[m.s + i] = all[i].second;
    }
    std::swap(Qa, Qb);
  }
  // eigenvalues ascending and the ordered eigenvectors
  std::vector<double> wv(n);
  for (int i = 0; i < n; ++i) wv[i] = D[order[i]];
  DH_HIP(hipMemcpyAsync(w, wv.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
  DevBuf<int>& dord = ws->dord;
  DH_TRY(dord.upload(order.data(), n, s));
  hipLaunchKernelGGL(gather_cols_kernel, dim3(std::min(cdiv(n, 256), 64u), n), dim3(256), 0, s, Qa,
                     ldq, n, dord.p, z, ldz);
  DH_LAUNCH_CHECK();
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace dftbhip
