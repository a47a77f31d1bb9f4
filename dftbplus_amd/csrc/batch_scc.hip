// The host-side steps of a molecule batch's SCC iteration as device kernels (batch_scc.hip):
// per-molecule Fermi filling (Efilling, src/dftbp/dftb/etemp.F90:59-194, one wave per
// molecule), output charges and band-structure energy, SCC energy, convergence test and
// modified Broyden mixing (src/dftbp/mixer/broydenmixer.F90:204-342, one thread per molecule),
// so that a whole iteration is a fixed sequence of launches on one stream that a HIP graph
// captures. The iteration number, the active flags and the count of molecules still iterating
// live on the device; the host reads two integers per iteration.
#include <algorithm>
#include <cmath>

#include "batch.h"
#include "common.h"

namespace dftbhip {

// ---------------------------------------------------------------------------------------------
// iteration bookkeeping
// ---------------------------------------------------------------------------------------------
// Start of an iteration: the iteration number, the counters read by the host, and dq = qin - q0
__global__ void bscc_begin_kernel(BatchSccDev d) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a == 0) {
    d.counters[0] += 1;            // iteration number
    d.counters[1] = 0;             // molecules still iterating after this iteration
    d.counters[2] = 0;             // failure code of this iteration
    d.counters[3] = 0x7fffffff;    // lowest failing molecule
  }
  if (a < d.nat) d.dq[a] = d.qin[a] - d.q0[a];
}

// molecules whose eigensolver failed in this iteration (the lowest index is reported)
__global__ void bscc_info_kernel(BatchSccDev d, const int* info) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.nmol || !d.active[m]) return;
  if (info[m] != 0) {
    atomicMin(&d.counters[3], m);
    atomicExch(&d.counters[2], 1);
  }
}

// ---------------------------------------------------------------------------------------------
// Fermi filling, one wave per molecule. Sums are per-lane Kahan sums combined by an xor
// butterfly, whose result is the same in every lane, so every lane takes the same branches.
// ---------------------------------------------------------------------------------------------
__device__ inline void kahan_add(double& s, double& c, double v) {
  const double y = v - c;
  const double t = s + y;
  c = (t - s) - y;
  s = t;
}

__device__ inline double wave_sum(double v) {
  for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline double electron_count_w(const double* e, int n, double ef, double kT) {
  double s = 0.0, c = 0.0;
  for (int j = threadIdx.x & 63; j < n; j += 64) kahan_add(s, c, 1.0 / (1.0 + exp((e[j] - ef) / kT)));
  return wave_sum(s - c);
}

// fillings (factor 2 and the rescale included), Eband, TS, E0 at ef (electronFill, etemp.F90:
// 294-455): a level past the first one with f <= 1e-15 is empty, and that one enters neither
// sum; returns the index of the last level with f >= eps (+1) through nocc
__device__ inline void electron_fill_w(const double* e, int n, double ef, double kT, double scale,
                                       double* f, double& eband, double& ts, int& nocc) {
  const double eps = 2.220446049250313e-16;
  double eb = 0.0, cb = 0.0, tsum = 0.0, ct = 0.0;
  int last = 0;
  for (int j = threadIdx.x & 63; j < n; j += 64) {
    const double x = (e[j] - ef) / kT;
    const double fj = x > 700.0 ? 0.0 : 1.0 / (1.0 + exp(x));
    double fp = 1.0;
    if (j > 0) {
      const double xp = (e[j - 1] - ef) / kT;
      fp = xp > 700.0 ? 0.0 : 1.0 / (1.0 + exp(xp));
    }
    const bool kept = j == 0 || fp > 1.0e-15;   // levels are ascending: f is non-increasing
    const double fv = kept ? fj : 0.0;
    if (kept && fj > 1.0e-15) {
      if (fj > eps && fj < 1.0 - eps)
        kahan_add(tsum, ct, -(fj * log(fj) + (1.0 - fj) * log(1.0 - fj)));
      kahan_add(eb, cb, fj * e[j]);
    }
    const double out = 2.0 * fv * scale;
    f[j] = out;
    if (fabs(out) >= eps) last = j + 1;
  }
  eband = wave_sum(eb - cb);
  ts = wave_sum(tsum - ct) * kT;
  for (int o = 1; o < 64; o <<= 1) last = max(last, __shfl_xor(last, o, 64));
  nocc = last;
}

__global__ __launch_bounds__(256) void bscc_fermi_kernel(BatchSccDev d, const double* w,
                                                         double* fill) {
  const int m = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (m >= d.nmol || !d.active[m]) return;
  const int lane = threadIdx.x & 63;
  const long o0 = d.orb0[m];
  const int n = (int)(d.orb0[m + 1] - o0);
  const double* e = w + o0;
  double* f = fill + o0;
  const double kT = fmax(d.kT, 1.0e-8);
  const double ne = 0.5 * d.nel[m];
  const double eps2 = 2.0 * 2.220446049250313e-16;
  double ef = 0.0, scale = 1.0;
  bool done = false;
  if (ne < 2.220446049250313e-16 || n == 0) {
    for (int j = lane; j < n; j += 64) f[j] = 0.0;
    if (lane == 0) {
      d.ef[m] = (n ? e[0] : 0.0) - 1000.0 * (kT + 1.1920928955078125e-07);
      d.eband[m] = d.ts[m] = d.e0[m] = 0.0;
      d.nocc[m] = 0;
    }
    return;
  }
  if (n <= ne) {
    ef = e[n - 1] + 1000.0 * (kT + 1.1920928955078125e-07);
    done = true;
  }
  if (!done && fabs(ne - rint(ne)) <= 1.0e-15) {
    int k = (int)ceil(ne - 1e-12);
    k = min(k, n - 1);
    ef = 0.5 * (e[k] + e[k - 1]);
    if (fabs(ne - electron_count_w(e, n, ef, kT)) <= 1.0e-7) done = true;
  }
  if (!done) {
    double upper = e[n - 1] + 0.01, lower = e[0] - 0.01;
    double nmax = electron_count_w(e, n, upper, kT), nmin = electron_count_w(e, n, lower, kT);
    for (int g = 0; g < 64 && nmin > ne; ++g) {
      lower = 2.0 * (lower - upper) + lower;
      nmin = electron_count_w(e, n, lower, kT);
    }
    for (int g = 0; g < 64 && nmax < ne; ++g) {
      upper = 2.0 * (upper - lower) + lower;
      nmax = electron_count_w(e, n, upper, kT);
    }
    ef = 0.5 * (upper + lower);
    double nc = electron_count_w(e, n, ef, kT);
    // the interval halves every step: at most ~2100 halvings reach the resolution of a double
    for (int g = 0; g < 4096 && fabs(ne - nc) > 1.0e-15 &&
                    fabs(upper - lower) >= fmax(fabs(ef) * eps2, eps2);
         ++g) {
      if ((nmax >= nmin) == (ne >= nc)) {
        lower = ef;
        nmin = nc;
      } else {
        upper = ef;
        nmax = nc;
      }
      ef = 0.5 * (upper + lower);
      nc = electron_count_w(e, n, ef, kT);
    }
    if (fabs(ne - nc) > 1.0e-7) {
      // the search did not converge: reported like an eigensolver failure
      if (lane == 0) {
        atomicMin(&d.counters[3], m);
        atomicExch(&d.counters[2], 2);
      }
      return;
    }
    if (ne > eps2 && fabs(ne - nc) > eps2) scale = ne / nc;
  }
  double eb, ts;
  int nocc;
  electron_fill_w(e, n, ef, kT, scale, f, eb, ts, nocc);
  if (lane == 0) {
    d.ef[m] = ef;
    d.eband[m] = 2.0 * eb;
    d.ts[m] = 2.0 * ts;
    d.e0[m] = 2.0 * (eb - 0.5 * ts);
    d.nocc[m] = nocc;
  }
}

// ---------------------------------------------------------------------------------------------
// output charges, band-structure (non-SCC) energy
// ---------------------------------------------------------------------------------------------
// qout per atom from the orbital populations, dq = qout - q0 (the charges the output potential
// is built from); atoms of inactive molecules keep their last values
__global__ void bscc_qout_kernel(BatchSccDev d, const long* atomStart, const int* atomMol,
                                 const double* qorb) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= d.nat) return;
  if (!d.active[atomMol[a]]) {
    d.dq[a] = d.qout[a] - d.q0[a];
    return;
  }
  double q = 0.0;
  for (long o = atomStart[a]; o < atomStart[a + 1]; ++o) q += qorb[o];
  d.qout[a] = q;
  d.dq[a] = q - d.q0[a];
}

// eNonScc per molecule: the sum of the orbital energies of Mulliken (one wave per molecule)
__global__ __launch_bounds__(256) void bscc_enonscc_kernel(BatchSccDev d, const double* eorb) {
  const int m = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (m >= d.nmol || !d.active[m]) return;
  double s = 0.0;
  for (long o = d.orb0[m] + (threadIdx.x & 63); o < d.orb0[m + 1]; o += 64) s += eorb[o];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) d.eNonScc[m] = s;
}

// ---------------------------------------------------------------------------------------------
// SCC energy, convergence and Broyden mixing: one thread per molecule, in the operation order
// of the host mixer (sequential dot products, LU with partial pivoting)
// ---------------------------------------------------------------------------------------------
__device__ double dotn(const double* a, const double* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

__global__ void bscc_mix_kernel(BatchSccDev d, const double* vout) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.nmol || !d.active[m]) return;
  const int it = d.counters[0];
  const int a0 = d.atom0[m], na = d.atom0[m + 1] - a0;
  double es = 0.0, er = 0.0;
  for (int a = a0; a < a0 + na; ++a) {
    es += 0.5 * vout[a] * d.dq[a];
    er = fmax(er, fabs(d.qout[a] - d.qin[a]));
  }
  d.eScc[m] = es;
  d.err[m] = er;
  d.iters[m] = it;
  if (er < d.sccTol && it >= d.minSccIter) {
    d.conv[m] = 1;
    d.active[m] = 0;
    return;
  }
  if (it == d.maxSccIter) {
    d.active[m] = 0;
    return;
  }
  atomicAdd(&d.counters[1], 1);
  // modified Broyden mixing of this molecule's charges (qin <- mixed)
  const int H = d.maxHist;
  double* q = d.qin + a0;
  double* dlast = d.dlast + a0;
  double* qlast = d.qlast + a0;
  double* ww = d.ww + (long)m * H;
  double* aa = d.aa + (long)m * H * H;
  double* B = d.B + (long)m * H * H;
  double* cc = d.cc + (long)m * H;
  double* dfu = d.tmp + a0;
  double* uu0 = d.uu;
  double* dF0 = d.dF;
  const long hs = d.nat;   // stride of the history vectors (all atoms of the batch)
  const int nn = ++d.mixIt[m];
  if (nn == 1) {
    for (int i = 0; i < na; ++i) {
      const double di = d.qout[a0 + i] - q[i];
      qlast[i] = q[i];
      dlast[i] = di;
      q[i] += d.alpha * di;
    }
    return;
  }
  if (nn - 1 > H) {   // more steps than the history holds (reference: an error)
    atomicMin(&d.counters[3], m);
    atomicExch(&d.counters[2], 3);
    return;
  }
  const int n1 = nn - 1;
  // the difference d = qout - qin is formed element by element where needed
  double dd = 0.0;
  for (int i = 0; i < na; ++i) {
    const double di = d.qout[a0 + i] - q[i];
    dd += di * di;
  }
  double wgt = sqrt(dd);
  wgt = (wgt > d.wfac / d.maxw) ? d.wfac / wgt : d.maxw;
  wgt = fmax(wgt, d.minw);
  ww[n1 - 1] = wgt;
  double nf = 0.0;
  for (int i = 0; i < na; ++i) {
    dfu[i] = (d.qout[a0 + i] - q[i]) - dlast[i];
    nf += dfu[i] * dfu[i];
  }
  double inv = fmax(sqrt(nf), 2.220446049250313e-16);
  inv = 1.0 / inv;
  for (int i = 0; i < na; ++i) dfu[i] *= inv;
  for (int ii = 0; ii < nn - 2; ++ii) {
    const double* dFi = dF0 + ii * hs + a0;
    aa[ii + (long)(n1 - 1) * H] = dotn(dFi, dfu, na);
    aa[(n1 - 1) + (long)ii * H] = aa[ii + (long)(n1 - 1) * H];
    double s = 0.0;
    for (int i = 0; i < na; ++i) s += dFi[i] * (d.qout[a0 + i] - q[i]);
    cc[ii] = ww[ii] * s;
  }
  aa[(n1 - 1) + (long)(n1 - 1) * H] = 1.0;
  {
    double s = 0.0;
    for (int i = 0; i < na; ++i) s += dfu[i] * (d.qout[a0 + i] - q[i]);
    cc[n1 - 1] = ww[n1 - 1] * s;
  }
  // beta(i, j) = w_i w_j a(i, j) + omega0^2 delta(i, j), symmetric: solve beta x = cc (LU with
  // partial pivoting, rows swapped as they are chosen)
  const double om2 = d.omega0 * d.omega0;
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j)
      B[i + (long)j * H] = ww[i] * ww[j] * aa[i + (long)j * H] + (i == j ? om2 : 0.0);
  for (int k = 0; k < n1; ++k) {
    int p = k;
    for (int i = k + 1; i < n1; ++i)
      if (fabs(B[i + (long)k * H]) > fabs(B[p + (long)k * H])) p = i;
    if (p != k) {
      for (int j = 0; j < n1; ++j) {
        const double t = B[k + (long)j * H];
        B[k + (long)j * H] = B[p + (long)j * H];
        B[p + (long)j * H] = t;
      }
      const double t = cc[k];
      cc[k] = cc[p];
      cc[p] = t;
    }
    for (int i = k + 1; i < n1; ++i) {
      const double fct = B[i + (long)k * H] / B[k + (long)k * H];
      for (int j = k; j < n1; ++j) B[i + (long)j * H] -= fct * B[k + (long)j * H];
      cc[i] -= fct * cc[k];
    }
  }
  for (int k = n1 - 1; k >= 0; --k) {
    double s = cc[k];
    for (int j = k + 1; j < n1; ++j) s -= B[k + (long)j * H] * cc[j];
    cc[k] = s / B[k + (long)k * H];
  }
  double* dFn = dF0 + (n1 - 1) * hs + a0;
  double* un = uu0 + (n1 - 1) * hs + a0;
  for (int i = 0; i < na; ++i) {
    dFn[i] = dfu[i];
    un[i] = d.alpha * dfu[i] + inv * (q[i] - qlast[i]);
  }
  for (int i = 0; i < na; ++i) {
    const double di = d.qout[a0 + i] - q[i];
    qlast[i] = q[i];
    dlast[i] = di;
    double v = q[i] + d.alpha * di;
    for (int ii = 0; ii < nn - 2; ++ii) v -= ww[ii] * cc[ii] * uu0[ii * hs + a0 + i];
    v -= ww[n1 - 1] * cc[n1 - 1] * un[i];
    q[i] = v;
  }
}

// reset of the SCC state: qin = q0, every molecule active, counters and mixers cleared
__global__ void bscc_reset_kernel(BatchSccDev d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < d.nat) {
    d.qin[t] = d.q0[t];
    d.qout[t] = d.q0[t];
  }
  if (t < d.nmol) {
    d.active[t] = 1;
    d.conv[t] = 0;
    d.iters[t] = 0;
    d.mixIt[t] = 0;
    d.err[t] = 0.0;
    d.eNonScc[t] = d.eScc[t] = 0.0;
    d.ef[t] = d.eband[t] = d.ts[t] = d.e0[t] = 0.0;
  }
  if (t == 0) {
    d.counters[0] = 0;
    d.counters[1] = d.nmol;
    d.counters[2] = 0;
    d.counters[3] = 0x7fffffff;
  }
}

// ---------------------------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------------------------
int launch_bscc_reset(hipStream_t s, const BatchSccDev& d) {
  const int n = std::max(d.nat, d.nmol);
  hipLaunchKernelGGL(bscc_reset_kernel, dim3(cdiv(std::max(n, 1), 256)), dim3(256), 0, s, d);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_bscc_begin(hipStream_t s, const BatchSccDev& d) {
  hipLaunchKernelGGL(bscc_begin_kernel, dim3(cdiv(std::max(d.nat, 1), 256)), dim3(256), 0, s, d);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_bscc_info(hipStream_t s, const BatchSccDev& d, const int* info) {
  hipLaunchKernelGGL(bscc_info_kernel, dim3(cdiv(d.nmol, 256)), dim3(256), 0, s, d, info);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_bscc_fermi(hipStream_t s, const BatchSccDev& d, const double* w, double* fill) {
  hipLaunchKernelGGL(bscc_fermi_kernel, dim3(cdiv(d.nmol, 4)), dim3(256), 0, s, d, w, fill);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_bscc_outputs(hipStream_t s, const BatchSccDev& d, const long* atomStart,
                        const int* atomMol, const double* qorb, const double* eorb) {
  hipLaunchKernelGGL(bscc_qout_kernel, dim3(cdiv(d.nat, 256)), dim3(256), 0, s, d, atomStart,
                     atomMol, qorb);
  DH_LAUNCH_CHECK();
  hipLaunchKernelGGL(bscc_enonscc_kernel, dim3(cdiv(d.nmol, 4)), dim3(256), 0, s, d, eorb);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_bscc_mix(hipStream_t s, const BatchSccDev& d, const double* vout) {
  hipLaunchKernelGGL(bscc_mix_kernel, dim3(cdiv(d.nmol, 64)), dim3(64), 0, s, d, vout);
  DH_LAUNCH_CHECK();
  return 0;
}

}  // namespace dftbhip
