// Sparse (atom-pair block) kernels of the SCC-DFTB path:
//  * Slater-Koster H0/S assembly over the neighbour blocks (nonscc.F90 buildH0/buildS):
//    8-point Neville interpolation on the equidistant grid, 5th-order tail to zero over the
//    last bohr (slakoeqgrid.F90:183-244), s/p rotation to the bond direction (sk.F90:35-289);
//  * sparse -> dense unpack (both triangles written, periodic images folded) and the dense ->
//    sparse pack of the lower triangle (sparse2dense.F90:266-321, 1024-1097);
//  * H = H0 + 1/2 S (V_mu + V_nu) (shift.F90:132-229), Mulliken charges and the non-SCC energy
//    (populations.F90:120-193);
//  * the pair part of the forces 2 sum(rho dH0 - W dS + rho (V_mu+V_nu)/2 dS)
//    (forces.F90:548-672) with finite-difference (nonscc.F90:428-484) or analytic SK derivatives.
//
// Orbital order within an atom: s, p_y, p_z, p_x. Species carry at most two shells with l <= 1,
// i.e. at most 4 orbitals per atom.
#include "common.h"
#include "gamma_dev.h"
#include "sparse.h"

#include <cmath>

namespace dftbhip {

constexpr double kDistFudge = 1.0;
constexpr double kDeltaR = 1.0e-5;

// Neville interpolation as polyInterU2 (interpolation.F90:182-239) on 8 uniform points
// xa_i = (i0 + i) * dx, values y_i = tab[(i0 - 1 + i) * nint + col].
__device__ double neville8(const double* tab, int nint, int col, int i0, double dx, double xx,
                           double* dydx) {
  double xa[8], cc[8], dd[8], cp[8], dp[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    xa[i] = (double)(i0 + i) * dx;
    cc[i] = dd[i] = tab[(long)(i0 - 1 + i) * nint + col];
    cp[i] = dp[i] = 0.0;
  }
  double delta[7];
#pragma unroll
  for (int i = 0; i < 7; ++i) delta[i] = 1.0 / (xa[i + 1] - xa[0]);
  int icl = (int)ceil((xx - xa[0]) * delta[0]);
  icl = min(max(icl, 1), 8);
  double yy = cc[icl - 1];
  double yp = 0.0;
  icl -= 1;
  for (int mm = 1; mm < 8; ++mm) {
    for (int ii = 0; ii < 8 - mm; ++ii) {
      const double r2 = (dd[ii] - cc[ii + 1]) * delta[mm - 1];
      const double r2p = (dp[ii] - cp[ii + 1]) * delta[mm - 1];
      // derivative of c = (x_i - x) r2 and d = (x_{i+m} - x) r2
      cp[ii] = -r2 + (xa[ii] - xx) * r2p;
      dp[ii] = -r2 + (xa[ii + mm] - xx) * r2p;
      cc[ii] = (xa[ii] - xx) * r2;
      dd[ii] = (xa[ii + mm] - xx) * r2;
    }
    if (2 * icl < 8 - mm) {
      yy += cc[icl];
      yp += cp[icl];
    } else {
      yy += dd[icl - 1];
      yp += dp[icl - 1];
      icl -= 1;
    }
  }
  if (dydx) *dydx = yp;
  return yy;
}

// SK integrals (and optionally their radial derivatives) at distance r.
__device__ void sk_integrals(const double* tab, int nrows, int nint, double dx, double r,
                             double* out, double* dout) {
  const double rmax = nrows * dx + kDistFudge;
  const int ind = (int)floor(r / dx);
  for (int c = 0; c < nint; ++c) {
    out[c] = 0.0;
    if (dout) dout[c] = 0.0;
  }
  if (r >= rmax) return;
  if (ind < nrows) {
    int ilast = min(nrows, ind + 4);
    ilast = max(ilast, 8);
    for (int c = 0; c < nint; ++c) {
      double d1;
      out[c] = neville8(tab, nint, c, ilast - 7, dx, r, dout ? &d1 : nullptr);
      if (dout) dout[c] = d1;
    }
  } else {
    const int i0 = nrows - 7;
    const double x8 = nrows * dx;
    const double dr = r - rmax;
    for (int c = 0; c < nint; ++c) {
      const double y0 = neville8(tab, nint, c, i0, dx, x8 - kDeltaR, nullptr);
      const double y2 = neville8(tab, nint, c, i0, dx, x8 + kDeltaR, nullptr);
      const double y1 = tab[(long)(nrows - 1) * nint + c];
      const double y1p = (y2 - y0) / (2.0 * kDeltaR);
      const double y1pp = (y2 + y0 - 2.0 * y1) / (kDeltaR * kDeltaR);
      // poly5ToZero (interpolation.F90:31-67) with dx = -distFudge
      const double ddx = -kDistFudge;
      const double dx1 = y1p * ddx, dx2 = y1pp * ddx * ddx;
      const double a3 = 10.0 * y1 - 4.0 * dx1 + 0.5 * dx2;
      const double a4 = -15.0 * y1 + 7.0 * dx1 - 1.0 * dx2;
      const double a5 = 6.0 * y1 - 3.0 * dx1 + 0.5 * dx2;
      const double xr = dr * (-1.0 / kDistFudge);
      out[c] = ((a5 * xr + a4) * xr + a3) * xr * xr * xr;
      if (dout) {
        const double dpoly = ((5.0 * a5 * xr + 4.0 * a4) * xr + 3.0 * a3) * xr * xr;
        dout[c] = dpoly * (-1.0 / kDistFudge);
      }
    }
  }
}

// Rotated block hh[row * 4 + col] (row: orbital of atom 2, col: orbital of atom 1) for the bond
// unit vector u = (l, m, n); optional derivative dh[c][row*4+col] w.r.t. the bond vector R.
__device__ void sk_block(const SpeciesDev& s1, const SpeciesDev& s2, const double* sk,
                         const double* dsk, double r, const double u[3], double* hh,
                         double* dh /* 3 x 16 or null */) {
  const double uv[3] = {u[1], u[2], u[0]};   // orbital order y, z, x
  for (int i = 0; i < 16; ++i) hh[i] = 0.0;
  if (dh)
    for (int i = 0; i < 48; ++i) dh[i] = 0.0;
  int ind = 0, icol = 0;
  for (int a = 0; a < s1.nshell; ++a) {
    const int l1 = s1.ang[a];
    int irow = 0;
    for (int b = 0; b < s2.nshell; ++b) {
      const int l2 = s2.ang[b];
      if (l1 == 0 && l2 == 0) {
        hh[irow * 4 + icol] = sk[ind];
        if (dh)
          for (int c = 0; c < 3; ++c) dh[c * 16 + irow * 4 + icol] = dsk[ind] * u[c];
      } else if (l1 + l2 == 1) {
        const double sg = (l1 == 0) ? 1.0 : -1.0;   // (-1)^(l1+l2) transpose for p on atom 1
        for (int p = 0; p < 3; ++p) {
          const int rr = (l1 == 0) ? irow + p : irow;
          const int cc = (l1 == 0) ? icol : icol + p;
          hh[rr * 4 + cc] = sg * uv[p] * sk[ind];
          if (dh) {
            const int pc = (p == 0) ? 1 : (p == 1 ? 2 : 0);   // cartesian index of orbital p
            for (int c = 0; c < 3; ++c) {
              const double du = ((pc == c ? 1.0 : 0.0) - u[pc] * u[c]) / r;
              dh[c * 16 + rr * 4 + cc] = sg * (du * sk[ind] + uv[p] * u[c] * dsk[ind]);
            }
          }
        }
      } else {
        const double sg = sk[ind], pi = sk[ind + 1];
        for (int p = 0; p < 3; ++p)
          for (int q = 0; q < 3; ++q) {
            const double val = (sg - pi) * uv[p] * uv[q] + (p == q ? pi : 0.0);
            hh[(irow + p) * 4 + icol + q] = val;
            if (dh) {
              const int pc = (p == 0) ? 1 : (p == 1 ? 2 : 0);
              const int qc = (q == 0) ? 1 : (q == 1 ? 2 : 0);
              for (int c = 0; c < 3; ++c) {
                const double dup = ((pc == c ? 1.0 : 0.0) - u[pc] * u[c]) / r;
                const double duq = ((qc == c ? 1.0 : 0.0) - u[qc] * u[c]) / r;
                dh[c * 16 + (irow + p) * 4 + icol + q] =
                    (dsk[ind] - dsk[ind + 1]) * u[c] * uv[p] * uv[q] +
                    (p == q ? dsk[ind + 1] * u[c] : 0.0) + (sg - pi) * (dup * uv[q] + uv[p] * duq);
              }
            }
          }
      }
      ind += min(l1, l2) + 1;
      irow += 2 * l2 + 1;
    }
    icol += 2 * l1 + 1;
  }
}

__device__ __forceinline__ void pair_block(const SkDev& sk, int sp1, int sp2, bool ham,
                                           const double vec[3], double* hh, double* dh) {
  const double r = sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
  const double u[3] = {vec[0] / r, vec[1] / r, vec[2] / r};
  const int pair = sp1 * sk.nsp + sp2;
  const double* tab = (ham ? sk.htab : sk.stab) + sk.tabOff[pair];
  const int nint = sk.nint[pair];
  double v[8], dv[8];
  sk_integrals(tab, sk.nrows, nint, sk.dist, r, v, dh ? dv : nullptr);
  sk_block(sk.sp[sp1], sk.sp[sp2], v, dv, r, u, hh, dh);
}

__global__ void build_h0s_kernel(SkDev sk, PairsDev pr, const double* coords, double* h0,
                                 double* s) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= pr.nblk) return;
  const int a1 = pr.at1[k], i2 = pr.img2[k], a2 = pr.at2[k];
  const int sp1 = pr.species[a1], sp2 = pr.species[a2];
  const SpeciesDev& S1 = sk.sp[sp1];
  const SpeciesDev& S2 = sk.sp[sp2];
  const int n1 = S1.norb, n2 = S2.norb;
  const long st = pr.start[k];
  if (i2 == a1) {
    for (int c = 0; c < n1; ++c)
      for (int r = 0; r < n2; ++r) {
        h0[st + r + c * n2] = (r == c) ? S1.onsite[c] : 0.0;
        s[st + r + c * n2] = (r == c) ? 1.0 : 0.0;
      }
    return;
  }
  double vec[3];
  for (int c = 0; c < 3; ++c) vec[c] = coords[3 * (long)i2 + c] - coords[3 * (long)a1 + c];
  double hh[16], ss[16];
  pair_block(sk, sp1, sp2, true, vec, hh, nullptr);
  pair_block(sk, sp1, sp2, false, vec, ss, nullptr);
  for (int c = 0; c < n1; ++c)
    for (int r = 0; r < n2; ++r) {
      h0[st + r + c * n2] = hh[r * 4 + c];
      s[st + r + c * n2] = ss[r * 4 + c];
    }
}

// The per-iteration sparse kernels below give each SK block kBlkLanes lanes, lane e holding
// element e = r + c n2 of the column-major n2 x n1 block (lanes past n1 n2 idle): the blocks of a
// wave are consecutive in the sparse arrays, so their loads and stores are contiguous.
constexpr int kBlkLanes = 16;

// Dense fill of both triangles; blocks of periodic images of the same pair are summed.
template <typename T>
__global__ void unpack_kernel(PairsDev pr, const int* norbAtom, const long* iAtomStart,
                              const double* x, T* dense, long ld) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int k = (int)(t / kBlkLanes), e = (int)(t % kBlkLanes);
  if (k >= pr.nblk) return;
  const int a1 = pr.at1[k], a2 = pr.at2[k];
  const int n1 = norbAtom[a1], n2 = norbAtom[a2];
  if (e >= n1 * n2) return;
  const int r = e % n2, c = e / n2;
  const long o1 = iAtomStart[a1], o2 = iAtomStart[a2];
  const T v = (T)x[pr.start[k] + e];
  atomicAdd(&dense[(o2 + r) + (o1 + c) * ld], v);
  if (a1 != a2) atomicAdd(&dense[(o1 + c) + (o2 + r) * ld], v);
}

template <typename T>
__global__ void pack_kernel(PairsDev pr, const int* norbAtom, const long* iAtomStart,
                            const T* dense, long ld, double* x) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int k = (int)(t / kBlkLanes), e = (int)(t % kBlkLanes);
  if (k >= pr.nblk) return;
  const int a1 = pr.at1[k], a2 = pr.at2[k];
  const int n1 = norbAtom[a1], n2 = norbAtom[a2];
  if (e >= n1 * n2) return;
  const int r = e % n2, c = e / n2;
  const long o1 = iAtomStart[a1], o2 = iAtomStart[a2];
  long rr = r, cc = c;
  if (a1 == a2 && r < c) {   // on-site block symmetrised from its lower triangle
    rr = c;
    cc = r;
  }
  x[pr.start[k] + e] = (double)dense[(o2 + rr) + (o1 + cc) * ld];
}

// phase factor exp(i 2 pi kf.T) of block k; at a time-reversal-invariant point (m = 2 kf
// integral) it is (-1)^(m.T), taken exactly from the parity
__device__ inline void block_phase(const double* blkT, int k, double kx, double ky, double kz,
                                   int real2k, int m0, int m1, int m2, double& sn, double& cs) {
  if (real2k) {
    const long par = llround(m0 * blkT[3 * k] + m1 * blkT[3 * k + 1] + m2 * blkT[3 * k + 2]);
    cs = (par & 1) ? -1.0 : 1.0;
    sn = 0.0;
  } else {
    sincos(2.0 * M_PI * (kx * blkT[3 * k] + ky * blkT[3 * k + 1] + kz * blkT[3 * k + 2]), &sn,
           &cs);
  }
}

// Complex matrix at the k-point kf (fractional) into its real 2n x 2n embedding
// [[Re, -Im], [Im, Re]]: blocks times exp(i 2 pi kf.T) at (atom2 rows, atom1 columns) and the
// Hermitian counterpart; on-site blocks contribute their lower triangle with a real diagonal
// ('L' storage of unpackHS_cmplx_kpts, sparse2dense.F90:185-262). At a time-reversal-invariant
// point (real2k) the matrix is real: n x n, both triangles. The dense array is zeroed
// beforehand.
__global__ void unpack_kpt_kernel(PairsDev pr, const int* norbAtom, const long* iAtomStart,
                                  const double* blkT, double kx, double ky, double kz,
                                  int real2k, int m0, int m1, int m2, const double* x,
                                  double* dense, long ld, long n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= pr.nblk) return;
  const int a1 = pr.at1[k], a2 = pr.at2[k];
  const int n1 = norbAtom[a1], n2 = norbAtom[a2];
  const long o1 = iAtomStart[a1], o2 = iAtomStart[a2];
  const long st = pr.start[k];
  double sn, cs;
  block_phase(blkT, k, kx, ky, kz, real2k, m0, m1, m2, sn, cs);
  for (int c = 0; c < n1; ++c)
    for (int r = 0; r < n2; ++r) {
      if (a1 == a2 && r < c) continue;
      const double v = x[st + r + c * n2];
      const double re = v * cs, im = v * sn;
      const long i = o2 + r, j = o1 + c;
      if (real2k) {
        atomicAdd(&dense[i + j * ld], re);
        if (i != j) atomicAdd(&dense[j + i * ld], re);
        continue;
      }
      atomicAdd(&dense[i + j * ld], re);
      atomicAdd(&dense[(n + i) + (n + j) * ld], re);
      if (i == j) continue;
      atomicAdd(&dense[(n + i) + j * ld], im);
      atomicAdd(&dense[i + (n + j) * ld], -im);
      atomicAdd(&dense[j + i * ld], re);
      atomicAdd(&dense[(n + j) + (n + i) * ld], re);
      atomicAdd(&dense[(n + j) + i * ld], -im);
      atomicAdd(&dense[j + (n + i) * ld], im);
    }
}

// x += w Re(exp(-i 2 pi kf.T) P) per block, P = TL + i BL read from the lower triangle of the
// embedded 2n x 2n matrix; on-site blocks completed from their lower triangle by conjugation
// (packHS_cmplx_kpts, sparse2dense.F90:820-918)
__global__ void pack_kpt_kernel(PairsDev pr, const int* norbAtom, const long* iAtomStart,
                                const double* blkT, double kx, double ky, double kz, int real2k,
                                int m0, int m1, int m2, double w, const double* dense, long ld,
                                long n, double* x) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= pr.nblk) return;
  const int a1 = pr.at1[k], a2 = pr.at2[k];
  const int n1 = norbAtom[a1], n2 = norbAtom[a2];
  const long o1 = iAtomStart[a1], o2 = iAtomStart[a2];
  const long st = pr.start[k];
  double sn, cs;
  block_phase(blkT, k, kx, ky, kz, real2k, m0, m1, m2, sn, cs);
  for (int c = 0; c < n1; ++c)
    for (int r = 0; r < n2; ++r) {
      double re, im = 0.0;
      if (a1 == a2 && r < c) {
        const long i = o2 + c, j = o1 + r;
        re = dense[i + j * ld];
        if (!real2k) im = -dense[(n + i) + j * ld];
      } else {
        const long i = o2 + r, j = o1 + c;
        re = dense[i + j * ld];
        if (!real2k) im = dense[(n + i) + j * ld];
      }
      x[st + r + c * n2] += w * (re * cs + im * sn);
    }
}

__global__ void shift_kernel(PairsDev pr, const int* norbAtom, const long* iAtomStart,
                             const double* h0, const double* s, const double* vorb, double* h) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int k = (int)(t / kBlkLanes), e = (int)(t % kBlkLanes);
  if (k >= pr.nblk) return;
  const int a1 = pr.at1[k], a2 = pr.at2[k];
  const int n1 = norbAtom[a1], n2 = norbAtom[a2];
  if (e >= n1 * n2) return;
  const int r = e % n2, c = e / n2;
  const long o1 = iAtomStart[a1], o2 = iAtomStart[a2];
  const long i = pr.start[k] + e;
  h[i] = h0[i] + 0.5 * s[i] * (vorb[o2 + r] + vorb[o1 + c]);
}

// q_mu += sum_nu S rho (column sums to atom 1, row sums to atom 2 for off-site blocks);
// the same contraction with H0 gives the non-SCC energy per orbital. A block's column and row
// sums are gathered by shuffles inside its lane group (every lane takes part), then one lane
// per column (r = 0) and per row (c = 0) adds them.
__global__ void mulliken_kernel(PairsDev pr, const int* norbAtom, const long* iAtomStart,
                                const double* s, const double* h0, const double* rho,
                                double* qorb, double* eorb) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int k = (int)(t / kBlkLanes), e = (int)(t % kBlkLanes);
  const int lane = threadIdx.x & 63;
  int a1 = 0, a2 = 0, n1 = 1, n2 = 1;
  double pq = 0.0, pe = 0.0;
  const bool blk = k < pr.nblk;
  if (blk) {
    a1 = pr.at1[k];
    a2 = pr.at2[k];
    n1 = norbAtom[a1];
    n2 = norbAtom[a2];
    if (e < n1 * n2) {
      const long i = pr.start[k] + e;
      pq = s[i] * rho[i];
      pe = h0 ? h0[i] * rho[i] : 0.0;
    }
  }
  const bool valid = blk && e < n1 * n2;
  const int r = e % n2, c = e / n2;
  double cq = 0.0, ce = 0.0, rq = 0.0, re = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    // column c: lanes e + j (r = 0 lanes), row r: lanes e + j n2 (c = 0 lanes)
    const double vq = __shfl(pq, min(lane + j, 63), 64);
    const double ve = __shfl(pe, min(lane + j, 63), 64);
    const double wq = __shfl(pq, min(lane + j * n2, 63), 64);
    const double we = __shfl(pe, min(lane + j * n2, 63), 64);
    if (j < n2) {
      cq += vq;
      ce += ve;
    }
    if (j < n1) {
      rq += wq;
      re += we;
    }
  }
  if (!valid) return;
  const long o1 = iAtomStart[a1], o2 = iAtomStart[a2];
  if (r == 0) {
    atomicAdd(&qorb[o1 + c], cq);
    if (eorb) atomicAdd(&eorb[o1 + c], ce);
  }
  if (c == 0 && a1 != a2) {
    atomicAdd(&qorb[o2 + r], rq);
    if (eorb) atomicAdd(&eorb[o2 + r], re);
  }
}

// Pair forces: gradient contribution of every off-site block. mode 0: central finite
// differences with step delta on the bond vector (the reference default); mode 1: analytic
// derivative of the interpolated and rotated blocks (the adjoint of the assembly).
__global__ void pair_force_kernel(SkDev sk, PairsDev pr, const int* norbAtom,
                                  const long* iAtomStart, const double* coords, const double* rho,
                                  const double* erho, const double* vorb, int mode, double delta,
                                  double* grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= pr.nblk) return;
  const int a1 = pr.at1[k], i2 = pr.img2[k], a2 = pr.at2[k];
  if (a1 == a2) return;
  const int sp1 = pr.species[a1], sp2 = pr.species[a2];
  const int n1 = norbAtom[a1], n2 = norbAtom[a2];
  const long o1 = iAtomStart[a1], o2 = iAtomStart[a2];
  const long st = pr.start[k];
  double vec[3];
  for (int c = 0; c < 3; ++c) vec[c] = coords[3 * (long)i2 + c] - coords[3 * (long)a1 + c];
  double g[3] = {0, 0, 0};
  if (mode == 0) {
    for (int c = 0; c < 3; ++c) {
      double hp[16], hm[16], sp[16], sm[16];
      double vp[3] = {vec[0], vec[1], vec[2]}, vm[3] = {vec[0], vec[1], vec[2]};
      vp[c] += delta;
      vm[c] -= delta;
      pair_block(sk, sp1, sp2, true, vp, hp, nullptr);
      pair_block(sk, sp1, sp2, true, vm, hm, nullptr);
      pair_block(sk, sp1, sp2, false, vp, sp, nullptr);
      pair_block(sk, sp1, sp2, false, vm, sm, nullptr);
      double acc = 0.0;
      for (int cc = 0; cc < n1; ++cc)
        for (int r = 0; r < n2; ++r) {
          const long e = st + r + cc * n2;
          const double dH = (hm[r * 4 + cc] - hp[r * 4 + cc]) / (2.0 * delta);
          const double dS = (sm[r * 4 + cc] - sp[r * 4 + cc]) / (2.0 * delta);
          const double vs = 0.5 * (vorb[o2 + r] + vorb[o1 + cc]);
          acc += rho[e] * dH - erho[e] * dS + rho[e] * vs * dS;
        }
      g[c] = 2.0 * acc;
    }
  } else {
    double hh[16], dh[48], ss[16], ds[48];
    pair_block(sk, sp1, sp2, true, vec, hh, dh);
    pair_block(sk, sp1, sp2, false, vec, ss, ds);
    for (int c = 0; c < 3; ++c) {
      double acc = 0.0;
      for (int cc = 0; cc < n1; ++cc)
        for (int r = 0; r < n2; ++r) {
          const long e = st + r + cc * n2;
          // derivative w.r.t. atom 1 = - d/dR
          const double dH = -dh[c * 16 + r * 4 + cc];
          const double dS = -ds[c * 16 + r * 4 + cc];
          const double vs = 0.5 * (vorb[o2 + r] + vorb[o1 + cc]);
          acc += rho[e] * dH - erho[e] * dS + rho[e] * vs * dS;
        }
      g[c] = 2.0 * acc;
    }
  }
  for (int c = 0; c < 3; ++c) {
    atomicAdd(&grad[3 * (long)a1 + c], g[c]);
    atomicAdd(&grad[3 * (long)a2 + c], -g[c]);
  }
}

__global__ void invr_cluster_kernel(int nat, const double* coords, double* g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nat) return;
  double v = 0.0;
  if (i != j) {
    const double dx = coords[3 * i] - coords[3 * j], dy = coords[3 * i + 1] - coords[3 * j + 1],
                 dz = coords[3 * i + 2] - coords[3 * j + 2];
    v = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
  }
  g[i + (long)j * nat] = v;
}

// Ewald real-space sum over translations, plus the constant -pi/(V alpha^2) and the self term.
__global__ void ewald_real_kernel(int nat, const double* coords, int ntr, const double* tr,
                                  double alpha, double maxr, double cst, double self, double* g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nat) return;
  const double dx = coords[3 * i] - coords[3 * j], dy = coords[3 * i + 1] - coords[3 * j + 1],
               dz = coords[3 * i + 2] - coords[3 * j + 2];
  double acc = 0.0;
  for (int t = 0; t < ntr; ++t) {
    const double x = dx - tr[3 * t], y = dy - tr[3 * t + 1], z = dz - tr[3 * t + 2];
    const double r = sqrt(x * x + y * y + z * z);
    if (r <= maxr && r > 1.0e-8) acc += erfc(alpha * r) / r;
  }
  acc += cst;
  if (i == j) acc += self;
  g[i + (long)j * nat] += acc;
}

// structure factor matrices C(i,g) = cos(G.r_i), S(i,g) = sin(G.r_i)
__global__ void strucfac_kernel(int nat, const double* coords, int ng, const double* G,
                                double* cs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y;
  if (i >= nat) return;
  const double ph =
      G[3 * g] * coords[3 * i] + G[3 * g + 1] * coords[3 * i + 1] + G[3 * g + 2] * coords[3 * i + 2];
  double sn, cn;
  sincos(ph, &sn, &cn);
  cs[i + (long)g * nat] = cn;
  cs[i + (long)(g + ng) * nat] = sn;
}

// gamma(jc, i) and gamma(i, jc) -= expGamma for every neighbour within its cutoff
__global__ void short_gamma_kernel(NeighDev nb, const double* coords, const double* hubbu,
                                   const double* gcut, int nsp, int nat, double* g) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb.n) return;
  const int i = nb.at1[k], j = nb.img2[k], jc = nb.at2[k];
  const int si = nb.species[i], sj = nb.species[jc];
  const double dx = coords[3 * (long)i] - coords[3 * (long)j],
               dy = coords[3 * (long)i + 1] - coords[3 * (long)j + 1],
               dz = coords[3 * (long)i + 2] - coords[3 * (long)j + 2];
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  if (j != i && r > gcut[sj * nsp + si]) return;
  const double v = expgamma(r, hubbu[sj], hubbu[si]);
  atomicAdd(&g[jc + (long)i * nat], -v);
  if (jc != i) atomicAdd(&g[i + (long)jc * nat], -v);
}

// gradient of 1/2 dq^T gamma dq from the short-range part and (cluster) 1/R part per neighbour
__global__ void short_gamma_grad_kernel(NeighDev nb, const double* coords, const double* hubbu,
                                        const double* gcut, int nsp, const double* dq,
                                        double* grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb.n) return;
  const int i = nb.at1[k], j = nb.img2[k], jc = nb.at2[k];
  if (jc == i) return;
  const int si = nb.species[i], sj = nb.species[jc];
  const double d[3] = {coords[3 * (long)i] - coords[3 * (long)j],
                       coords[3 * (long)i + 1] - coords[3 * (long)j + 1],
                       coords[3 * (long)i + 2] - coords[3 * (long)j + 2]};
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (r > gcut[sj * nsp + si]) return;
  const double f = -expgamma_prime(r, hubbu[sj], hubbu[si]) * dq[i] * dq[jc] / r;
  for (int c = 0; c < 3; ++c) {
    atomicAdd(&grad[3 * (long)i + c], f * d[c]);
    atomicAdd(&grad[3 * (long)jc + c], -f * d[c]);
  }
}

// cluster 1/R gradient: one thread per atom i, loop over j
__global__ void invr_cluster_grad_kernel(int nat, const double* coords, const double* dq,
                                         double* grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nat) return;
  double g[3] = {0, 0, 0};
  for (int j = 0; j < nat; ++j) {
    if (j == i) continue;
    const double d[3] = {coords[3 * i] - coords[3 * j], coords[3 * i + 1] - coords[3 * j + 1],
                         coords[3 * i + 2] - coords[3 * j + 2]};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double f = -dq[i] * dq[j] / (r2 * sqrt(r2));
    for (int c = 0; c < 3; ++c) g[c] += f * d[c];
  }
  for (int c = 0; c < 3; ++c) grad[3 * i + c] += g[c];
}

// Ewald gradient: real space over translations + reciprocal from structure factors
// Ewald gradient of atom i (one 256-thread workgroup per atom): real-space sum over all atoms
// and lattice translations within maxr, reciprocal sum over the G vectors with the structure
// factors C_G = sum_j dq_j cos(G r_j), S_G = sum_j dq_j sin(G r_j) (CS = [C; S])
__global__ __launch_bounds__(256) void ewald_grad_kernel(int nat, const double* coords, int ntr,
                                                         const double* tr, double alpha,
                                                         double maxr, int ng, const double* G,
                                                         const double* w, const double* CS,
                                                         const double* dq, double* grad) {
  __shared__ double red[3][4];
  const int i = blockIdx.x;
  if (i >= nat) return;
  const int tid = threadIdx.x;
  const double xi = coords[3 * i], yi = coords[3 * i + 1], zi = coords[3 * i + 2];
  const double ta = 2.0 * alpha / sqrt(M_PI);
  const double maxr2 = maxr * maxr;
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;
  for (int j = tid; j < nat; j += 256) {
    const double dx = xi - coords[3 * j], dy = yi - coords[3 * j + 1], dz = zi - coords[3 * j + 2];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int t = 0; t < ntr; ++t) {
      const double x = dx - tr[3 * t], y = dy - tr[3 * t + 1], z = dz - tr[3 * t + 2];
      const double r2 = x * x + y * y + z * z;
      if (r2 <= maxr2 && r2 > 1.0e-16) {
        const double r = sqrt(r2);
        const double dfdr = -(ta * exp(-alpha * alpha * r2) / r + erfc(alpha * r) / r2);
        const double f = dfdr / r;
        a0 += f * x;
        a1 += f * y;
        a2 += f * z;
      }
    }
    g0 += a0 * dq[j];
    g1 += a1 * dq[j];
    g2 += a2 * dq[j];
  }
  for (int q = tid; q < ng; q += 256) {
    const double ph = G[3 * q] * xi + G[3 * q + 1] * yi + G[3 * q + 2] * zi;
    double sn, cn;
    sincos(ph, &sn, &cn);
    const double fac = (sn * CS[q] - cn * CS[ng + q]) * w[q];
    g0 -= fac * G[3 * q];
    g1 -= fac * G[3 * q + 1];
    g2 -= fac * G[3 * q + 2];
  }
  for (int o = 32; o > 0; o >>= 1) {
    g0 += __shfl_down(g0, o, 64);
    g1 += __shfl_down(g1, o, 64);
    g2 += __shfl_down(g2, o, 64);
  }
  const int lane = tid & 63, wv = tid >> 6;
  if (lane == 0) {
    red[0][wv] = g0;
    red[1][wv] = g1;
    red[2][wv] = g2;
  }
  __syncthreads();
  if (tid < 3)
    grad[3 * i + tid] += dq[i] * (red[tid][0] + red[tid][1] + red[tid][2] + red[tid][3]);
}

// repulsive energy and gradient over neighbours within the repulsive cutoff
__device__ void rep_value(const RepDev& rp, double r, double* e, double* de) {
  *e = 0.0;
  *de = 0.0;
  if (r < 1.0e-2 || r >= rp.cutoff) return;
  if (r < rp.xs[0]) {
    const double ex = exp(-rp.expc[0] * r + rp.expc[1]);
    *e = ex + rp.expc[2];
    *de = -rp.expc[0] * ex;
    return;
  }
  int im = 0;
  while (im + 1 < rp.nint && rp.xs[im + 1] <= r) ++im;
  const double dr = r - rp.xs[im];
  const double* c = (im < rp.nint - 1) ? rp.coef + 4 * im : rp.last;
  const int nc = (im < rp.nint - 1) ? 4 : 6;
  double en = 0.0, dv = 0.0;
  for (int q = nc - 1; q >= 0; --q) en = en * dr + c[q];
  for (int q = nc - 1; q >= 1; --q) dv = dv * dr + q * c[q];
  *e = en;
  *de = dv;
}

__global__ void repulsive_kernel(NeighDev nb, const double* coords, const RepDev* reps, int nsp,
                                 double* eatom, double* grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb.n) return;
  const int i = nb.at1[k], j = nb.img2[k], jc = nb.at2[k];
  if (j == i) return;
  const double d[3] = {coords[3 * (long)i] - coords[3 * (long)j],
                       coords[3 * (long)i + 1] - coords[3 * (long)j + 1],
                       coords[3 * (long)i + 2] - coords[3 * (long)j + 2]};
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const RepDev& rp = reps[nb.species[i] * nsp + nb.species[jc]];
  if (r >= rp.cutoff) return;
  double e, de;
  rep_value(rp, r, &e, &de);
  atomicAdd(&eatom[i], 0.5 * e);
  if (jc != i) {
    atomicAdd(&eatom[jc], 0.5 * e);
    for (int c = 0; c < 3; ++c) {
      atomicAdd(&grad[3 * (long)i + c], de * d[c] / r);
      atomicAdd(&grad[3 * (long)jc + c], -de * d[c] / r);
    }
  }
}

// ------------------------------------------------------------------------------------------
// launch wrappers
// ------------------------------------------------------------------------------------------
int launch_build_h0s(hipStream_t s, const SkDev& sk, const PairsDev& pr, const double* coords,
                     double* h0, double* ov) {
  if (pr.nblk == 0) return 0;
  hipLaunchKernelGGL(build_h0s_kernel, dim3(cdiv(pr.nblk, 128)), dim3(128), 0, s, sk, pr, coords,
                     h0, ov);
  DH_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int launch_unpack(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                  const double* x, T* dense, long ld, int n) {
  DH_TRY(fill2d<T>(s, dense, ld, n, n, T(0)));
  if (pr.nblk == 0) return 0;
  hipLaunchKernelGGL((unpack_kernel<T>), dim3(cdiv((long)pr.nblk * kBlkLanes, 256)),
                     dim3(256), 0, s, pr, norb,
                     iAtomStart, x, dense, ld);
  DH_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int launch_pack(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                const T* dense, long ld, double* x) {
  if (pr.nblk == 0) return 0;
  hipLaunchKernelGGL((pack_kernel<T>), dim3(cdiv((long)pr.nblk * kBlkLanes, 256)),
                     dim3(256), 0, s, pr, norb,
                     iAtomStart, dense, ld, x);
  DH_LAUNCH_CHECK();
  return 0;
}

bool kpoint_is_real(const double* kf) {
  for (int q = 0; q < 3; ++q)
    if (std::fabs(2.0 * kf[q] - std::nearbyint(2.0 * kf[q])) > 1e-12) return false;
  return true;
}

int kpoint_dim(const double* kf, int n) { return kpoint_is_real(kf) ? n : 2 * n; }

int launch_unpack_kpt(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                      const double* blkT, const double* kf, const double* x, double* dense,
                      long ld, int n) {
  const int nd = kpoint_dim(kf, n);
  DH_CHECK(ld >= nd, "unpack_kpt: leading dimension too small");
  DH_TRY(fill2d<double>(s, dense, ld, nd, nd, 0.0));
  if (pr.nblk == 0) return 0;
  const int re = kpoint_is_real(kf) ? 1 : 0;
  hipLaunchKernelGGL(unpack_kpt_kernel, dim3(cdiv(pr.nblk, 128)), dim3(128), 0, s, pr, norb,
                     iAtomStart, blkT, kf[0], kf[1], kf[2], re, (int)std::nearbyint(2.0 * kf[0]),
                     (int)std::nearbyint(2.0 * kf[1]), (int)std::nearbyint(2.0 * kf[2]), x,
                     dense, ld, (long)n);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_pack_kpt(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                    const double* blkT, const double* kf, double w, const double* dense, long ld,
                    int n, double* x) {
  if (pr.nblk == 0) return 0;
  DH_CHECK(ld >= kpoint_dim(kf, n), "pack_kpt: leading dimension too small");
  const int re = kpoint_is_real(kf) ? 1 : 0;
  hipLaunchKernelGGL(pack_kpt_kernel, dim3(cdiv(pr.nblk, 128)), dim3(128), 0, s, pr, norb,
                     iAtomStart, blkT, kf[0], kf[1], kf[2], re, (int)std::nearbyint(2.0 * kf[0]),
                     (int)std::nearbyint(2.0 * kf[1]), (int)std::nearbyint(2.0 * kf[2]), w,
                     dense, ld, (long)n, x);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_shift(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                 const double* h0, const double* ov, const double* vorb, double* h) {
  if (pr.nblk == 0) return 0;
  hipLaunchKernelGGL(shift_kernel, dim3(cdiv((long)pr.nblk * kBlkLanes, 256)),
                     dim3(256), 0, s, pr, norb,
                     iAtomStart, h0, ov, vorb, h);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_mulliken(hipStream_t s, const PairsDev& pr, const int* norb, const long* iAtomStart,
                    const double* ov, const double* h0, const double* rho, double* qorb,
                    double* eorb, int nOrb) {
  DH_TRY(fill<double>(s, qorb, nOrb, 0.0));
  if (eorb) DH_TRY(fill<double>(s, eorb, nOrb, 0.0));
  if (pr.nblk == 0) return 0;
  hipLaunchKernelGGL(mulliken_kernel, dim3(cdiv((long)pr.nblk * kBlkLanes, 256)),
                     dim3(256), 0, s, pr, norb,
                     iAtomStart, ov, h0, rho, qorb, eorb);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_pair_forces(hipStream_t s, const SkDev& sk, const PairsDev& pr, const int* norb,
                       const long* iAtomStart, const double* coords, const double* rho,
                       const double* erho, const double* vorb, int mode, double delta,
                       double* grad) {
  if (pr.nblk == 0) return 0;
  hipLaunchKernelGGL(pair_force_kernel, dim3(cdiv(pr.nblk, 64)), dim3(64), 0, s, sk, pr, norb,
                     iAtomStart, coords, rho, erho, vorb, mode, delta, grad);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_invr_cluster(hipStream_t s, int nat, const double* coords, double* g) {
  hipLaunchKernelGGL(invr_cluster_kernel, dim3(cdiv(nat, 256), nat), dim3(256), 0, s, nat, coords,
                     g);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_ewald_real(hipStream_t s, int nat, const double* coords, int ntr, const double* tr,
                      double alpha, double maxr, double cst, double self, double* g) {
  hipLaunchKernelGGL(ewald_real_kernel, dim3(cdiv(nat, 256), nat), dim3(256), 0, s, nat, coords,
                     ntr, tr, alpha, maxr, cst, self, g);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_strucfac(hipStream_t s, int nat, const double* coords, int ng, const double* G,
                    double* cs) {
  if (ng == 0) return 0;
  hipLaunchKernelGGL(strucfac_kernel, dim3(cdiv(nat, 256), ng), dim3(256), 0, s, nat, coords, ng,
                     G, cs);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_short_gamma(hipStream_t s, const NeighDev& nb, const double* coords,
                       const double* hubbu, const double* gcut, int nsp, int nat, double* g) {
  if (nb.n == 0) return 0;
  hipLaunchKernelGGL(short_gamma_kernel, dim3(cdiv(nb.n, 256)), dim3(256), 0, s, nb, coords, hubbu,
                     gcut, nsp, nat, g);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_short_gamma_grad(hipStream_t s, const NeighDev& nb, const double* coords,
                            const double* hubbu, const double* gcut, int nsp, const double* dq,
                            double* grad) {
  if (nb.n == 0) return 0;
  hipLaunchKernelGGL(short_gamma_grad_kernel, dim3(cdiv(nb.n, 256)), dim3(256), 0, s, nb, coords,
                     hubbu, gcut, nsp, dq, grad);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_gamma_grad(hipStream_t s, const NeighDev& nb, const double* coords,
                      const double* hubbu, const double* gcut, int nsp, int nat, bool periodic,
                      int ntr, const double* tr, double alpha, double maxr, int ng,
                      const double* G, const double* w, const double* CSq, const double* dq,
                      double* grad) {
  if (nb.n > 0) {
    hipLaunchKernelGGL(short_gamma_grad_kernel, dim3(cdiv(nb.n, 256)), dim3(256), 0, s, nb, coords,
                       hubbu, gcut, nsp, dq, grad);
    DH_LAUNCH_CHECK();
  }
  if (!periodic) {
    hipLaunchKernelGGL(invr_cluster_grad_kernel, dim3(cdiv(nat, 64)), dim3(64), 0, s, nat, coords,
                       dq, grad);
  } else {
    hipLaunchKernelGGL(ewald_grad_kernel, dim3(nat), dim3(256), 0, s, nat, coords, ntr, tr, alpha,
                       maxr, ng, G, w, CSq, dq, grad);
  }
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_repulsive(hipStream_t s, const NeighDev& nb, const double* coords, const RepDev* reps,
                     int nsp, double* eatom, double* grad) {
  if (nb.n == 0) return 0;
  hipLaunchKernelGGL(repulsive_kernel, dim3(cdiv(nb.n, 256)), dim3(256), 0, s, nb, coords, reps,
                     nsp, eatom, grad);
  DH_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Block-sparse products with a transposed dense operand: YH^T = (H X)^T and YS^T = (S X)^T for
// X given as Xt = X^T (nvec x n, column-major, so the nvec coefficients of one orbital are
// contiguous). H and S are the matrices unpack builds from the sparse blocks (image blocks of one
// atom pair summed).
//  * spmm_fold_kernel: once per H and S, the row view's groups (row atom A, column atom B) become
//    dense 4 x 4 blocks of both matrices (image blocks summed with their k-point phases +-1,
//    transposed where the stored block is (B, A), zero-padded past the atoms' orbitals), 32
//    consecutive doubles per group in row order, with the column atom's first orbital beside;
//  * spmm_fold_t_kernel: a workgroup of 4 waves owns one row atom and 256 kVpl consecutive
//    vectors (kVpl per lane, 64 apart). Per group the wave reads the block's 32 values as
//    wave-uniform scalar loads from contiguous memory, the 4 coefficient rows of the column atom
//    (kVpl coalesced loads each) and does 32 kVpl FMAs; the outputs of the atom's 4 orbitals
//    for both matrices stay in registers. Workgroups are numbered so that consecutive ones of an
//    XCD (IDs congruent mod 8) walk the atoms of one vector tile, so the coefficient rows a tile
//    reads stay in that XCD's L2 while its neighbouring row atoms read them.
// ---------------------------------------------------------------------------------------------
constexpr int kXcds = 8;
static int g_spmmVpl = 1;
int set_spmm_cols(int vpl) {
  DH_CHECK(vpl == 1 || vpl == 2 || vpl == 4, "spmm: 1, 2 or 4 vectors per lane");
  g_spmmVpl = vpl;
  return 0;
}

__global__ __launch_bounds__(256) void spmm_fold_kernel(
    const int* __restrict__ gPtr, const int2* __restrict__ grp, const int2* __restrict__ entS,
    const int* __restrict__ norbAtom, const long* __restrict__ iAtomStart,
    const long* __restrict__ blkStart, const double* __restrict__ h,
    const double* __restrict__ ov, const double* __restrict__ blkT, int m0, int m1, int m2,
    double* __restrict__ fold, int* __restrict__ colOff) {
  const int A = blockIdx.x;
  const int nA = norbAtom[A];
  const int g0 = gPtr[A], g1 = gPtr[A + 1];
  for (int t = threadIdx.x; t < (g1 - g0) * 32; t += blockDim.x) {
    const int g = g0 + (t >> 5), el = t & 31;
    const int mat = el >> 4, r = (el >> 2) & 3, c = el & 3;
    const int2 gr = grp[g];
    const int B = entS[gr.x].y, nB = norbAtom[B];
    const double* __restrict__ src = mat ? ov : h;
    double v = 0.0;
    if (r < nA && c < nB) {
      for (int e = gr.x; e < gr.y; ++e) {
        const int k = entS[e].x >> 1, tr = entS[e].x & 1;
        const long q = blkStart[k] + (tr ? (c + (long)r * nB) : (r + (long)c * nA));
        double x = src[q];
        if (blkT) {
          const long par =
              llround(m0 * blkT[3 * k] + m1 * blkT[3 * k + 1] + m2 * blkT[3 * k + 2]);
          if (par & 1) x = -x;
        }
        v += x;
      }
    }
    fold[(long)g * 32 + el] = v;
    if (el == 0) colOff[g] = (int)iAtomStart[B];
  }
}

template <int kVpl>
__global__ __launch_bounds__(256) void spmm_fold_t_kernel(
    int nat, int ntile, int n, const int* __restrict__ gPtr, const int* __restrict__ colOff,
    const double* __restrict__ fold, const int* __restrict__ norbAtom,
    const long* __restrict__ iAtomStart, const double* __restrict__ Xt, long ldx, int nvec,
    double* __restrict__ yh, double* __restrict__ ys, long ldy) {
  const long w = blockIdx.x;
  const int xcd = (int)(w % kXcds);
  const long l = w / kXcds;
  const int tile = xcd + kXcds * (int)(l / nat);
  const int A = (int)(l % nat);
  if (tile >= ntile) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int jb = tile * 256 * kVpl + wave * 64 * kVpl + lane;
  const double* __restrict__ xj[kVpl];
#pragma unroll
  for (int v = 0; v < kVpl; ++v) xj[v] = Xt + min(jb + 64 * v, nvec - 1);
  double ah[kVpl][4], as[kVpl][4];
#pragma unroll
  for (int v = 0; v < kVpl; ++v)
#pragma unroll
    for (int r = 0; r < 4; ++r) ah[v][r] = as[v][r] = 0.0;
  const int g0 = gPtr[A], g1 = gPtr[A + 1];
  // the coefficient rows of the next group are loaded while the current one is multiplied
  // (rows past the column atom's orbitals meet zero block columns; clamped to stay inside Xt)
  double xn[kVpl][4];
  auto load_x = [&](int g) {
    const int oB = colOff[g];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const long row = min(oB + c, n - 1);
#pragma unroll
      for (int v = 0; v < kVpl; ++v) xn[v][c] = xj[v][row * ldx];
    }
  };
  if (g0 < g1) load_x(g0);
  for (int g = g0; g < g1; ++g) {
    double xv[kVpl][4];
#pragma unroll
    for (int v = 0; v < kVpl; ++v)
#pragma unroll
      for (int c = 0; c < 4; ++c) xv[v][c] = xn[v][c];
    if (g + 1 < g1) load_x(g + 1);
    const double* __restrict__ b = fold + (long)g * 32;
    double bh[16], bs[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      bh[q] = b[q];
      bs[q] = b[16 + q];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < kVpl; ++v) {
          ah[v][r] = fma(bh[r * 4 + c], xv[v][c], ah[v][r]);
          as[v][r] = fma(bs[r * 4 + c], xv[v][c], as[v][r]);
        }
  }
  const int nA = norbAtom[A];
  const long oA = iAtomStart[A];
#pragma unroll
  for (int v = 0; v < kVpl; ++v) {
    const int j = jb + 64 * v;
    if (j >= nvec) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (r >= nA) break;
      yh[j + (oA + r) * ldy] = ah[v][r];
      ys[j + (oA + r) * ldy] = as[v][r];
    }
  }
}

int launch_spmm_fold(hipStream_t s, const RowBlocksDev& rb, const int* norb,
                     const long* iAtomStart, const long* blkStart, const double* h,
                     const double* ov, const double* blkT, const double* kf, double* fold,
                     int* colOff) {
  if (rb.nat == 0) return 0;
  int m[3] = {0, 0, 0};
  if (blkT) {
    DH_CHECK(kf && kpoint_is_real(kf), "spmm: the k-point is not time-reversal invariant");
    for (int q = 0; q < 3; ++q) m[q] = (int)std::nearbyint(2.0 * kf[q]);
  }
  hipLaunchKernelGGL(spmm_fold_kernel, dim3(rb.nat), dim3(256), 0, s, rb.gPtr, rb.grp, rb.entS,
                     norb, iAtomStart, blkStart, h, ov, blkT, m[0], m[1], m[2], fold, colOff);
  DH_LAUNCH_CHECK();
  return 0;
}

int launch_spmm_hs_t(hipStream_t s, const RowBlocksDev& rb, const int* norb,
                     const long* iAtomStart, int n, const double* fold, const int* colOff,
                     const double* Xt, long ldx, int nvec, double* yh, double* ys, long ldy) {
  if (rb.nat == 0 || nvec <= 0) return 0;
  const int vpl = g_spmmVpl;
  const int ntile = (int)cdiv(nvec, 256 * vpl);
  const long tilesPad = (long)cdiv(ntile, kXcds) * kXcds;
  const long nwg = tilesPad * rb.nat;
  DH_CHECK(nwg < (1L << 31), "spmm: grid too large");
#define DH_SPMM(V_)                                                                           \
  hipLaunchKernelGGL(spmm_fold_t_kernel<V_>, dim3((unsigned)nwg), dim3(256), 0, s, rb.nat, ntile, \
                     n, rb.gPtr, colOff, fold, norb, iAtomStart, Xt, ldx, nvec, yh, ys, ldy)
  if (vpl == 1) DH_SPMM(1);
  else if (vpl == 2) DH_SPMM(2);
  else DH_SPMM(4);
#undef DH_SPMM
  DH_LAUNCH_CHECK();
  return 0;
}

// Sparse -> dense from the row view (unpackHS_real, sparse2dense.F90:266-321, both triangles):
// one wave per row atom, a lane per group of entries with one column atom (the image blocks of
// the pair, summed in registers), every element of the atom pair's block written once.
template <typename T>
__global__ __launch_bounds__(256) void unpack_rows_kernel(
    RowBlocksDev rb, const int* __restrict__ norbAtom, const long* __restrict__ iAtomStart,
    const long* __restrict__ blkStart, const double* __restrict__ x, T* __restrict__ dense,
    long ld) {
  const int A = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (A >= rb.nat) return;
  const int nA = norbAtom[A];
  const long oA = iAtomStart[A];
  const int g1 = rb.gPtr[A + 1];
  for (int g = rb.gPtr[A] + lane; g < g1; g += 64) {
    const int2 gr = rb.grp[g];
    const int B = rb.entS[gr.x].y;
    const int nB = norbAtom[B];
    const long oB = iAtomStart[B];
    double v[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) v[r][c] = 0.0;
    for (int p = gr.x; p < gr.y; ++p) {
      const int2 en = rb.entS[p];
      const int k = en.x >> 1, tr = en.x & 1;
      const long st = blkStart[k];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (r >= nA) break;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if (c >= nB) break;
          v[r][c] += x[st + (tr ? (c + (long)r * nB) : (r + (long)c * nA))];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (c >= nB) break;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (r >= nA) break;
        dense[(oA + r) + (oB + c) * ld] = (T)v[r][c];
      }
    }
  }
}

template <typename T>
int launch_unpack_rows(hipStream_t s, const RowBlocksDev& rb, const int* norb,
                       const long* iAtomStart, const long* blkStart, const double* x, T* dense,
                       long ld, int n) {
  DH_TRY(fill2d<T>(s, dense, ld, n, n, T(0)));
  if (rb.nat == 0) return 0;
  hipLaunchKernelGGL((unpack_rows_kernel<T>), dim3(cdiv(rb.nat, 4)), dim3(256), 0, s, rb, norb,
                     iAtomStart, blkStart, x, dense, ld);
  DH_LAUNCH_CHECK();
  return 0;
}
template int launch_unpack_rows<double>(hipStream_t, const RowBlocksDev&, const int*,
                                        const long*, const long*, const double*, double*, long,
                                        int);
template int launch_unpack_rows<float>(hipStream_t, const RowBlocksDev&, const int*, const long*,
                                       const long*, const double*, float*, long, int);

// Mulliken populations and band-structure energy per orbital (mullikenPerOrbital,
// populations.F90:120-193: the block of a pair (atom1, atom2) adds its column sums to atom1's
// orbitals and, when atom1 != atom2 as central atoms, its row sums to atom2's) from the row view:
// one wave per row atom, a lane per entry, the per-orbital partial sums combined by an xor
// butterfly, so every orbital is written once, without atomics, and the result does not depend
// on the order in which the hardware runs the waves. In the row view an entry (k, t, B) of row
// atom A holds element (r of A, c of B) at r + c nA (t = 0, pair (B, A)) or c + r nB (t = 1,
// pair (A, B)); a pair of an atom with itself or its image (t = 0, B = A) adds its column sums,
// i.e. element (c, r) of the stored block.
__global__ __launch_bounds__(256) void mulliken_rows_kernel(
    RowBlocksDev rb, const int* __restrict__ norbAtom, const long* __restrict__ iAtomStart,
    const long* __restrict__ blkStart, const double* __restrict__ ov,
    const double* __restrict__ h0, const double* __restrict__ rho, double* __restrict__ qorb,
    double* __restrict__ eorb) {
  const int A = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (A >= rb.nat) return;
  const int nA = norbAtom[A];
  double q[4] = {0.0, 0.0, 0.0, 0.0}, e[4] = {0.0, 0.0, 0.0, 0.0};
  const int p1 = rb.rowPtr[A + 1];
  for (int p = rb.rowPtr[A] + lane; p < p1; p += 64) {
    const int2 en = rb.ent[p];
    const int k = en.x >> 1, tr = en.x & 1, B = en.y;
    const int nB = norbAtom[B];
    const long st = blkStart[k];
    const bool self = !tr && B == A;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (r >= nA) break;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c >= nB) break;
        const long idx = st + (tr ? (c + (long)r * nB) : (self ? (c + (long)r * nA)
                                                              : (r + (long)c * nA)));
        const double d = rho[idx];
        q[r] = fma(ov[idx], d, q[r]);
        if (h0) e[r] = fma(h0[idx], d, e[r]);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
    for (int o = 32; o > 0; o >>= 1) {
      q[r] += __shfl_xor(q[r], o, 64);
      e[r] += __shfl_xor(e[r], o, 64);
    }
  if (lane == 0) {
    const long oA = iAtomStart[A];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (r >= nA) break;
      qorb[oA + r] = q[r];
      if (eorb) eorb[oA + r] = e[r];
    }
  }
}

int launch_mulliken_rows(hipStream_t s, const RowBlocksDev& rb, const int* norb,
                         const long* iAtomStart, const long* blkStart, const double* ov,
                         const double* h0, const double* rho, double* qorb, double* eorb) {
  if (rb.nat == 0) return 0;
  hipLaunchKernelGGL(mulliken_rows_kernel, dim3(cdiv(rb.nat, 4)), dim3(256), 0, s, rb, norb,
                     iAtomStart, blkStart, ov, h0, rho, qorb, eorb);
  DH_LAUNCH_CHECK();
  return 0;
}

template int launch_unpack<double>(hipStream_t, const PairsDev&, const int*, const long*,
                                   const double*, double*, long, int);
template int launch_unpack<float>(hipStream_t, const PairsDev&, const int*, const long*,
                                  const double*, float*, long, int);
template int launch_pack<double>(hipStream_t, const PairsDev&, const int*, const long*,
                                 const double*, long, double*);
template int launch_pack<float>(hipStream_t, const PairsDev&, const int*, const long*,
                                const float*, long, double*);

}  // namespace dftbhip
