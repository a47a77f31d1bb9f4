// Neighbour list and sparse block descriptor on the device: the half list of
// updateNeighbourList (src/dftbp/dftb/periodic.F90:446-696) and the block offsets of
// getSparseDescriptor (src/dftbp/dftb/sparse2dense.F90:3617-3662), in the order the host
// builder and the reference produce.
//
//   candidate images (translation t, atom a) -> index t * nat + a, binned on a grid of cells of
//   one cutoff over the atoms' bounding box grown by the cutoff (counting sort);
//   one thread per central atom counts, then writes, the candidates of the 27 surrounding
//   cells with a >= atom and d^2 <= cutoff^2; one workgroup per atom sorts them in LDS by
//   (d^2, index): the index order is (translation, atom);
//   images in use get numbers after the central atoms by a scan over the candidate index;
//   pairs within the SK cutoff and their block sizes are scanned into the block list.
//
// Squared distances are formed as the host builder forms them (differences, then a sum of
// squares from x to z, no fused multiply-add), so the sort keys and cutoff tests agree bit
// for bit.
#include <climits>
#include <cmath>

#include "neigh.h"

namespace dftbhip {

constexpr int kNlSortMax = 4096;    // neighbours of one atom sorted in LDS
constexpr double kNlMinDist2 = 1.0e-4;   // (0.01 bohr)^2: atoms closer are an input error

void cell_translations(const double* lat, double cutoff, std::vector<double>& tr) {
  tr.assign(3, 0.0);
  if (!lat) return;
  // reciprocal vectors / 2 pi: rows of inv(lat)^T
  const double* a = lat;
  const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
  double inv[9];
  inv[0] = (a[4] * a[8] - a[5] * a[7]) / det;
  inv[1] = -(a[1] * a[8] - a[2] * a[7]) / det;
  inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
  inv[3] = -(a[3] * a[8] - a[5] * a[6]) / det;
  inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
  inv[5] = -(a[0] * a[5] - a[2] * a[3]) / det;
  inv[6] = (a[3] * a[7] - a[4] * a[6]) / det;
  inv[7] = -(a[0] * a[7] - a[1] * a[6]) / det;
  inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
  int nmax[3];
  for (int i = 0; i < 3; ++i) {
    // row i of inv^T = column i of inv
    const double nr = std::sqrt(inv[i] * inv[i] + inv[3 + i] * inv[3 + i] + inv[6 + i] * inv[6 + i]);
    nmax[i] = (int)std::ceil(cutoff * nr) + 1;
  }
  for (int x = -nmax[0]; x <= nmax[0]; ++x)
    for (int y = -nmax[1]; y <= nmax[1]; ++y)
      for (int z = -nmax[2]; z <= nmax[2]; ++z) {
        if (x == 0 && y == 0 && z == 0) continue;
        for (int c = 0; c < 3; ++c) tr.push_back(x * lat[c] + y * lat[3 + c] + z * lat[6 + c]);
      }
}

struct NlGrid {
  double lo[3], hi[3];
  double cutoff;
  int nbin[3];
};

__device__ __forceinline__ double dist2_xyz(const double* x, const double* y) {
#pragma clang fp contract(off)
  double d2 = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double d = x[c] - y[c];
    d2 += d * d;
  }
  return d2;
}

__device__ __forceinline__ void bin_of(const NlGrid& g, const double* x, int* b) {
  for (int c = 0; c < 3; ++c) {
    const int v = (int)floor((x[c] - (g.lo[c] - g.cutoff)) / g.cutoff);
    b[c] = min(max(v, 0), g.nbin[c] - 1);
  }
}

// Exclusive scan of n values in one workgroup of 1024 threads (chunks of 1024, wave scans by
// shuffles, wave totals in LDS); the sum goes to *total.
template <typename T>
__global__ __launch_bounds__(1024) void scan_exclusive_kernel(const T* in, T* out, long n,
                                                              T* total) {
  __shared__ T wsum[16], wpre[16];
  __shared__ T carry;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = T(0);
  __syncthreads();
  for (long base = 0; base < n; base += 1024) {
    const long i = base + threadIdx.x;
    const T v = i < n ? in[i] : T(0);
    T x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const T y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
      T s = lane < 16 ? wsum[lane] : T(0);
      for (int o = 1; o < 16; o <<= 1) {
        const T y = __shfl_up(s, o, 64);
        if (lane >= o) s += y;
      }
      if (lane < 16) wpre[lane] = s - wsum[lane];
    }
    __syncthreads();
    const T excl = carry + wpre[w] + (x - v);
    if (i < n) out[i] = excl;
    __syncthreads();
    if (threadIdx.x == 1023) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

template <typename T>
static int scan_exclusive(hipStream_t s, const T* in, T* out, long n, T* total) {
  hipLaunchKernelGGL(scan_exclusive_kernel<T>, dim3(1), dim3(1024), 0, s, in, out, n, total);
  DH_LAUNCH_CHECK();
  return 0;
}

__global__ void nl_points_kernel(int nat, long P, const double* coords, const double* tr,
                                 NlGrid g, double* px, int* pbin, int* binCount) {
  const long pi = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (pi >= P) return;
  const int t = (int)(pi / nat), a = (int)(pi - (long)t * nat);
  double x[3];
  bool in = true;
  for (int c = 0; c < 3; ++c) {
    x[c] = coords[3 * a + c] + tr[3 * t + c];
    if (x[c] < g.lo[c] - g.cutoff || x[c] > g.hi[c] + g.cutoff) in = false;
    px[3 * pi + c] = x[c];
  }
  int bin = -1;
  if (in) {
    int b[3];
    bin_of(g, x, b);
    bin = (b[0] * g.nbin[1] + b[1]) * g.nbin[2] + b[2];
    atomicAdd(&binCount[bin], 1);
  }
  pbin[pi] = bin;
}

__global__ void nl_scatter_kernel(long P, const int* pbin, int* binFill, int* binItems) {
  const long pi = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (pi >= P) return;
  const int b = pbin[pi];
  if (b < 0) return;
  binItems[atomicAdd(&binFill[b], 1)] = (int)pi;
}

// kFill = false: count the candidates of each central atom (and flag atoms too close);
// kFill = true: write them (squared distance, candidate index) from fStart[atom]
template <bool kFill>
__global__ void nl_search_kernel(int nat, NlGrid g, const double* coords, const double* px,
                                 const int* binStart, const int* binCount, const int* binItems,
                                 double cut2, int* cnt, const long* fStart, double* fD2,
                                 int* fPi, int* err) {
  const int a2 = blockIdx.x * blockDim.x + threadIdx.x;
  if (a2 >= nat) return;
  const double* c2 = coords + 3 * a2;
  int b[3];
  bin_of(g, c2, b);
  const long o = kFill ? fStart[a2] : 0;
  int k = 0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        const int bx = b[0] + dx, by = b[1] + dy, bz = b[2] + dz;
        if (bx < 0 || by < 0 || bz < 0 || bx >= g.nbin[0] || by >= g.nbin[1] || bz >= g.nbin[2])
          continue;
        const int bin = (bx * g.nbin[1] + by) * g.nbin[2] + bz;
        const int q1 = binStart[bin] + binCount[bin];
        for (int q = binStart[bin]; q < q1; ++q) {
          const int pi = binItems[q];
          const int t = pi / nat, a = pi - t * nat;
          if (a < a2) continue;
          const double d2 = dist2_xyz(px + 3 * (long)pi, c2);
          if (d2 > cut2) continue;
          if (d2 < kNlMinDist2) {
            if (t == 0 && a == a2) continue;
            if (!kFill && atomicCAS(&err[0], 0, 1) == 0) {
              err[1] = a + 1;
              err[2] = a2 + 1;
            }
            continue;
          }
          if (kFill) {
            fD2[o + k] = d2;
            fPi[o + k] = pi;
          }
          ++k;
        }
      }
  if (!kFill) cnt[a2] = k;
}

// one workgroup per central atom: bitonic sort of its candidates by (d^2, candidate index) in
// LDS; candidates of other cells (index >= nat) are marked used
__global__ __launch_bounds__(256) void nl_sort_kernel(int nat, const int* cnt, const long* fStart,
                                                      double* fD2, int* fPi, int* used,
                                                      int* err) {
  __shared__ double kd[kNlSortMax];
  __shared__ int kp[kNlSortMax];
  const int a2 = blockIdx.x;
  const int n = cnt[a2];
  if (n > kNlSortMax) {
    if (threadIdx.x == 0) atomicMax(&err[3], n);
    return;
  }
  if (n == 0) return;
  int N2 = 1;
  while (N2 < n) N2 <<= 1;
  const long o = fStart[a2];
  for (int i = threadIdx.x; i < N2; i += blockDim.x) {
    kd[i] = i < n ? fD2[o + i] : INFINITY;
    kp[i] = i < n ? fPi[o + i] : INT_MAX;
  }
  __syncthreads();
  for (int k = 2; k <= N2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < N2; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          const bool gt = kd[i] > kd[l] || (kd[i] == kd[l] && kp[i] > kp[l]);
          if (gt == ((i & k) == 0)) {
            const double td = kd[i];
            kd[i] = kd[l];
            kd[l] = td;
            const int tp = kp[i];
            kp[i] = kp[l];
            kp[l] = tp;
          }
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    fD2[o + i] = kd[i];
    fPi[o + i] = kp[i];
    if (kp[i] >= nat) used[kp[i]] = 1;
  }
}

__global__ void nl_images_kernel(int nat, long P, const double* coords, const double* px,
                                 const int* used, const int* imgRank, double* coordsAll,
                                 int* img2cent) {
  const long pi = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (pi >= P) return;
  if (pi < nat) {
    for (int c = 0; c < 3; ++c) coordsAll[3 * pi + c] = coords[3 * pi + c];
    img2cent[pi] = (int)pi;
  } else if (used[pi]) {
    const long img = nat + imgRank[pi];
    for (int c = 0; c < 3; ++c) coordsAll[3 * img + c] = px[3 * pi + c];
    img2cent[img] = (int)(pi % nat);
  }
}

__global__ void nl_pairs_kernel(int nat, const int* cnt, const long* fStart, const double* fD2,
                                const int* fPi, const int* imgRank, int* pAt1, int* pImg2,
                                int* pAt2, double* pD2, int* pCount, long* pStart) {
  const int a2 = blockIdx.x * blockDim.x + threadIdx.x;
  if (a2 >= nat) return;
  const long o = fStart[a2], base = o + a2;
  const int n = cnt[a2];
  pAt1[base] = a2;
  pImg2[base] = a2;
  pAt2[base] = a2;
  pD2[base] = 0.0;
  pCount[a2] = n;
  pStart[a2] = base;
  if (a2 == nat - 1) pStart[nat] = base + 1 + n;
  for (int j = 0; j < n; ++j) {
    const int pi = fPi[o + j];
    const int t = pi / nat, a = pi - t * nat;
    const long q = base + 1 + j;
    pAt1[q] = a2;
    pImg2[q] = t == 0 ? a : nat + imgRank[pi];
    pAt2[q] = a;
    pD2[q] = fD2[o + j];
  }
}

__global__ void nl_block_flags_kernel(long nPairs, const double* pD2, const int* pAt1,
                                      const int* pAt2, const int* norb, double sk2, int* flag,
                                      long* size) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= nPairs) return;
  const bool in = pD2[q] <= sk2;
  flag[q] = in ? 1 : 0;
  size[q] = in ? (long)norb[pAt1[q]] * norb[pAt2[q]] : 0;
}

__global__ void nl_blocks_kernel(long nPairs, const int* flag, const int* bIdx, const long* bOff,
                                 const int* pAt1, const int* pImg2, const int* pAt2, int* bAt1,
                                 int* bImg2, int* bAt2, long* bStart) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= nPairs || !flag[q]) return;
  const int k = bIdx[q];
  bAt1[k] = pAt1[q];
  bImg2[k] = pImg2[q];
  bAt2[k] = pAt2[q];
  bStart[k] = bOff[q];
}

struct Inv9 {
  double v[9];
};

// integer cell translation of each block's image: rint((r_img - r_centre) L^-1)
__global__ void nl_block_translations_kernel(int nBlk, const int* img2, const int* at2,
                                             const double* coordsAll, const double* coords,
                                             Inv9 inv, double* T) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nBlk) return;
  const int img = img2[k], jc = at2[k];
  double d[3];
  for (int q = 0; q < 3; ++q) d[q] = coordsAll[3 * (long)img + q] - coords[3 * (long)jc + q];
  for (int q = 0; q < 3; ++q)
    T[3 * (long)k + q] = rint(d[0] * inv.v[q] + d[1] * inv.v[3 + q] + d[2] * inv.v[6 + q]);
}

int block_translations(hipStream_t s, int nBlk, const int* img2, const int* at2,
                       const double* coordsAll, const double* coords, const double* inv,
                       double* T) {
  if (nBlk <= 0) return 0;
  Inv9 iv;
  for (int q = 0; q < 9; ++q) iv.v[q] = inv[q];
  hipLaunchKernelGGL(nl_block_translations_kernel, dim3(cdiv(nBlk, 256)), dim3(256), 0, s, nBlk,
                     img2, at2, coordsAll, coords, iv, T);
  DH_LAUNCH_CHECK();
  return 0;
}

static int read_back(hipStream_t s, void* dst, const void* src, size_t bytes) {
  DH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  DH_HIP(hipStreamSynchronize(s));
  return 0;
}

int device_neighbours(hipStream_t s, int nat, const double* coords, const double* lat,
                      double cutoff, double skcut, const int* norb, DevNeighbourList& nl) {
  DH_CHECK(nat >= 1 && coords && norb && cutoff > 0.0 && skcut <= cutoff,
           "neighbour list: bad arguments");
  std::vector<double> tr;
  cell_translations(lat, cutoff, tr);
  const long ntr = (long)tr.size() / 3;
  const long P = ntr * nat;
  DH_CHECK(P < INT_MAX / 2, "neighbour list: too many candidate images");
  NlGrid g;
  g.cutoff = cutoff;
  for (int c = 0; c < 3; ++c) {
    g.lo[c] = 1e300;
    g.hi[c] = -1e300;
  }
  for (int a = 0; a < nat; ++a)
    for (int c = 0; c < 3; ++c) {
      g.lo[c] = std::min(g.lo[c], coords[3 * a + c]);
      g.hi[c] = std::max(g.hi[c], coords[3 * a + c]);
    }
  long nbins = 1;
  for (int c = 0; c < 3; ++c) {
    g.nbin[c] = std::max(1, (int)std::floor((g.hi[c] - g.lo[c] + 2 * cutoff) / cutoff));
    nbins *= g.nbin[c];
  }
  DH_CHECK(nbins < INT_MAX / 2, "neighbour list: too many grid cells");
  nl.nat = nat;
  DH_TRY(nl.coordsD.upload(coords, 3 * (size_t)nat, s));
  DH_TRY(nl.trD.upload(tr.data(), tr.size(), s));
  DH_TRY(nl.norbD.upload(norb, nat, s));
  DH_TRY(nl.px.alloc(3 * (size_t)P));
  DH_TRY(nl.pbin.alloc(P));
  DH_TRY(nl.binItems.alloc(P));
  DH_TRY(nl.binCount.alloc(nbins));
  DH_TRY(nl.binStart.alloc(nbins));
  DH_TRY(nl.binFill.alloc(nbins));
  DH_TRY(nl.err.alloc(4));
  DH_TRY(nl.scal.alloc(4));
  DH_TRY(nl.iscal.alloc(4));
  DH_TRY(nl.cnt.alloc(nat));
  DH_TRY(nl.fStart.alloc(nat));
  DH_TRY(nl.used.alloc(P));
  DH_TRY(nl.imgRank.alloc(P));
  DH_TRY(fill<int>(s, nl.binCount.p, nbins, 0));
  DH_TRY(fill<int>(s, nl.err.p, 4, 0));
  DH_TRY(fill<int>(s, nl.used.p, P, 0));
  // candidates into grid cells
  hipLaunchKernelGGL(nl_points_kernel, dim3(cdiv(P, 256)), dim3(256), 0, s, nat, P, nl.coordsD.p,
                     nl.trD.p, g, nl.px.p, nl.pbin.p, nl.binCount.p);
  DH_LAUNCH_CHECK();
  DH_TRY(scan_exclusive<int>(s, nl.binCount.p, nl.binStart.p, nbins, nl.iscal.p));
  DH_HIP(hipMemcpyAsync(nl.binFill.p, nl.binStart.p, nbins * sizeof(int),
                        hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(nl_scatter_kernel, dim3(cdiv(P, 256)), dim3(256), 0, s, P, nl.pbin.p,
                     nl.binFill.p, nl.binItems.p);
  DH_LAUNCH_CHECK();
  // neighbours per central atom, their offsets
  const double cut2 = cutoff * cutoff;
  hipLaunchKernelGGL(nl_search_kernel<false>, dim3(cdiv(nat, 64)), dim3(64), 0, s, nat, g,
                     nl.coordsD.p, nl.px.p, nl.binStart.p, nl.binCount.p, nl.binItems.p, cut2,
                     nl.cnt.p, nullptr, nullptr, nullptr, nl.err.p);
  DH_LAUNCH_CHECK();
  DH_TRY(nl.size.alloc(nat));
  std::vector<long> cl(nat);   // counts widened to long for the offsets (read by the upload)
  {
    std::vector<int> c(nat);
    DH_TRY(read_back(s, c.data(), nl.cnt.p, nat * sizeof(int)));
    int err[4];
    DH_TRY(read_back(s, err, nl.err.p, sizeof(err)));
    if (err[0]) {
      set_error("Atoms " + std::to_string(err[1]) + " and " + std::to_string(err[2]) +
                " too close together");
      return kErrArg;
    }
    nl.maxNeighbour = 0;
    for (int a = 0; a < nat; ++a) {
      cl[a] = c[a];
      nl.maxNeighbour = std::max(nl.maxNeighbour, c[a]);
    }
    if (nl.maxNeighbour > kNlSortMax) {
      set_error("neighbour list: an atom has " + std::to_string(nl.maxNeighbour) +
                " neighbours, more than " + std::to_string(kNlSortMax));
      return kErrArg;
    }
    DH_TRY(nl.size.upload(cl.data(), nat, s));
  }
  DH_TRY(scan_exclusive<long>(s, nl.size.p, nl.fStart.p, nat, nl.scal.p));
  long F = 0;
  DH_TRY(read_back(s, &F, nl.scal.p, sizeof(long)));
  DH_TRY(nl.fD2.alloc(std::max<long>(F, 1)));
  DH_TRY(nl.fPi.alloc(std::max<long>(F, 1)));
  hipLaunchKernelGGL(nl_search_kernel<true>, dim3(cdiv(nat, 64)), dim3(64), 0, s, nat, g,
                     nl.coordsD.p, nl.px.p, nl.binStart.p, nl.binCount.p, nl.binItems.p, cut2,
                     nl.cnt.p, nl.fStart.p, nl.fD2.p, nl.fPi.p, nl.err.p);
  DH_LAUNCH_CHECK();
  hipLaunchKernelGGL(nl_sort_kernel, dim3(nat), dim3(256), 0, s, nat, nl.cnt.p, nl.fStart.p,
                     nl.fD2.p, nl.fPi.p, nl.used.p, nl.err.p);
  DH_LAUNCH_CHECK();
  // image numbers
  DH_TRY(scan_exclusive<int>(s, nl.used.p, nl.imgRank.p, P, nl.iscal.p));
  int nImg = 0;
  DH_TRY(read_back(s, &nImg, nl.iscal.p, sizeof(int)));
  nl.nAll = nat + nImg;
  DH_TRY(nl.coordsAll.alloc(3 * (size_t)nl.nAll));
  DH_TRY(nl.img2cent.alloc(nl.nAll));
  hipLaunchKernelGGL(nl_images_kernel, dim3(cdiv(P, 256)), dim3(256), 0, s, nat, P, nl.coordsD.p,
                     nl.px.p, nl.used.p, nl.imgRank.p, nl.coordsAll.p, nl.img2cent.p);
  DH_LAUNCH_CHECK();
  // pairs, self first per atom
  nl.nPairs = (int)(F + nat);
  DH_TRY(nl.pAt1.alloc(nl.nPairs));
  DH_TRY(nl.pImg2.alloc(nl.nPairs));
  DH_TRY(nl.pAt2.alloc(nl.nPairs));
  DH_TRY(nl.pD2.alloc(nl.nPairs));
  DH_TRY(nl.pCount.alloc(nat));
  DH_TRY(nl.pStart.alloc(nat + 1));
  hipLaunchKernelGGL(nl_pairs_kernel, dim3(cdiv(nat, 256)), dim3(256), 0, s, nat, nl.cnt.p,
                     nl.fStart.p, nl.fD2.p, nl.fPi.p, nl.imgRank.p, nl.pAt1.p, nl.pImg2.p,
                     nl.pAt2.p, nl.pD2.p, nl.pCount.p, nl.pStart.p);
  DH_LAUNCH_CHECK();
  // blocks within the SK cutoff and their offsets
  DH_TRY(nl.flag.alloc(nl.nPairs));
  DH_TRY(nl.bIdx.alloc(nl.nPairs));
  DH_TRY(nl.size.alloc(nl.nPairs));
  DH_TRY(nl.bOff.alloc(nl.nPairs));
  hipLaunchKernelGGL(nl_block_flags_kernel, dim3(cdiv(nl.nPairs, 256)), dim3(256), 0, s,
                     (long)nl.nPairs, nl.pD2.p, nl.pAt1.p, nl.pAt2.p, nl.norbD.p, skcut * skcut,
                     nl.flag.p, nl.size.p);
  DH_LAUNCH_CHECK();
  DH_TRY(scan_exclusive<int>(s, nl.flag.p, nl.bIdx.p, nl.nPairs, nl.iscal.p));
  DH_TRY(scan_exclusive<long>(s, nl.size.p, nl.bOff.p, nl.nPairs, nl.scal.p));
  DH_TRY(read_back(s, &nl.nBlk, nl.iscal.p, sizeof(int)));
  DH_TRY(read_back(s, &nl.nnz, nl.scal.p, sizeof(long)));
  DH_TRY(nl.bAt1.alloc(std::max(nl.nBlk, 1)));
  DH_TRY(nl.bImg2.alloc(std::max(nl.nBlk, 1)));
  DH_TRY(nl.bAt2.alloc(std::max(nl.nBlk, 1)));
  DH_TRY(nl.bStart.alloc(std::max(nl.nBlk, 1)));
  hipLaunchKernelGGL(nl_blocks_kernel, dim3(cdiv(nl.nPairs, 256)), dim3(256), 0, s,
                     (long)nl.nPairs, nl.flag.p, nl.bIdx.p, nl.bOff.p, nl.pAt1.p, nl.pImg2.p,
                     nl.pAt2.p, nl.bAt1.p, nl.bImg2.p, nl.bAt2.p, nl.bStart.p);
  DH_LAUNCH_CHECK();
  int err[4];
  DH_TRY(read_back(s, err, nl.err.p, sizeof(err)));
  DH_CHECK(err[3] == 0, "neighbour list: neighbour count above the sort capacity");
  return 0;
}

}  // namespace dftbhip
