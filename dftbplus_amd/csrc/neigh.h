// Neighbour list and sparse block descriptor built on the device (neigh.hip).
#pragma once

#include <string>
#include <vector>

#include "common.h"

namespace dftbhip {

// Lattice translations that can hold atoms within cutoff of the central cell, the zero
// translation first (3 doubles each; lat: the three lattice vectors, or null: cluster).
void cell_translations(const double* lat, double cutoff, std::vector<double>& tr);

// The half neighbour list and the SK block list, resident on the device. Pairs are ordered as
// the reference's iNeighbour: per central atom a, first a itself, then the images of atoms
// b >= a within the cutoff by increasing squared distance (ties by cell translation, then atom).
// Images are numbered after the central atoms in (translation, atom) order.
struct DevNeighbourList {
  int nat = 0, nAll = 0, nPairs = 0, nBlk = 0, maxNeighbour = 0;
  long nnz = 0;
  DevBuf<double> coordsAll;          // 3 nAll
  DevBuf<int> img2cent;              // nAll
  DevBuf<int> pAt1, pImg2, pAt2;     // all pairs (nPairs = nat + neighbours)
  DevBuf<double> pD2;                // squared distance of each pair
  DevBuf<int> pCount;                // neighbours of each atom, self excluded
  DevBuf<long> pStart;               // first pair of each atom (nat + 1)
  DevBuf<int> bAt1, bImg2, bAt2;     // pairs within the SK cutoff
  DevBuf<long> bStart;               // their block offsets in the sparse arrays
  // work arrays
  DevBuf<double> px, fD2;
  DevBuf<int> pbin, binCount, binStart, binFill, binItems, cnt, fPi, used, imgRank, flag, err;
  DevBuf<long> fStart, size, bOff, scal;
  DevBuf<double> coordsD, trD;
  DevBuf<int> norbD, bIdx, iscal;
};

// Builds the list for nat atoms at coords (host, 3 nat) with lattice lat (host, 9 doubles, or
// null) and cutoff; blocks within skcut with norb orbitals per atom (host). Errors: atoms closer
// than 0.01 bohr (kErrArg), more than 4096 neighbours of one atom (kErrArg).
int device_neighbours(hipStream_t s, int nat, const double* coords, const double* lat,
                      double cutoff, double skcut, const int* norb, DevNeighbourList& nl);

// T (3 per block) = rint((coordsAll[img2] - coords[at2]) inv), inv = L^-1 (host, row-major
// as the lattice rows), for the k-point phases of the blocks
int block_translations(hipStream_t s, int nBlk, const int* img2, const int* at2,
                       const double* coordsAll, const double* coords, const double* inv,
                       double* T);

}  // namespace dftbhip
