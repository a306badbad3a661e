// Cell / Verlet list state shared by the list builder and the pair-force, learned-correction
// and Ewald real-space kernels.
#pragma once
#include "common.h"

namespace accu {

struct CellGeo {
  double lo[3];
  double span[3];
  double box[3];
  int nc[3];
  int periodic;
  int dedup;  // some axis has < 3 cells: stencil cells repeat and must be de-duplicated
};

struct NList {
  int n = 0;
  long long ncells = 0;
  CellGeo geo{};
  double rc2 = 0.0;
  // per-atom arrays (capacity cap_n)
  long long cap_n = 0;
  int* cell = nullptr;
  int* sorted = nullptr;
  int* tmp = nullptr;   // atom ids 0..n-1 (values of the sort by cell)
  int* skey = nullptr;  // sorted cell ids
  int* nn = nullptr;
  int* halfcnt = nullptr;
  // per-cell arrays (capacity cap_cells)
  long long cap_cells = 0;
  int* count = nullptr;
  int* start = nullptr;
  // ELL neighbour list
  int* nbr = nullptr;
  long long nbr_cap = 0;
  int stride = 0;
  int max_nbrs = 0;
  int last_max = 0;
  int* dev_scalars = nullptr;  // [0] = max neighbour count of the last build
  long long* half_off = nullptr;
  long long half_cap = 0;
  void* scratch = nullptr;  // temporary storage of the library sort and scans
  size_t scratch_cap = 0;

  int ensure(int natoms, long long cells);
  int ensure_scratch(size_t bytes);
  int ensure_ell(int natoms, int nbrs);
  void release();
};

int make_geo(CellGeo& g, const double* lo, const double* span, const int* ncell, const double* box, int periodic);
int launch_bin_sort(NList& h, const double* pos, int stride_pos, int n, const CellGeo& g, hipStream_t st);
// exact = 1: minimum image with the true division (bit-exact pair sets, the drop-in
// path); exact = 0: reciprocal multiply, identical for every |d| < L/2
int launch_build_ell(NList& h, const double* pos, int stride_pos, int n, const CellGeo& g, double rc2,
                     const int* sorted, hipStream_t st, int exact = 1);

}  // namespace accu
