// Cell binning, a stable sort by cell (rocPRIM radix sort), and the Verlet (ELL) neighbour list.
//
// Follows accusim/forces.py:279-321 (cell_list_pairs) and 78-98 (VerletList):
//   cell = clamp(trunc(((pos - lo) / span) * ncell), 0, ncell - 1), flat C-order index,
//   atoms stably sorted by flat cell (argsort kind="stable"),
//   27-cell stencil (wrapped if periodic, clipped if not, de-duplicated),
//   pair kept when the minimum-image |d|^2 < rc^2, d = pos[j] - pos[i].
// Every float64 expression that decides membership is evaluated in the reference's
// operation order without contraction, so pair sets are bit-exact.
//
// Layout: the list is "ELL" by sorted slot t: nbr[k * stride + t] = atom id of the
// k-th neighbour of atom sorted[t], k < nn[t]. Both directions are stored (full list),
// so force kernels gather and never scatter. stride = n rounded up to 64.
#include "common.h"
#include "nlist.h"
#include "../../include/accusim_hip.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstring>

namespace accu {

// ------------------------------------------------------------------ kernels

__global__ void k_zero_int(int* __restrict__ a, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// Flat cell of every atom (forces.py:292-296), its count, and the identity values of the
// sort that follows.
__global__ void k_cell_assign(const double* __restrict__ pos, int stride_pos, int n, CellGeo g,
                              int* __restrict__ cell, int* __restrict__ count, int* __restrict__ idx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double t;
    {
#pragma clang fp contract(off)
      t = ((pos[(size_t)stride_pos * i + d] - g.lo[d]) / g.span[d]) * (double)g.nc[d];
    }
    int m = g.nc[d] - 1;
    int ci;
    if (!(t >= 0.0)) ci = 0;  // negative (trunc then max 0) or NaN
    else if (t >= (double)m) ci = m;
    else ci = (int)t;
    c[d] = ci;
  }
  int flat = (c[0] * g.nc[1] + c[1]) * g.nc[2] + c[2];
  cell[i] = flat;
  idx[i] = i;
  atomicAdd(&count[flat], 1);
}

__device__ __forceinline__ void unflat(int f, const int* nc, int* c) {
  c[2] = f % nc[2];
  f /= nc[2];
  c[1] = f % nc[1];
  c[0] = f / nc[1];
}

// Build the full ELL list. Thread = sorted slot t, atom i = sorted[t] (or t when
// sorted == nullptr, i.e. the caller has already put atoms in cell order).
__global__ void k_build_ell(const double* __restrict__ pos, int stride_pos, int n, CellGeo g,
                            double rc2, const int* __restrict__ cell, const int* __restrict__ start,
                            const int* __restrict__ sorted, int* __restrict__ nn, int* __restrict__ nbr,
                            int stride, int max_nbrs, int* __restrict__ maxcount, int exact) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  double inv[3] = {1.0 / g.box[0], 1.0 / g.box[1], 1.0 / g.box[2]};
  if (t >= n) return;
  const int i = sorted ? sorted[t] : t;
  const double xi = pos[(size_t)stride_pos * i], yi = pos[(size_t)stride_pos * i + 1],
               zi = pos[(size_t)stride_pos * i + 2];
  int ci[3];
  unflat(cell[i], g.nc, ci);
  int seen[27];
  int nseen = 0;
  int cnt = 0;
  for (int o = 0; o < 27; ++o) {
    int c[3] = {ci[0] + o / 9 - 1, ci[1] + (o / 3) % 3 - 1, ci[2] + o % 3 - 1};
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (g.periodic) {
        c[d] = ((c[d] % g.nc[d]) + g.nc[d]) % g.nc[d];
      } else if (c[d] < 0 || c[d] >= g.nc[d]) {
        ok = false;
      }
    }
    if (!ok) continue;
    int f = (c[0] * g.nc[1] + c[1]) * g.nc[2] + c[2];
    if (g.dedup) {
      bool dup = false;
      for (int q = 0; q < nseen; ++q) dup |= (seen[q] == f);
      if (dup) continue;
      seen[nseen++] = f;
    }
    const int a = start[f], b = start[f + 1];
    for (int s = a; s < b; ++s) {
      const int j = sorted ? sorted[s] : s;
      if (j == i) continue;
      double dx, dy, dz;
      {
#pragma clang fp contract(off)
        dx = pos[(size_t)stride_pos * j] - xi;
        dy = pos[(size_t)stride_pos * j + 1] - yi;
        dz = pos[(size_t)stride_pos * j + 2] - zi;
      }
      double r2;
      if (g.periodic) {
        // reciprocal-multiply minimum image first: its r2 is within a few ulps of the
        // exact one (it picks another image only at a rounding tie, |d| = L/2, where both
        // images are L/2 away), so only a candidate within 1e-10 of rc2 needs the exact
        // expression with the true division to decide membership bit for bit
        r2 = r2_np(minimg_fast(dx, g.box[0], inv[0]), minimg_fast(dy, g.box[1], inv[1]),
                   minimg_fast(dz, g.box[2], inv[2]));
        if (exact && fabs(r2 - rc2) <= 1e-10 * rc2)
          r2 = r2_np(minimg_exact(dx, g.box[0]), minimg_exact(dy, g.box[1]), minimg_exact(dz, g.box[2]));
      } else {
        r2 = r2_np(dx, dy, dz);
      }
      if (r2 < rc2) {
        if (cnt < max_nbrs) nbr[(size_t)cnt * stride + t] = j;
        ++cnt;
      }
    }
  }
  // the stored count never exceeds the capacity (readers index nbr with it); the true
  // maximum goes to maxcount so the host can detect an overflow and regrow
  nn[t] = cnt < max_nbrs ? cnt : max_nbrs;
  atomicMax(maxcount, cnt);
}

// The same list built one home cell per block: the candidates of the cell's 27-cell stencil
// (in the order above: stencil cell, then slot) are staged into the LDS in chunks, and
// every atom of the cell (one per thread) tests them there. All lanes of a wave read the
// same candidate, so the LDS reads broadcast, and the global gathers of a candidate happen
// once per cell instead of once per atom. Entries, their order and the counts are those of
// k_build_ell.
constexpr int kEB = 256;   // threads per block
constexpr int kECh = 1024;  // candidates per LDS chunk (28 KiB)
__global__ void __launch_bounds__(kEB) k_build_ell_cell(const double* __restrict__ pos, int stride_pos, CellGeo g,
                                                        double rc2, const int* __restrict__ start,
                                                        const int* __restrict__ sorted, int* __restrict__ nn,
                                                        int* __restrict__ nbr, int stride, int max_nbrs,
                                                        int* __restrict__ maxcount, int exact) {
  __shared__ double lx[kECh], ly[kECh], lz[kECh];
  __shared__ int lid[kECh];
  __shared__ int cf[27], cst[28];
  __shared__ int ncell;
  const int f = blockIdx.x;
  const int h0 = start[f], h1 = start[f + 1];
  if (h0 == h1) return;
  if (threadIdx.x == 0) {  // the stencil, de-duplicated and clipped as in k_build_ell
    int ci[3];
    unflat(f, g.nc, ci);
    int m = 0, tot = 0;
    for (int o = 0; o < 27; ++o) {
      int c[3] = {ci[0] + o / 9 - 1, ci[1] + (o / 3) % 3 - 1, ci[2] + o % 3 - 1};
      bool ok = true;
      for (int d = 0; d < 3; ++d) {
        if (g.periodic) c[d] = ((c[d] % g.nc[d]) + g.nc[d]) % g.nc[d];
        else if (c[d] < 0 || c[d] >= g.nc[d]) ok = false;
      }
      if (!ok) continue;
      const int q = (c[0] * g.nc[1] + c[1]) * g.nc[2] + c[2];
      bool dup = false;
      if (g.dedup)
        for (int k = 0; k < m; ++k) dup |= (cf[k] == q);
      if (dup) continue;
      cf[m] = q;
      cst[m] = tot;
      tot += start[q + 1] - start[q];
      ++m;
    }
    cst[m] = tot;
    ncell = m;
  }
  __syncthreads();
  const int m = ncell, total = cst[m];
  const double inv[3] = {1.0 / g.box[0], 1.0 / g.box[1], 1.0 / g.box[2]};
  for (int a0 = h0; a0 < h1; a0 += blockDim.x) {
    const int t = a0 + (int)threadIdx.x;
    const bool on = t < h1;
    const int i = on ? sorted[t] : -1;
    const double xi = on ? pos[(size_t)stride_pos * i] : 0.0, yi = on ? pos[(size_t)stride_pos * i + 1] : 0.0,
                 zi = on ? pos[(size_t)stride_pos * i + 2] : 0.0;
    int cnt = 0;
    for (int base = 0; base < total; base += kECh) {
      const int len = min(kECh, total - base);
      __syncthreads();  // the previous chunk is no longer read
      for (int k = threadIdx.x; k < len; k += blockDim.x) {
        const int q = base + k;
        int lo = 0, hi = m - 1;  // the stencil cell r with cst[r] <= q < cst[r + 1]
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (cst[mid] <= q) lo = mid;
          else hi = mid - 1;
        }
        const int j = sorted[start[cf[lo]] + (q - cst[lo])];
        lx[k] = pos[(size_t)stride_pos * j];
        ly[k] = pos[(size_t)stride_pos * j + 1];
        lz[k] = pos[(size_t)stride_pos * j + 2];
        lid[k] = j;
      }
      __syncthreads();
      if (!on) continue;
      auto test = [&](int j, double xj, double yj, double zj) {
        if (j == i) return;
        double dx, dy, dz;
        {
#pragma clang fp contract(off)
          dx = xj - xi;
          dy = yj - yi;
          dz = zj - zi;
        }
        double r2;
        if (g.periodic) {  // as in k_build_ell: the fast image, the exact one near rc2
          r2 = r2_np(minimg_fast(dx, g.box[0], inv[0]), minimg_fast(dy, g.box[1], inv[1]),
                     minimg_fast(dz, g.box[2], inv[2]));
          if (exact && fabs(r2 - rc2) <= 1e-10 * rc2)
            r2 = r2_np(minimg_exact(dx, g.box[0]), minimg_exact(dy, g.box[1]), minimg_exact(dz, g.box[2]));
        } else {
          r2 = r2_np(dx, dy, dz);
        }
        if (r2 < rc2) {
          if (cnt < max_nbrs) nbr[(size_t)cnt * stride + t] = j;
          ++cnt;
        }
      };
      int k = 0;
      for (; k + 4 <= len; k += 4) {  // four candidates' LDS reads in flight together
        double X[4], Y[4], Z[4];
        int J[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          X[u] = lx[k + u];
          Y[u] = ly[k + u];
          Z[u] = lz[k + u];
          J[u] = lid[k + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) test(J[u], X[u], Y[u], Z[u]);
      }
      for (; k < len; ++k) test(lid[k], lx[k], ly[k], lz[k]);
    }
    if (on) {
      nn[t] = cnt < max_nbrs ? cnt : max_nbrs;
      atomicMax(maxcount, cnt);
    }
  }
}

__global__ void k_count_half(int n, const int* __restrict__ sorted, const int* __restrict__ nn,
                             const int* __restrict__ nbr, int stride, int* __restrict__ cnt) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) cnt[n] = 0;  // the scan's extra slot gives the total
  if (t >= n) return;
  const int i = sorted ? sorted[t] : t;
  int c = 0;
  for (int k = 0; k < nn[t]; ++k) c += (nbr[(size_t)k * stride + t] > i);
  cnt[t] = c;
}

__global__ void k_emit_half(int n, const int* __restrict__ sorted, const int* __restrict__ nn,
                            const int* __restrict__ nbr, int stride, const long long* __restrict__ off,
                            long long* __restrict__ iout, long long* __restrict__ jout) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = sorted ? sorted[t] : t;
  long long o = off[t];
  for (int k = 0; k < nn[t]; ++k) {
    int j = nbr[(size_t)k * stride + t];
    if (j > i) {
      iout[o] = i;
      jout[o] = j;
      ++o;
    }
  }
}

// max_i |minimum-image(pos_i - ref_i)|^2 (forces.py:90-93), as the bit pattern of a
// non-negative double so atomicMax orders it.
__global__ void k_maxdisp2(const double* __restrict__ pos, int sp, const double* __restrict__ ref, int sr,
                           int n, Box box, unsigned long long* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    double d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      {
#pragma clang fp contract(off)
        d[k] = pos[(size_t)sp * i + k] - ref[(size_t)sr * i + k];
      }
      if (box.periodic) d[k] = minimg_exact(d[k], box.L[k]);
    }
    v = r2_np(d[0], d[1], d[2]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(v));
}

// ------------------------------------------------------------------ host side

static int realloc_ints(int** p, long long count) {
  if (*p) HIP_TRY(hipFree(*p));
  *p = nullptr;
  HIP_TRY(hipMalloc(p, count * sizeof(int)));
  return ACCU_OK;
}

int NList::ensure(int natoms, long long cells) {
  int rc;
  if (natoms + 1 > cap_n) {
    long long c = natoms + 1 + natoms / 4 + 64;
    int** arrs[] = {&cell, &sorted, &tmp, &skey, &nn, &halfcnt};
    for (int** a : arrs)
      if ((rc = realloc_ints(a, c))) return rc;
    cap_n = c;
  }
  if (cells + 1 > cap_cells) {
    long long c = cells + 1 + cells / 4 + 64;
    int** arrs[] = {&count, &start};
    for (int** a : arrs)
      if ((rc = realloc_ints(a, c))) return rc;
    cap_cells = c;
  }
  if (!dev_scalars) {
    HIP_TRY(hipMalloc(&dev_scalars, 64));
    HIP_TRY(hipMemset(dev_scalars, 0, 64));
    HIP_TRY(hipDeviceSynchronize());  // the null-stream fill completes before any caller stream runs
  }
  return ACCU_OK;
}

int NList::ensure_ell(int natoms, int nbrs) {
  int st = (natoms + 63) / 64 * 64;
  long long need = (long long)st * nbrs;
  if (need > nbr_cap) {
    if (nbr) HIP_TRY(hipFree(nbr));
    HIP_TRY(hipMalloc(&nbr, need * sizeof(int)));
    nbr_cap = need;
  }
  stride = st;
  max_nbrs = nbrs;
  return ACCU_OK;
}

void NList::release() {
  int* ps[] = {cell, sorted, tmp, skey, nn, halfcnt, count, start, nbr, dev_scalars};
  for (int* p : ps)
    if (p) (void)hipFree(p);
  if (half_off) (void)hipFree(half_off);
  if (scratch) (void)hipFree(scratch);
}

int NList::ensure_scratch(size_t bytes) {
  if (bytes <= scratch_cap && scratch) return ACCU_OK;
  if (scratch) HIP_TRY(hipFree(scratch));
  scratch = nullptr;
  HIP_TRY(hipMalloc(&scratch, bytes + 4096));
  scratch_cap = bytes + 4096;
  return ACCU_OK;
}

// Binning and the stable sort by cell (forces.py:292-298: np.argsort(cell, kind="stable")):
// counts per cell (atomics: the count is order-free), their exclusive scan to cell starts
// (start[ncells] = n), and a stable radix sort of (cell, atom) pairs, so atoms of a cell
// keep their index order.
int launch_bin_sort(NList& h, const double* pos, int stride_pos, int n, const CellGeo& g, hipStream_t st) {
  const int T = 256;
  const long long nc = h.ncells;
  hipLaunchKernelGGL(k_zero_int, dim3((unsigned)((nc + 1 + T - 1) / T)), dim3(T), 0, st, h.count, nc + 1);
  hipLaunchKernelGGL(k_cell_assign, dim3((n + T - 1) / T), dim3(T), 0, st, pos, stride_pos, n, g, h.cell, h.count,
                     h.tmp);
  LAUNCH_CHECK();
  size_t b1 = 0, b2 = 0;
  HIP_TRY(rocprim::exclusive_scan(nullptr, b1, h.count, h.start, 0, (size_t)nc + 1, rocprim::plus<int>(), st));
  unsigned bits = 1;
  while (bits < 31 && (1ll << bits) < nc) ++bits;
  // cell ids are non-negative: sorted as unsigned keys over their low `bits` bits
  const unsigned* keys_in = reinterpret_cast<const unsigned*>(h.cell);
  unsigned* keys_out = reinterpret_cast<unsigned*>(h.skey);
  HIP_TRY(rocprim::radix_sort_pairs(nullptr, b2, keys_in, keys_out, h.tmp, h.sorted, (unsigned)n, 0u, bits, st));
  int rc = h.ensure_scratch(b1 > b2 ? b1 : b2);
  if (rc) return rc;
  b1 = b2 = h.scratch_cap;
  HIP_TRY(rocprim::exclusive_scan(h.scratch, b1, h.count, h.start, 0, (size_t)nc + 1, rocprim::plus<int>(), st));
  HIP_TRY(rocprim::radix_sort_pairs(h.scratch, b2, keys_in, keys_out, h.tmp, h.sorted, (unsigned)n, 0u, bits, st));
  return ACCU_OK;
}

int launch_build_ell(NList& h, const double* pos, int stride_pos, int n, const CellGeo& g, double rc2,
                     const int* sorted, hipStream_t st, int exact) {
  if (sorted && h.ncells < (1LL << 31)) {  // one block per home cell over the cell-sorted atoms
    hipLaunchKernelGGL(k_build_ell_cell, dim3((unsigned)h.ncells), dim3(kEB), 0, st, pos, stride_pos, g, rc2, h.start,
                       sorted, h.nn, h.nbr, h.stride, h.max_nbrs, h.dev_scalars, exact);
  } else {
    const int T = 256;
    hipLaunchKernelGGL(k_build_ell, dim3((n + T - 1) / T), dim3(T), 0, st, pos, stride_pos, n, g, rc2, h.cell,
                       h.start, sorted, h.nn, h.nbr, h.stride, h.max_nbrs, h.dev_scalars, exact);
  }
  LAUNCH_CHECK();
  return ACCU_OK;
}

int make_geo(CellGeo& g, const double* lo, const double* span, const int* ncell, const double* box, int periodic) {
  long long tot = 1;
  g.dedup = 0;
  for (int d = 0; d < 3; ++d) {
    ARG_CHECK(ncell[d] >= 1, "ncell[%d] = %d must be >= 1", d, ncell[d]);
    ARG_CHECK(span[d] > 0, "span[%d] must be positive", d);
    g.lo[d] = lo[d];
    g.span[d] = span[d];
    g.nc[d] = ncell[d];
    g.box[d] = box ? box[d] : 0.0;
    tot *= ncell[d];
    if (ncell[d] < 3) g.dedup = 1;
  }
  ARG_CHECK(tot < (1LL << 30), "too many cells (%lld)", tot);
  ARG_CHECK(!periodic || box, "periodic list needs a box");
  g.periodic = periodic;
  return ACCU_OK;
}

}  // namespace accu

using namespace accu;

extern "C" {

int accu_nlist_create(accu_nlist** out) {
  ARG_CHECK(out, "null handle pointer");
  *out = reinterpret_cast<accu_nlist*>(new NList());
  return ACCU_OK;
}

int accu_nlist_destroy(accu_nlist* h) {
  if (h) {
    NList* p = reinterpret_cast<NList*>(h);
    p->release();
    delete p;
  }
  return ACCU_OK;
}

int accu_nlist_build(accu_nlist* hh, const double* pos, int n, const double* lo, const double* span,
                     const int* ncell, const double* box, int periodic, double rc, void* stream) {
  ARG_CHECK(hh && pos && lo && span && ncell, "null argument");
  ARG_CHECK(n >= 0, "n must be >= 0");
  NList& h = *reinterpret_cast<NList*>(hh);
  hipStream_t st = (hipStream_t)stream;
  CellGeo g;
  int rc_ = make_geo(g, lo, span, ncell, box, periodic);
  if (rc_) return rc_;
  h.geo = g;
  h.n = n;
  h.ncells = (long long)ncell[0] * ncell[1] * ncell[2];
  {
#pragma clang fp contract(off)
    h.rc2 = rc * rc;  // forces.py:320 compares with rc ** 2
  }
  if ((rc_ = h.ensure(n, h.ncells))) return rc_;
  if (n == 0) {
    h.last_max = 0;
    return ACCU_OK;
  }
  if ((rc_ = launch_bin_sort(h, pos, 3, n, g, st))) return rc_;
  int guess = h.max_nbrs > 0 ? h.max_nbrs : 64;
  for (int attempt = 0; attempt < 3; ++attempt) {
    if ((rc_ = h.ensure_ell(n, guess))) return rc_;
    HIP_TRY(hipMemsetAsync(h.dev_scalars, 0, sizeof(int), st));
    if ((rc_ = launch_build_ell(h, pos, 3, n, g, h.rc2, h.sorted, st))) return rc_;
    int mx = 0;
    HIP_TRY(hipMemcpyAsync(&mx, h.dev_scalars, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    h.last_max = mx;
    if (mx <= guess) return ACCU_OK;
    guess = mx + mx / 8 + 8;
  }
  set_error("neighbour list did not fit after regrowing");
  return ACCU_EOVERFLOW;
}

int accu_nlist_info(accu_nlist* hh, long long* out) {
  ARG_CHECK(hh && out, "null argument");
  NList& h = *reinterpret_cast<NList*>(hh);
  out[0] = h.n;
  out[1] = h.ncells;
  out[2] = h.stride;
  out[3] = h.max_nbrs;
  out[4] = h.last_max;
  out[5] = h.geo.nc[0];
  out[6] = h.geo.nc[1];
  out[7] = h.geo.nc[2];
  return ACCU_OK;
}

int accu_nlist_cells(accu_nlist* hh, int* cell_out, int* sorted_out, int* nn_out, void* stream) {
  ARG_CHECK(hh, "null handle");
  NList& h = *reinterpret_cast<NList*>(hh);
  hipStream_t st = (hipStream_t)stream;
  if (h.n == 0) return ACCU_OK;
  if (cell_out) HIP_TRY(hipMemcpyAsync(cell_out, h.cell, h.n * sizeof(int), hipMemcpyDeviceToDevice, st));
  if (sorted_out) HIP_TRY(hipMemcpyAsync(sorted_out, h.sorted, h.n * sizeof(int), hipMemcpyDeviceToDevice, st));
  if (nn_out) HIP_TRY(hipMemcpyAsync(nn_out, h.nn, h.n * sizeof(int), hipMemcpyDeviceToDevice, st));
  return ACCU_OK;
}

int accu_nlist_count_half(accu_nlist* hh, long long* npairs, void* stream) {
  ARG_CHECK(hh && npairs, "null argument");
  NList& h = *reinterpret_cast<NList*>(hh);
  hipStream_t st = (hipStream_t)stream;
  *npairs = 0;
  if (h.n == 0) return ACCU_OK;
  if (!h.half_off || h.half_cap < h.n + 1) {
    if (h.half_off) HIP_TRY(hipFree(h.half_off));
    HIP_TRY(hipMalloc(&h.half_off, (h.n + 1) * sizeof(long long)));
    h.half_cap = h.n + 1;
  }
  const int T = 256;
  hipLaunchKernelGGL(k_count_half, dim3((h.n + T - 1) / T), dim3(T), 0, st, h.n, h.sorted, h.nn, h.nbr, h.stride,
                     h.halfcnt);
  LAUNCH_CHECK();
  size_t bytes = 0;
  HIP_TRY(rocprim::exclusive_scan(nullptr, bytes, h.halfcnt, h.half_off, 0ll, (size_t)h.n + 1,
                                  rocprim::plus<long long>(), st));
  int rc = h.ensure_scratch(bytes);
  if (rc) return rc;
  bytes = h.scratch_cap;
  HIP_TRY(rocprim::exclusive_scan(h.scratch, bytes, h.halfcnt, h.half_off, 0ll, (size_t)h.n + 1,
                                  rocprim::plus<long long>(), st));
  HIP_TRY(hipMemcpyAsync(npairs, h.half_off + h.n, sizeof(long long), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return ACCU_OK;
}

int accu_nlist_export_half(accu_nlist* hh, long long* i_out, long long* j_out, void* stream) {
  ARG_CHECK(hh && i_out && j_out, "null argument");
  NList& h = *reinterpret_cast<NList*>(hh);
  ARG_CHECK(h.half_off, "call accu_nlist_count_half first");
  if (h.n == 0) return ACCU_OK;
  const int T = 256;
  hipLaunchKernelGGL(k_emit_half, dim3((h.n + T - 1) / T), dim3(T), 0, (hipStream_t)stream, h.n, h.sorted, h.nn,
                     h.nbr, h.stride, h.half_off, i_out, j_out);
  LAUNCH_CHECK();
  return ACCU_OK;
}

int accu_max_disp2(const double* pos, const double* ref, int n, const double* box, int periodic, double* out_host,
                   void* stream) {
  ARG_CHECK(pos && ref && out_host, "null argument");
  ARG_CHECK(!periodic || box, "periodic needs a box");
  hipStream_t st = (hipStream_t)stream;
  *out_host = 0.0;
  if (n <= 0) return ACCU_OK;
  Workspace* ws = workspace_for(st);
  if (!ws) return ACCU_ERUNTIME;
  int rc = ws->ensure(16);
  if (rc) return rc;
  Box b{};
  b.periodic = periodic;
  for (int d = 0; d < 3; ++d) {
    b.L[d] = box ? box[d] : 1.0;
    b.inv[d] = 1.0 / b.L[d];
  }
  unsigned long long* o = reinterpret_cast<unsigned long long*>(ws->out);
  HIP_TRY(hipMemsetAsync(o, 0, sizeof(unsigned long long), st));
  const int T = 256;
  hipLaunchKernelGGL(k_maxdisp2, dim3((n + T - 1) / T), dim3(T), 0, st, pos, 3, ref, 3, n, b, o);
  LAUNCH_CHECK();
  unsigned long long bits = 0;
  HIP_TRY(hipMemcpyAsync(&bits, o, sizeof bits, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  double v;
  memcpy(&v, &bits, sizeof v);
  *out_host = v;
  return ACCU_OK;
}

}  // extern "C"
