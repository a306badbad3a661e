// O(N^2) softened Newtonian gravity, j-particles staged through LDS tiles.
// forces.py:38-48: F_i = sum_j G m_i m_j (r_j - r_i) / (r^2 + eps^2)^(3/2),
//                  U   = -1/2 sum_{i != j} G m_i m_j / sqrt(r^2 + eps^2).
// tree.py:81-87 (direct_subset) is the same sum restricted to target rows.
#include "common.h"
#include "gravity.h"
#include "../../include/accusim_hip.h"

#include <algorithm>

namespace accu {

constexpr int kGT = 256;  // threads per block = j-tile size
// direct_subset-sized calls (fewer target blocks than kSplitBelow, no energy) spread the
// sources over about kSplitBlocks workgroups
constexpr int kSplitBelow = 64, kSplitBlocks = 1024;

Box make_box(const double* box, int periodic);

// Particle of target row r: >= 0 for a valid row, -1 past the last row, -2 for a target
// index outside [0, n) (its force row is written as NaN and nothing is read for it).
__device__ __forceinline__ int target_of(const long long* tgt, int r, int ntgt, int n) {
  if (r >= ntgt) return -1;
  const long long t = tgt ? tgt[r] : r;
  return (t >= 0 && t < n) ? (int)t : -2;
}

__device__ __forceinline__ void store_bad_row(double* F, int r) {
  const double nan = __builtin_nan("");
  F[3 * (size_t)r] = nan;
  F[3 * (size_t)r + 1] = nan;
  F[3 * (size_t)r + 2] = nan;
}

// FP64: every pair term and sum in float64.
__global__ void __launch_bounds__(kGT) k_grav64(const double* __restrict__ pos, const double* __restrict__ m, int n,
                                                const long long* __restrict__ tgt, int ntgt, double G, double eps2,
                                                Box box, double* __restrict__ F, double* partials,
                                                unsigned* counter, double* Uout) {
  __shared__ double4 tile[kGT];
  __shared__ double red[16];
  const int r = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int t = target_of(tgt, r, ntgt, n);
  const bool active = t >= 0;
  const int i = active ? t : 0;
  if (t == -2) store_bad_row(F, r);
  const double xi = pos[3 * (size_t)i], yi = pos[3 * (size_t)i + 1], zi = pos[3 * (size_t)i + 2];
  double fx = 0, fy = 0, fz = 0, phi = 0;
  for (int base = 0; base < n; base += kGT) {
    const int j = base + threadIdx.x;
    if (j < n) tile[threadIdx.x] = make_double4(pos[3 * (size_t)j], pos[3 * (size_t)j + 1], pos[3 * (size_t)j + 2], m[j]);
    __syncthreads();
    const int cnt = (n - base) < kGT ? (n - base) : kGT;
    if (__any(i >= base && i < base + cnt))
      grav64_tile<true>(tile, cnt, base, i, xi, yi, zi, eps2, box, fx, fy, fz, phi);
    else
      grav64_tile<false>(tile, cnt, base, i, xi, yi, zi, eps2, box, fx, fy, fz, phi);
    __syncthreads();
  }
  double U[1] = {0.0};
  if (active) {
    const double gm = G * m[i];
    F[3 * (size_t)r] = gm * fx;
    F[3 * (size_t)r + 1] = gm * fy;
    F[3 * (size_t)r + 2] = gm * fz;
    U[0] = -0.5 * gm * phi;
  }
  if (Uout) grid_sum<1>(U, partials, counter, Uout, red);
}

// Few targets (direct_subset: 200 sampled stars against all of them), no energy: the
// sources are split over `nsplit` ranges of whole tiles so that many workgroups share the
// sum. Block b takes target block b % nb and source split b / nb, and writes its partial
// force sums to part[(split * ntgt + r) * 3 + d]; k_grav_split_sum adds the splits in order.
__global__ void __launch_bounds__(kGT) k_grav64_split(const double* __restrict__ pos, const double* __restrict__ m,
                                                      int n, const long long* __restrict__ tgt, int ntgt, int nb,
                                                      int tiles_per_split, double eps2, Box box,
                                                      double* __restrict__ part) {
  __shared__ double4 tile[kGT];
  const int tb = blockIdx.x % nb, sp = blockIdx.x / nb;
  const int r = tb * kGT + threadIdx.x;
  const int t = target_of(tgt, r, ntgt, n);
  const bool active = t >= 0;
  const int i = active ? t : 0;
  const double xi = pos[3 * (size_t)i], yi = pos[3 * (size_t)i + 1], zi = pos[3 * (size_t)i + 2];
  double fx = 0, fy = 0, fz = 0, phi = 0;
  const int b0 = sp * tiles_per_split * kGT;
  const int b1 = min(n, b0 + tiles_per_split * kGT);
  for (int base = b0; base < b1; base += kGT) {
    const int j = base + threadIdx.x;
    if (j < n) tile[threadIdx.x] = make_double4(pos[3 * (size_t)j], pos[3 * (size_t)j + 1], pos[3 * (size_t)j + 2], m[j]);
    __syncthreads();
    const int cnt = (n - base) < kGT ? (n - base) : kGT;
    if (__any(i >= base && i < base + cnt))
      grav64_tile<true>(tile, cnt, base, i, xi, yi, zi, eps2, box, fx, fy, fz, phi);
    else
      grav64_tile<false>(tile, cnt, base, i, xi, yi, zi, eps2, box, fx, fy, fz, phi);
    __syncthreads();
  }
  if (active) {
    double* o = part + ((size_t)sp * ntgt + r) * 3;
    o[0] = fx;
    o[1] = fy;
    o[2] = fz;
  }
}

__global__ void k_grav_split_sum(const double* __restrict__ m, int n, const long long* __restrict__ tgt, int ntgt,
                                 int nsplit, double G, const double* __restrict__ part, double* __restrict__ F) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = target_of(tgt, r, ntgt, n);
  if (t == -2) store_bad_row(F, r);
  if (t < 0) return;
  double f[3] = {0.0, 0.0, 0.0};
  for (int sp = 0; sp < nsplit; ++sp)
    for (int d = 0; d < 3; ++d) f[d] += part[((size_t)sp * ntgt + r) * 3 + d];
  const double gm = G * m[t];
  for (int d = 0; d < 3; ++d) F[3 * (size_t)r + d] = gm * f[d];
}

// Mixed: float32 pair terms summed per tile in float32, tile sums added in float64.
// Positions are taken relative to `origin` (the centre of mass) before rounding.
__global__ void __launch_bounds__(kGT) k_grav32(const double* __restrict__ pos, const double* __restrict__ m, int n,
                                                const long long* __restrict__ tgt, int ntgt, double G, double eps2,
                                                double ox, double oy, double oz, double* __restrict__ F,
                                                double* partials, unsigned* counter, double* Uout) {
  __shared__ float4 tile[kGT];
  __shared__ double red[16];
  const int r = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int t = target_of(tgt, r, ntgt, n);
  const bool active = t >= 0;
  const int i = active ? t : 0;
  if (t == -2) store_bad_row(F, r);
  const float xi = (float)(pos[3 * (size_t)i] - ox), yi = (float)(pos[3 * (size_t)i + 1] - oy),
              zi = (float)(pos[3 * (size_t)i + 2] - oz);
  const float e2 = (float)eps2;
  double fx = 0, fy = 0, fz = 0, phi = 0;
  for (int base = 0; base < n; base += kGT) {
    const int j = base + threadIdx.x;
    if (j < n)
      tile[threadIdx.x] = make_float4((float)(pos[3 * (size_t)j] - ox), (float)(pos[3 * (size_t)j + 1] - oy),
                                      (float)(pos[3 * (size_t)j + 2] - oz), (float)m[j]);
    __syncthreads();
    const int cnt = (n - base) < kGT ? (n - base) : kGT;
    float gx = 0.f, gy = 0.f, gz = 0.f, gp = 0.f;
    if (__any(i >= base && i < base + cnt))
      grav32_tile<true>(tile, cnt, base, i, xi, yi, zi, e2, gx, gy, gz, gp);
    else
      grav32_tile<false>(tile, cnt, base, i, xi, yi, zi, e2, gx, gy, gz, gp);
    fx += (double)gx;
    fy += (double)gy;
    fz += (double)gz;
    phi += (double)gp;
    __syncthreads();
  }
  double U[1] = {0.0};
  if (active) {
    const double gm = G * m[i];
    F[3 * (size_t)r] = gm * fx;
    F[3 * (size_t)r + 1] = gm * fy;
    F[3 * (size_t)r + 2] = gm * fz;
    U[0] = -0.5 * gm * phi;
  }
  if (Uout) grid_sum<1>(U, partials, counter, Uout, red);
}

// ---- launch planning shared by both entry points ---------------------------------

GravPlan plan_gravity(int n, int ntgt, bool has_targets, bool has_U, int precision) {
  GravPlan g;
  g.nb = (ntgt + kGT - 1) / kGT;
  const int ntiles = (n + kGT - 1) / kGT;
  if (precision == ACCU_FP64 && has_targets && !has_U && g.nb < kSplitBelow && ntiles > 1) {
    // few target blocks: split the sources so that about kSplitBlocks workgroups share them
    const int want = std::max(1, std::min(ntiles, kSplitBlocks / g.nb));
    g.tps = (ntiles + want - 1) / want;
    g.nsplit = (ntiles + g.tps - 1) / g.tps;
    g.need = (size_t)g.nsplit * ntgt * 3;
  } else {
    g.need = has_U ? (size_t)g.nb : 0;
  }
  return g;
}

// Enqueue the kernels of one call. partials holds g.need doubles; counter is the grid-sum
// arrival counter (zero on entry, left zero by the sum).
int launch_gravity(const GravPlan& g, const double* pos, const double* mass, int n, const long long* targets,
                   int ntgt, double G, double eps2, const double* box, int periodic, int precision,
                   const double* origin, double* F, double* U, double* partials, unsigned* counter, hipStream_t st) {
  if (g.nsplit > 0) {
    Box b = make_box(box, periodic);
    hipLaunchKernelGGL(k_grav64_split, dim3(g.nb * g.nsplit), dim3(kGT), 0, st, pos, mass, n, targets, ntgt, g.nb,
                       g.tps, eps2, b, partials);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_grav_split_sum, dim3(g.nb), dim3(kGT), 0, st, mass, n, targets, ntgt, g.nsplit, G, partials,
                       F);
  } else if (precision == ACCU_FP64) {
    Box b = make_box(box, periodic);
    hipLaunchKernelGGL(k_grav64, dim3(g.nb), dim3(kGT), 0, st, pos, mass, n, targets, ntgt, G, eps2, b, F, partials,
                       counter, U);
  } else {
    double o[3] = {0, 0, 0};
    if (origin)
      for (int d = 0; d < 3; ++d) o[d] = origin[d];
    hipLaunchKernelGGL(k_grav32, dim3(g.nb), dim3(kGT), 0, st, pos, mass, n, targets, ntgt, G, eps2, o[0], o[1], o[2],
                       F, partials, counter, U);
  }
  LAUNCH_CHECK();
  return ACCU_OK;
}

int check_gravity_args(const double* pos, const double* mass, double* F, int precision, int periodic) {
  ARG_CHECK(pos && mass && F, "null argument");
  ARG_CHECK(precision == ACCU_FP64 || precision == ACCU_FP32_PAIR, "unknown precision %d", precision);
  ARG_CHECK(!(precision == ACCU_FP32_PAIR && periodic), "float32 gravity supports open boundaries only");
  return ACCU_OK;
}

}  // namespace accu

using namespace accu;

extern "C" {

int accu_gravity_direct(const double* pos, const double* mass, int n, const long long* targets, int ntgt, double G,
                        double eps2, const double* box, int periodic, int precision, const double* origin,
                        double* F, double* U, void* stream) {
  int rc = check_gravity_args(pos, mass, F, precision, periodic);
  if (rc) return rc;
  if (!targets) ntgt = n;
  if (n <= 0 || ntgt <= 0) return ACCU_OK;
  const GravPlan g = plan_gravity(n, ntgt, targets != nullptr, U != nullptr, precision);
  Workspace* ws = workspace_for((hipStream_t)stream);
  if (!ws) return ACCU_ERUNTIME;
  rc = ws->ensure(g.need);
  if (rc) return rc;
  return launch_gravity(g, pos, mass, n, targets, ntgt, G, eps2, box, periodic, precision, origin, F, U, ws->partials,
                        ws->counter, (hipStream_t)stream);
}

}  // extern "C"
