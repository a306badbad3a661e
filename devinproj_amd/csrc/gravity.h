// The LDS-tile loops of the direct gravity sum (forces.py:38-48), shared by the force
// entry points and the device-resident N-body engine, and the launch planning shared by
// accu_gravity_direct and its allocation-free variant.
#pragma once
#include "common.h"

namespace accu {

// One LDS tile of j-particles against target i. The self pair (j == i, excluded by the
// reference's r^2 = inf on the diagonal) can only lie in a tile that holds a target of the
// wave, so only those tiles (SELF, chosen wave-uniformly) pay the index compare and select
// per pair; the others run without them.
template <bool SELF>
__device__ __forceinline__ void grav64_tile(const double4* tile, int cnt, int base, int i, double xi, double yi,
                                            double zi, double eps2, const Box& box, double& fx, double& fy,
                                            double& fz, double& phi) {
  for (int k = 0; k < cnt; ++k) {
    const double4 pj = tile[k];
    double dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
    if (box.periodic) {
      dx = minimg_fast(dx, box.L[0], box.inv[0]);
      dy = minimg_fast(dy, box.L[1], box.inv[1]);
      dz = minimg_fast(dz, box.L[2], box.inv[2]);
    }
    const double r2 = r2_np(dx, dy, dz) + eps2;
    double inv = rsqrt64(r2);
    if constexpr (SELF) inv = (base + k == i) ? 0.0 : inv;
    const double mi3 = pj.w * inv * inv * inv;
    fx = fma(mi3, dx, fx);
    fy = fma(mi3, dy, fy);
    fz = fma(mi3, dz, fz);
    phi = fma(pj.w, inv, phi);
  }
}

template <bool SELF>
__device__ __forceinline__ void grav32_tile(const float4* tile, int cnt, int base, int i, float xi, float yi,
                                            float zi, float e2, float& gx, float& gy, float& gz, float& gp) {
  for (int k = 0; k < cnt; ++k) {
    const float4 pj = tile[k];
    const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, e2)));
    float inv = __builtin_amdgcn_rsqf(r2);
    if constexpr (SELF) inv = (base + k == i) ? 0.f : inv;
    const float mi = pj.w * inv;
    const float mi3 = mi * inv * inv;
    gx = fmaf(mi3, dx, gx);
    gy = fmaf(mi3, dy, gy);
    gz = fmaf(mi3, dz, gz);
    gp += mi;
  }
}

// Scratch that one call needs, in doubles: per-target-block energy partials for the
// grid-wide sum, or the per-split force partials of the few-target path.
struct GravPlan {
  int nb = 0;       // target blocks
  int nsplit = 0;   // > 0: few-target path with this many source splits
  int tps = 0;      // tiles per split
  size_t need = 0;  // doubles of scratch
};

GravPlan plan_gravity(int n, int ntgt, bool has_targets, bool has_U, int precision);
int check_gravity_args(const double* pos, const double* mass, double* F, int precision, int periodic);
// Enqueue the kernels of one call. partials holds plan.need doubles; counter is the
// grid-sum arrival counter (zero on entry, left zero by the sum).
int launch_gravity(const GravPlan& g, const double* pos, const double* mass, int n, const long long* targets,
                   int ntgt, double G, double eps2, const double* box, int periodic, int precision,
                   const double* origin, double* F, double* U, double* partials, unsigned* counter, hipStream_t st);

}  // namespace accu
