// Launch planning of the direct gravity sum, shared by accu_gravity_direct and its
// allocation-free variant.
#pragma once
#include "common.h"

namespace accu {

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
