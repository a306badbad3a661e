// Device-resident molecular dynamics: fused force + velocity-Verlet / Yoshida4 step.
//
// One launch per (sub)step s does, for every atom (integrators.py:25-31 regrouped):
//   F_s = F(x_s)                         (LJ and/or DSF over the Verlet list)
//   a_s = F_s / m;  v += (dt_s / 2) a_s   (closing half kick of substep s)
//   record PE(x_s), KE(v), momentum      (simulation.py:76-93, when asked)
//   v += (dt_{s+1} / 2) a_s;  x_{s+1} = (x_s + dt_{s+1} v) mod box  (next substep's
//                                        opening half kick, drift and wrap)
//   flag a rebuild when |min-image(x_{s+1} - x_build)|^2 >= (skin/2)^2 (forces.py:90-94)
// Positions are double-buffered (read x_s, write x_{s+1}). A flag-gated chain of
// kernels then re-bins, re-sorts all per-atom state into cell order and rebuilds the
// list; it costs only empty launches on steps that need no rebuild, so a whole run
// can be captured into one hipGraph.
#include "common.h"
#include "nlist.h"
#include "pairkern.h"
#include "../../include/accusim_hip.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace accu {

constexpr int kMT = 256;
constexpr int kLogW = 8;  // PE, KE, Px, Py, Pz, rebuilds-so-far, spare, spare

Box make_box(const double* box, int periodic);

struct MDParams {
  int n;
  int ns;
  double mass[kMaxSpecies];
  double charge[kMaxSpecies];
  int has_lj, has_coul;
  LJParams lj;
  CoulParams coul;
  Box box;
  double half_skin2;  // (skin / 2)^2
};

__device__ __forceinline__ double np_mod(double x, double L) {
  // numpy remainder for float64 (sign of the divisor)
  double r = fmod(x, L);
  if (r != 0.0) {
    if ((r < 0.0) != (L < 0.0)) r += L;
  } else {
    r = copysign(0.0, L);
  }
  return r;
}

// Accumulators of one atom's pair sums: float64 totals plus float32 partials that the
// mixed path flushes into the totals after every batch of neighbours.
struct PairAcc {
  double fx = 0, fy = 0, fz = 0, u = 0;   // LJ
  double cx = 0, cy = 0, cz = 0, cu = 0;  // Coulomb (times k at the end)
  float gx = 0.f, gy = 0.f, gz = 0.f, gu = 0.f, hx = 0.f, hy = 0.f, hz = 0.f, hu = 0.f;
  __device__ __forceinline__ void flush() {
    fx += gx; fy += gy; fz += gz; u += gu;
    cx += hx; cy += hy; cz += hz; cu += hu;
    gx = gy = gz = gu = hx = hy = hz = hu = 0.f;
  }
};

// One neighbour j (position xj, species in xj.w) of atom i.
template <typename P, bool LJ_ON, bool COUL_ON, bool ONE_SP>
__device__ __forceinline__ void md_pair(const MDParams& p, const LJ<P>& lj, const double4& xi, int row, double qi,
                                        const double4& xj, PairAcc& a) {
  double dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
  if (p.box.periodic) {
    dx = minimg_fast(dx, p.box.L[0], p.box.inv[0]);
    dy = minimg_fast(dy, p.box.L[1], p.box.inv[1]);
    dz = minimg_fast(dz, p.box.L[2], p.box.inv[2]);
  }
  const int tj = ONE_SP ? 0 : (int)xj.w;
  if constexpr (sizeof(P) == 8) {
    const double r2 = r2_np(dx, dy, dz);
    if constexpr (LJ_ON) {
      if constexpr (ONE_SP) lj_pair64(dx, dy, dz, r2, p.lj.eps[0], p.lj.sig2[0], p.lj.rc2[0], p.lj.shift, a.fx, a.fy, a.fz, a.u);
      else lj_pair64(dx, dy, dz, r2, lj.eps[row + tj], lj.sig2[row + tj], lj.rc2[row + tj], lj.shift, a.fx, a.fy, a.fz, a.u);
    }
    if constexpr (COUL_ON) coul_pair64(p.coul, dx, dy, dz, r2, qi * p.charge[tj], a.cx, a.cy, a.cz, a.cu);
  } else {
    const float fdx = (float)dx, fdy = (float)dy, fdz = (float)dz;
    const float r2 = r2_npf(fdx, fdy, fdz);
    if constexpr (LJ_ON) {
      if constexpr (ONE_SP)
        lj_pair32(fdx, fdy, fdz, r2, (float)p.lj.eps[0], (float)p.lj.sig2[0], (float)p.lj.rc2[0], (float)p.lj.shift,
                  a.gx, a.gy, a.gz, a.gu);
      else
        lj_pair32(fdx, fdy, fdz, r2, lj.eps[row + tj], lj.sig2[row + tj], lj.rc2[row + tj], lj.shift, a.gx, a.gy, a.gz,
                  a.gu);
    }
    if constexpr (COUL_ON)
      coul_pair32((float)p.coul.alpha, (float)p.coul.rc, (float)p.coul.c, (float)p.coul.erc_rc, (float)p.coul.fshift,
                  p.coul.shifted, fdx, fdy, fdz, r2, (float)(qi * p.charge[tj]), a.hx, a.hy, a.hz, a.hu);
  }
}

template <typename P, bool LJ_ON, bool COUL_ON, bool ONE_SP>
__global__ void __launch_bounds__(kMT) k_md_fused(MDParams p, const double4* __restrict__ pos,
                                                  double4* __restrict__ pos_next, double4* __restrict__ vel,
                                                  double4* __restrict__ Fst, const int* __restrict__ sp,
                                                  const double4* __restrict__ ref, const int* __restrict__ nn,
                                                  const int* __restrict__ nbr, int stride, double h_cur,
                                                  double h_next, double dt_next, int do_drift, int record,
                                                  int* __restrict__ flag, double* partials, unsigned* counter,
                                                  double* log, int* log_count, const int* nrebuild) {
  __shared__ double lds[3 * kMaxSpecies * kMaxSpecies];
  __shared__ double red[kLogW * 4];
  LJ<P> lj;
  if constexpr (LJ_ON && !ONE_SP) {
    lj.sp = sp;
    lj.stage(p.lj, lds);
  }
  __syncthreads();
  const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  bool moved_far = false;
  if (i < p.n) {
    const double4 xi = pos[i];
    const int ti = ONE_SP ? 0 : (int)xi.w;
    const int row = ti * p.ns;
    const double qi = p.charge[ti];
    const int cnt = nn[i];
    const int* __restrict__ col = nbr + i;
    PairAcc a;
    // batches of 4 neighbours: the four index loads, then the four position gathers,
    // are independent, so each lane keeps several memory requests in flight
    int k = 0;
    for (; k + 4 <= cnt; k += 4) {
      const int j0 = col[(size_t)k * stride], j1 = col[(size_t)(k + 1) * stride];
      const int j2 = col[(size_t)(k + 2) * stride], j3 = col[(size_t)(k + 3) * stride];
      const double4 x0 = pos[j0], x1 = pos[j1], x2 = pos[j2], x3 = pos[j3];
      md_pair<P, LJ_ON, COUL_ON, ONE_SP>(p, lj, xi, row, qi, x0, a);
      md_pair<P, LJ_ON, COUL_ON, ONE_SP>(p, lj, xi, row, qi, x1, a);
      md_pair<P, LJ_ON, COUL_ON, ONE_SP>(p, lj, xi, row, qi, x2, a);
      md_pair<P, LJ_ON, COUL_ON, ONE_SP>(p, lj, xi, row, qi, x3, a);
      if constexpr (sizeof(P) == 4) a.flush();
    }
    for (; k < cnt; ++k) md_pair<P, LJ_ON, COUL_ON, ONE_SP>(p, lj, xi, row, qi, pos[col[(size_t)k * stride]], a);
    a.flush();
    double fx = a.fx, fy = a.fy, fz = a.fz, u = a.u;
    double cx = a.cx, cy = a.cy, cz = a.cz, cu = a.cu;
    if constexpr (COUL_ON) {
      cu -= qi * qi * p.coul.self;
      fx += p.coul.k * cx;
      fy += p.coul.k * cy;
      fz += p.coul.k * cz;
      u += p.coul.k * cu;
    }
    const double m = p.mass[ti];
    const double ax = fx / m, ay = fy / m, az = fz / m;
    double4 v = vel[i];
    {
#pragma clang fp contract(off)
      v.x = v.x + h_cur * ax;
      v.y = v.y + h_cur * ay;
      v.z = v.z + h_cur * az;
    }
    if (record) {
      acc[0] = u;
      acc[1] = 0.5 * m * ((v.x * v.x + v.z * v.z) + v.y * v.y);
      acc[2] = m * v.x;
      acc[3] = m * v.y;
      acc[4] = m * v.z;
    }
    Fst[i] = make_double4(fx, fy, fz, 0.0);
    if (do_drift) {
      double4 x = xi;
      {
#pragma clang fp contract(off)
        v.x = v.x + h_next * ax;
        v.y = v.y + h_next * ay;
        v.z = v.z + h_next * az;
        x.x = x.x + dt_next * v.x;
        x.y = x.y + dt_next * v.y;
        x.z = x.z + dt_next * v.z;
      }
      if (p.box.periodic) {
        x.x = np_mod(x.x, p.box.L[0]);
        x.y = np_mod(x.y, p.box.L[1]);
        x.z = np_mod(x.z, p.box.L[2]);
      }
      pos_next[i] = x;
      const double4 r = ref[i];
      double d0, d1, d2;
      {
#pragma clang fp contract(off)
        d0 = x.x - r.x;
        d1 = x.y - r.y;
        d2 = x.z - r.z;
      }
      if (p.box.periodic) {
        d0 = minimg_exact(d0, p.box.L[0]);
        d1 = minimg_exact(d1, p.box.L[1]);
        d2 = minimg_exact(d2, p.box.L[2]);
      }
      moved_far = !(r2_np(d0, d1, d2) < p.half_skin2);
    }
    vel[i] = v;
  }
  if (do_drift && __any(moved_far) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
  if (!record) return;
  // deterministic grid reduction of the record, appended to the log
  block_sum<5>(acc, red);
  __shared__ int am_last;
  if (threadIdx.x == 0) {
    for (int k = 0; k < 5; ++k) partials[(size_t)blockIdx.x * 5 + k] = acc[k];
    __threadfence();
    am_last = (atomicAdd(counter, 1u) == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return;
  __threadfence();
  double tot[5] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
    for (int k = 0; k < 5; ++k)
      tot[k] += __hip_atomic_load(&partials[(size_t)b * 5 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  block_sum<5>(tot, red);
  if (threadIdx.x == 0) {
    const int idx = *log_count;
    double* o = log + (size_t)idx * kLogW;
    for (int k = 0; k < 5; ++k) o[k] = tot[k];
    o[5] = (double)*nrebuild;
    o[6] = 0.0;
    o[7] = 0.0;
    *log_count = idx + 1;
    *counter = 0u;
  }
}

// Opening half kick + drift of the first substep (state at a full step on entry).
__global__ void __launch_bounds__(kMT) k_md_drift(MDParams p, const double4* __restrict__ pos,
                                                  double4* __restrict__ pos_next, double4* __restrict__ vel,
                                                  const double4* __restrict__ Fst, const int* __restrict__ sp,
                                                  const double4* __restrict__ ref, double h_next, double dt_next,
                                                  int* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool moved_far = false;
  if (i < p.n) {
    const double m = p.mass[sp[i]];
    const double4 f = Fst[i];
    const double ax = f.x / m, ay = f.y / m, az = f.z / m;
    double4 v = vel[i], x = pos[i];
    {
#pragma clang fp contract(off)
      v.x = v.x + h_next * ax;
      v.y = v.y + h_next * ay;
      v.z = v.z + h_next * az;
      x.x = x.x + dt_next * v.x;
      x.y = x.y + dt_next * v.y;
      x.z = x.z + dt_next * v.z;
    }
    if (p.box.periodic) {
      x.x = np_mod(x.x, p.box.L[0]);
      x.y = np_mod(x.y, p.box.L[1]);
      x.z = np_mod(x.z, p.box.L[2]);
    }
    vel[i] = v;
    pos_next[i] = x;
    const double4 r = ref[i];
    double d0, d1, d2;
    {
#pragma clang fp contract(off)
      d0 = x.x - r.x;
      d1 = x.y - r.y;
      d2 = x.z - r.z;
    }
    if (p.box.periodic) {
      d0 = minimg_exact(d0, p.box.L[0]);
      d1 = minimg_exact(d1, p.box.L[1]);
      d2 = minimg_exact(d2, p.box.L[2]);
    }
    moved_far = !(r2_np(d0, d1, d2) < p.half_skin2);
  }
  if (__any(moved_far) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// Gather per-atom state into cell order (flag-gated).
__global__ void k_md_permute(int n, const int* __restrict__ sorted, const double4* __restrict__ pos,
                             const double4* __restrict__ vel, const double4* __restrict__ F, const int* __restrict__ sp,
                             const int* __restrict__ perm, const int* __restrict__ cell, double4* __restrict__ pos2,
                             double4* __restrict__ vel2, double4* __restrict__ F2, int* __restrict__ sp2,
                             int* __restrict__ perm2, int* __restrict__ cell2, const int* __restrict__ flag) {
  if (*flag == 0) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = sorted[t];
  pos2[t] = pos[i];
  vel2[t] = vel[i];
  F2[t] = F[i];
  sp2[t] = sp[i];
  perm2[t] = perm[i];
  cell2[t] = cell[i];
}

__global__ void k_md_copyback(int n, double4* __restrict__ pos, double4* __restrict__ ref, double4* __restrict__ vel,
                              double4* __restrict__ F, int* __restrict__ sp, int* __restrict__ perm,
                              int* __restrict__ cell, const double4* __restrict__ pos2,
                              const double4* __restrict__ vel2, const double4* __restrict__ F2,
                              const int* __restrict__ sp2, const int* __restrict__ perm2,
                              const int* __restrict__ cell2, const int* __restrict__ flag) {
  if (*flag == 0) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  pos[t] = pos2[t];
  ref[t] = pos2[t];
  vel[t] = vel2[t];
  F[t] = F2[t];
  sp[t] = sp2[t];
  perm[t] = perm2[t];
  cell[t] = cell2[t];
}

__global__ void k_md_end_rebuild(int* flag, int* nrebuild) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (*flag) {
      *nrebuild += 1;
      *flag = 0;
    }
  }
}

__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_md_scatter_user(int n, const int* __restrict__ perm, const double4* __restrict__ a,
                                  double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int u = perm[t];
  const double4 v = a[t];
  out[3 * (size_t)u] = v.x;
  out[3 * (size_t)u + 1] = v.y;
  out[3 * (size_t)u + 2] = v.z;
}

struct MD {
  MDParams p{};
  int precision = ACCU_FP64;
  double rc_list = 0.0, skin = 0.0;
  CellGeo geo{};
  NList nl;
  hipStream_t st = nullptr;
  // state (slot order)
  double4 *pos = nullptr, *pos_b = nullptr, *vel = nullptr, *F = nullptr, *ref = nullptr;
  int *sp = nullptr, *perm = nullptr;
  // rebuild scratch
  double4 *pos2 = nullptr, *vel2 = nullptr, *F2 = nullptr;
  int *sp2 = nullptr, *perm2 = nullptr, *cell2 = nullptr;
  int* ints = nullptr;  // [0] flag, [1] nrebuild, [2] log_count
  double* log = nullptr;
  int log_cap = 0;
  double* partials = nullptr;
  unsigned* counter = nullptr;
  bool ready = false;
  // optional HIP-event timing of every fused launch (host-driven runs only)
  bool timing = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  // graph
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  long long graph_key[4] = {0, 0, 0, 0};
  double4* graph_pos = nullptr;

  int alloc(int n) {
    auto A = [&](void** q, size_t bytes) -> int {
      HIP_TRY(hipMalloc(q, bytes));
      return ACCU_OK;
    };
    int rc;
    const size_t d4 = (size_t)n * sizeof(double4), i4 = (size_t)n * sizeof(int);
    void** d4s[] = {(void**)&pos, (void**)&pos_b, (void**)&vel, (void**)&F, (void**)&ref,
                    (void**)&pos2, (void**)&vel2, (void**)&F2};
    for (void** q : d4s)
      if ((rc = A(q, d4))) return rc;
    void** i4s[] = {(void**)&sp, (void**)&perm, (void**)&sp2, (void**)&perm2, (void**)&cell2};
    for (void** q : i4s)
      if ((rc = A(q, i4))) return rc;
    if ((rc = A((void**)&ints, 64))) return rc;
    HIP_TRY(hipMemset(ints, 0, 64));
    const int nb = (n + kMT - 1) / kMT;
    if ((rc = A((void**)&partials, (size_t)nb * 8 * sizeof(double)))) return rc;
    if ((rc = A((void**)&counter, 64))) return rc;
    HIP_TRY(hipMemset(counter, 0, 64));
    return ACCU_OK;
  }

  void release() {
    void* ps[] = {pos, pos_b, vel, F, ref, pos2, vel2, F2, sp, perm, sp2, perm2, cell2, ints, log, partials, counter};
    for (void* q : ps)
      if (q) (void)hipFree(q);
    nl.release();
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (st) (void)hipStreamDestroy(st);
  }

  template <typename P, bool ONE>
  void launch_fused_t(const double4* x, double4* xn, double hc, double hn, double dtn, int drift, int rec) {
    const int nb = (p.n + kMT - 1) / kMT;
    if (p.has_lj && p.has_coul)
      hipLaunchKernelGGL((k_md_fused<P, true, true, ONE>), dim3(nb), dim3(kMT), 0, st, p, x, xn, vel, F, sp, ref,
                         nl.nn, nl.nbr, nl.stride, hc, hn, dtn, drift, rec, ints, partials, counter, log, ints + 2,
                         ints + 1);
    else if (p.has_lj)
      hipLaunchKernelGGL((k_md_fused<P, true, false, ONE>), dim3(nb), dim3(kMT), 0, st, p, x, xn, vel, F, sp, ref,
                         nl.nn, nl.nbr, nl.stride, hc, hn, dtn, drift, rec, ints, partials, counter, log, ints + 2,
                         ints + 1);
    else
      hipLaunchKernelGGL((k_md_fused<P, false, true, ONE>), dim3(nb), dim3(kMT), 0, st, p, x, xn, vel, F, sp, ref,
                         nl.nn, nl.nbr, nl.stride, hc, hn, dtn, drift, rec, ints, partials, counter, log, ints + 2,
                         ints + 1);
  }

  void launch_fused(const double4* x, double4* xn, double hc, double hn, double dtn, int drift, int rec) {
    const bool one = p.ns == 1;
    if (precision == ACCU_FP64) {
      if (one) launch_fused_t<double, true>(x, xn, hc, hn, dtn, drift, rec);
      else launch_fused_t<double, false>(x, xn, hc, hn, dtn, drift, rec);
    } else {
      if (one) launch_fused_t<float, true>(x, xn, hc, hn, dtn, drift, rec);
      else launch_fused_t<float, false>(x, xn, hc, hn, dtn, drift, rec);
    }
  }

  // Flag-gated rebuild chain: bin, sort, permute state into cell order, rebuild the list.
  void launch_rebuild(double4* x) {
    const int n = p.n, T = 256, nb = (n + T - 1) / T;
    launch_bin_sort(nl, reinterpret_cast<const double*>(x), 4, n, geo, ints, st);
    hipLaunchKernelGGL(k_md_permute, dim3(nb), dim3(T), 0, st, n, nl.sorted, x, vel, F, sp, perm, nl.cell, pos2, vel2,
                       F2, sp2, perm2, cell2, ints);
    hipLaunchKernelGGL(k_md_copyback, dim3(nb), dim3(T), 0, st, n, x, ref, vel, F, sp, perm, nl.cell, pos2, vel2, F2,
                       sp2, perm2, cell2, ints);
    launch_build_ell(nl, reinterpret_cast<const double*>(x), 4, n, geo, nl.rc2, nullptr, ints, st, 0);
    hipLaunchKernelGGL(k_md_end_rebuild, dim3(1), dim3(64), 0, st, ints, ints + 1);
  }

  // Enqueue `steps` steps with substep weights w[0..nsub) (VV: {1}; Yoshida4: {w1, w0, w1}).
  // On entry x = pos holds a full-step state with F valid; on exit the same holds.
  void enqueue_steps(int steps, double dt, const double* w, int nsub, int record_every, int first_record_step) {
    double4* cur = pos;
    double4* nxt = pos_b;
    const int nb = (p.n + kMT - 1) / kMT;
    // opening drift of the first substep
    const double dt0 = w[0] * dt;
    hipLaunchKernelGGL(k_md_drift, dim3(nb), dim3(kMT), 0, st, p, cur, nxt, vel, F, sp, ref, 0.5 * dt0, dt0, ints);
    std::swap(cur, nxt);
    launch_rebuild(cur);
    for (int s = 0; s < steps; ++s) {
      for (int k = 0; k < nsub; ++k) {
        const double dtc = w[k] * dt;
        const bool last = (s == steps - 1) && (k == nsub - 1);
        const double dtn = w[(k + 1) % nsub] * dt;
        const int step_no = first_record_step + s + 1;
        const int rec = (k == nsub - 1) && record_every > 0 && (step_no % record_every == 0 || s == steps - 1);
        if (timing) {
          if (ev_used + 2 > ev.size()) {
            hipEvent_t a, b;
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            ev.push_back(a);
            ev.push_back(b);
          }
          (void)hipEventRecord(ev[ev_used], st);
        }
        launch_fused(cur, nxt, 0.5 * dtc, 0.5 * dtn, dtn, last ? 0 : 1, rec);
        if (timing) {
          (void)hipEventRecord(ev[ev_used + 1], st);
          ev_used += 2;
        }
        if (!last) {
          std::swap(cur, nxt);
          launch_rebuild(cur);
        }
      }
    }
    // an odd number of swaps leaves the state in pos_b: swap the buffers' roles
    if (cur != pos) std::swap(pos, pos_b);
  }
};

}  // namespace accu

using namespace accu;

extern "C" {

int accu_md_create(accu_md** out, int n, const double* pos_host, const double* vel_host, const int* sp_host, int ns,
                   const double* mass_host, const double* charge_host, const double* box_host, int periodic) {
  ARG_CHECK(out && pos_host && vel_host && sp_host && mass_host, "null argument");
  ARG_CHECK(n >= 2, "need at least 2 atoms");
  ARG_CHECK(ns >= 1 && ns <= kMaxSpecies, "number of species %d outside [1, %d]", ns, kMaxSpecies);
  ARG_CHECK(periodic && box_host, "the device MD engine needs a periodic box");
  for (int i = 0; i < n; ++i)
    ARG_CHECK(sp_host[i] >= 0 && sp_host[i] < ns, "species index %d of atom %d outside [0, %d)", sp_host[i], i, ns);
  MD* m = new MD();
  m->p.n = n;
  m->p.ns = ns;
  for (int k = 0; k < ns; ++k) {
    m->p.mass[k] = mass_host[k];
    m->p.charge[k] = charge_host ? charge_host[k] : 0.0;
  }
  m->p.box = make_box(box_host, periodic);
  int rc = m->alloc(n);
  if (rc) {
    m->release();
    delete m;
    return rc;
  }
  hipError_t e = hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error("hipStreamCreate: %s", hipGetErrorString(e));
    m->release();
    delete m;
    return ACCU_ERUNTIME;
  }
  std::vector<double4> h4(n);
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) {
    // the species index rides in pos.w so one gather brings position and type
    h4[i] = make_double4(pos_host[3 * (size_t)i], pos_host[3 * (size_t)i + 1], pos_host[3 * (size_t)i + 2],
                         (double)sp_host[i]);
    perm[i] = i;
  }
  HIP_TRY(hipMemcpy(m->pos, h4.data(), n * sizeof(double4), hipMemcpyHostToDevice));
  for (int i = 0; i < n; ++i)
    h4[i] = make_double4(vel_host[3 * (size_t)i], vel_host[3 * (size_t)i + 1], vel_host[3 * (size_t)i + 2], 0.0);
  HIP_TRY(hipMemcpy(m->vel, h4.data(), n * sizeof(double4), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(m->sp, sp_host, n * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(m->perm, perm.data(), n * sizeof(int), hipMemcpyHostToDevice));
  *out = reinterpret_cast<accu_md*>(m);
  return ACCU_OK;
}

int accu_md_destroy(accu_md* h) {
  if (h) {
    MD* m = reinterpret_cast<MD*>(h);
    m->release();
    delete m;
  }
  return ACCU_OK;
}

int accu_md_set_lj(accu_md* h, int ns, const double* eps, const double* sig2, const double* rc2, double shift) {
  ARG_CHECK(h && eps && sig2 && rc2, "null argument");
  MD& m = *reinterpret_cast<MD*>(h);
  ARG_CHECK(ns == m.p.ns, "species table size %d != %d", ns, m.p.ns);
  m.p.has_lj = 1;
  m.p.lj.ns = ns;
  m.p.lj.shift = shift;
  for (int k = 0; k < ns * ns; ++k) {
    m.p.lj.eps[k] = eps[k];
    m.p.lj.sig2[k] = sig2[k];
    m.p.lj.rc2[k] = rc2[k];
  }
  return ACCU_OK;
}

int accu_md_set_dsf(accu_md* h, double alpha, double rcut, double erc, double k) {
  ARG_CHECK(h, "null handle");
  MD& m = *reinterpret_cast<MD*>(h);
  CoulParams& c = m.p.coul;
  c = CoulParams{};
  c.alpha = alpha;
  c.rc = rcut;
  c.c = 2 * alpha / std::sqrt(M_PI);
  c.erc_rc = erc / rcut;
  c.fshift = erc / (rcut * rcut) + c.c * std::exp(-(alpha * rcut) * (alpha * rcut)) / rcut;
  c.self = erc / (2 * rcut) + alpha / std::sqrt(M_PI);
  c.k = k;
  c.shifted = 1;
  m.p.has_coul = 1;
  return ACCU_OK;
}

// rc_list includes the skin; the list is rebuilt when an atom has moved skin/2.
// slack: neighbour capacity = max count at the first build * (1 + slack) + 16.
int accu_md_init(accu_md* h, double rc_list, double skin, int precision, double slack, int log_cap) {
  ARG_CHECK(h, "null handle");
  MD& m = *reinterpret_cast<MD*>(h);
  ARG_CHECK(m.p.has_lj || m.p.has_coul, "no interaction set");
  ARG_CHECK(precision == ACCU_FP64 || precision == ACCU_FP32_PAIR, "unknown precision %d", precision);
  ARG_CHECK(rc_list > 0 && skin >= 0, "bad list radius");
  m.precision = precision;
  m.rc_list = rc_list;
  m.skin = skin;
  m.p.half_skin2 = skin > 0 ? (0.5 * skin) * (0.5 * skin) : -1.0;  // skin 0: rebuild every step
  const int n = m.p.n;
  double lo[3] = {0, 0, 0}, span[3];
  int nc[3];
  for (int d = 0; d < 3; ++d) {
    span[d] = m.p.box.L[d];
    nc[d] = (int)std::floor(span[d] / rc_list);
    nc[d] = nc[d] < 1 ? 1 : nc[d];
    if (nc[d] < 3) nc[d] = 1;
  }
  int rc = make_geo(m.geo, lo, span, nc, m.p.box.L, 1);
  if (rc) return rc;
  m.nl.n = n;
  m.nl.ncells = (long long)nc[0] * nc[1] * nc[2];
  m.nl.geo = m.geo;
  m.nl.rc2 = rc_list * rc_list;
  if ((rc = m.nl.ensure(n, m.nl.ncells))) return rc;
  if (m.log) HIP_TRY(hipFree(m.log));
  m.log_cap = log_cap > 16 ? log_cap : 16;
  HIP_TRY(hipMalloc(&m.log, (size_t)m.log_cap * kLogW * sizeof(double)));
  // first build: sort state into cell order, measure the neighbour count, size the list
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, m.st, m.ints, 1);
  if ((rc = launch_bin_sort(m.nl, reinterpret_cast<const double*>(m.pos), 4, n, m.geo, m.ints, m.st))) return rc;
  const int T = 256, nb = (n + T - 1) / T;
  hipLaunchKernelGGL(k_md_permute, dim3(nb), dim3(T), 0, m.st, n, m.nl.sorted, m.pos, m.vel, m.F, m.sp, m.perm,
                     m.nl.cell, m.pos2, m.vel2, m.F2, m.sp2, m.perm2, m.cell2, m.ints);
  hipLaunchKernelGGL(k_md_copyback, dim3(nb), dim3(T), 0, m.st, n, m.pos, m.ref, m.vel, m.F, m.sp, m.perm, m.nl.cell,
                     m.pos2, m.vel2, m.F2, m.sp2, m.perm2, m.cell2, m.ints);
  if ((rc = m.nl.ensure_ell(n, 1))) return rc;
  HIP_TRY(hipMemsetAsync(m.nl.dev_scalars, 0, sizeof(int), m.st));
  if ((rc = launch_build_ell(m.nl, reinterpret_cast<const double*>(m.pos), 4, n, m.geo, m.nl.rc2, nullptr, m.ints,
                             m.st, 0)))
    return rc;
  int mx = 0;
  HIP_TRY(hipMemcpyAsync(&mx, m.nl.dev_scalars, sizeof(int), hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipStreamSynchronize(m.st));
  const int cap = (int)(mx * (1.0 + slack)) + 16;
  if ((rc = m.nl.ensure_ell(n, cap))) return rc;
  if ((rc = launch_build_ell(m.nl, reinterpret_cast<const double*>(m.pos), 4, n, m.geo, m.nl.rc2, nullptr, m.ints,
                             m.st, 0)))
    return rc;
  hipLaunchKernelGGL(k_md_end_rebuild, dim3(1), dim3(64), 0, m.st, m.ints, m.ints + 1);
  // forces at t = 0 and the first record (no kick, no drift)
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, m.st, m.ints + 2, 0);
  m.launch_fused(m.pos, m.pos_b, 0.0, 0.0, 0.0, 0, 1);
  LAUNCH_CHECK();
  HIP_TRY(hipStreamSynchronize(m.st));
  m.ready = true;
  return ACCU_OK;
}

// Run `steps` steps (weights w[0..nsub) of dt per substep), recording every
// `record_every` steps and the last. use_graph != 0 captures the launches into a
// hipGraph once per (steps, dt, nsub, record_every) and replays it.
int accu_md_run(accu_md* h, int steps, double dt, const double* w_host, int nsub, int record_every, int use_graph) {
  ARG_CHECK(h && w_host, "null argument");
  MD& m = *reinterpret_cast<MD*>(h);
  ARG_CHECK(m.ready, "call accu_md_init first");
  ARG_CHECK(steps >= 1 && nsub >= 1 && nsub <= 8, "bad steps/nsub");
  int cnt = 0;
  HIP_TRY(hipMemcpyAsync(&cnt, m.ints + 2, sizeof(int), hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipStreamSynchronize(m.st));
  const int nrec = record_every > 0 ? steps / record_every + 2 : 0;
  ARG_CHECK(cnt + nrec <= m.log_cap, "energy log full (%d + %d > %d records)", cnt, nrec, m.log_cap);
  if (!use_graph) {
    m.enqueue_steps(steps, dt, w_host, nsub, record_every, 0);
    LAUNCH_CHECK();
    return ACCU_OK;
  }
  // The captured launches have the position buffers' roles baked in. A block of
  // steps swaps the roles steps*nsub times (the opening drift plus every substep but
  // the last), so an even count returns the state to the buffer it started from and
  // the same graph can be replayed back to back.
  ARG_CHECK((steps * nsub) % 2 == 0, "graph mode needs steps * substeps even (got %d)", steps * nsub);
  double wsum = 0;
  for (int k = 0; k < nsub; ++k) wsum += w_host[k] * (k + 1);
  long long key[4] = {steps, (long long)(dt * 1e30), (long long)(wsum * 1e12) * 16 + nsub, record_every};
  const bool same = m.gexec && memcmp(key, m.graph_key, sizeof key) == 0 && m.graph_pos == m.pos;
  if (!same) {
    if (m.gexec) HIP_TRY(hipGraphExecDestroy(m.gexec));
    if (m.graph) HIP_TRY(hipGraphDestroy(m.graph));
    m.gexec = nullptr;
    m.graph = nullptr;
    double4* p0 = m.pos;
    HIP_TRY(hipStreamBeginCapture(m.st, hipStreamCaptureModeThreadLocal));
    m.enqueue_steps(steps, dt, w_host, nsub, record_every, 0);
    hipError_t e = hipStreamEndCapture(m.st, &m.graph);
    if (e != hipSuccess) {
      set_error("graph capture failed: %s", hipGetErrorString(e));
      return ACCU_ERUNTIME;
    }
    HIP_TRY(hipGraphInstantiate(&m.gexec, m.graph, nullptr, nullptr, 0));
    if (m.pos != p0) {
      set_error("internal: graph block did not return the state to its buffer");
      return ACCU_ERUNTIME;
    }
    m.graph_pos = p0;
    memcpy(m.graph_key, key, sizeof key);
  }
  HIP_TRY(hipGraphLaunch(m.gexec, m.st));
  return ACCU_OK;
}

int accu_md_sync(accu_md* h) {
  ARG_CHECK(h, "null handle");
  MD& m = *reinterpret_cast<MD*>(h);
  HIP_TRY(hipStreamSynchronize(m.st));
  return ACCU_OK;
}

// Copy the state back in user order (host [n][3] arrays, any may be NULL) and the
// log records. stats: [0] rebuilds, [1] records, [2] max neighbours seen, [3] capacity.
int accu_md_get(accu_md* h, double* pos_host, double* vel_host, double* F_host, double* log_host, int log_max,
                long long* stats) {
  ARG_CHECK(h, "null handle");
  MD& m = *reinterpret_cast<MD*>(h);
  const int n = m.p.n;
  double* tmp = nullptr;
  HIP_TRY(hipMallocAsync((void**)&tmp, (size_t)n * 3 * sizeof(double), m.st));
  double4* srcs[3] = {m.pos, m.vel, m.F};
  double* dsts[3] = {pos_host, vel_host, F_host};
  for (int k = 0; k < 3; ++k) {
    if (!dsts[k]) continue;
    hipLaunchKernelGGL(k_md_scatter_user, dim3((n + 255) / 256), dim3(256), 0, m.st, n, m.perm, srcs[k], tmp);
    HIP_TRY(hipMemcpyAsync(dsts[k], tmp, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, m.st));
  }
  int ints[3] = {0, 0, 0};
  HIP_TRY(hipMemcpyAsync(ints, m.ints, sizeof ints, hipMemcpyDeviceToHost, m.st));
  int mx = 0;
  HIP_TRY(hipMemcpyAsync(&mx, m.nl.dev_scalars, sizeof(int), hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipStreamSynchronize(m.st));
  HIP_TRY(hipFreeAsync(tmp, m.st));
  const int nrec = ints[2] < log_max ? ints[2] : log_max;
  if (log_host && nrec > 0)
    HIP_TRY(hipMemcpy(log_host, m.log, (size_t)nrec * kLogW * sizeof(double), hipMemcpyDeviceToHost));
  if (stats) {
    stats[0] = ints[1];
    stats[1] = ints[2];
    stats[2] = mx;
    stats[3] = m.nl.max_nbrs;
    // [4] = current number of list entries (both directions)
    std::vector<int> nn(n);
    HIP_TRY(hipMemcpy(nn.data(), m.nl.nn, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    long long tot = 0;
    for (int v : nn) tot += v;
    stats[4] = tot;
  }
  if (mx > m.nl.max_nbrs) {
    set_error("neighbour list overflow: %d neighbours > capacity %d (raise the slack)", mx, m.nl.max_nbrs);
    return ACCU_EOVERFLOW;
  }
  return ACCU_OK;
}

// Raw device pointers of the slot-ordered state (for timing / profiling harnesses).
int accu_md_stream(accu_md* h, void** stream_out) {
  ARG_CHECK(h && stream_out, "null argument");
  *stream_out = reinterpret_cast<MD*>(h)->st;
  return ACCU_OK;
}

}  // extern "C"

extern "C" {

// Enable / disable HIP-event timing of each fused launch (host-driven runs only).
int accu_md_set_timing(accu_md* h, int enable) {
  ARG_CHECK(h, "null handle");
  MD& m = *reinterpret_cast<MD*>(h);
  m.timing = enable != 0;
  m.ev_used = 0;
  return ACCU_OK;
}

// Synchronise and return the summed duration (ms) and count of timed fused launches
// since the last accu_md_set_timing; resets the count.
int accu_md_timing(accu_md* h, double* total_ms, long long* launches) {
  ARG_CHECK(h && total_ms && launches, "null argument");
  MD& m = *reinterpret_cast<MD*>(h);
  HIP_TRY(hipStreamSynchronize(m.st));
  double tot = 0.0;
  for (size_t k = 0; k + 1 < m.ev_used; k += 2) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, m.ev[k], m.ev[k + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *launches = (long long)(m.ev_used / 2);
  m.ev_used = 0;
  return ACCU_OK;
}

}  // extern "C"
