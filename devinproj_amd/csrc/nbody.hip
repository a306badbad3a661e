// Device-resident N-body runs: Simulation.run (simulation.py:95-112) with VelocityVerlet or
// Yoshida4 (integrators.py:19-47) over direct Gravity (forces.py:31-48) or BarnesHut
// (tree.py:19-78), with the state in HBM between calls and the record sums of
// simulation.py:76-93 (potential and kinetic energy, momentum, angular momentum) reduced in
// float64 on the device.
//
// Direct gravity: ONE launch per (sub)step. Each block sums the forces on its 256 targets
// (j-particles through LDS tiles, as accu_gravity_direct), then does the closing half kick,
// the record terms of a recorded step, and the next substep's opening kick and drift into
// the other position buffer. Positions are double-buffered (a parity word on the device
// says which buffer is current), so no block's drift races another block's reads, and a
// Yoshida4 step is three launches. A run of steps can be captured as one hipGraph.
// Barnes-Hut: per substep the device tree build (one host synchronisation for the node
// count, so these runs are not captured), the walk into F, and one per-particle launch for
// the kicks, record terms and drift.
//
// The step is the reference's, per substep of weight w (dt_k = w dt, h_k = 0.5 dt_k):
//   v += h_k a; x += dt_k v; x %= box (periodic); a = F / m; v += h_k a.
// Fused across substeps: the launch that computes a closes substep k and opens k + 1.
#include "common.h"
#include "gravity.h"
#include "../../include/accusim_hip.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace accu {

Box make_box(const double* box, int periodic);

namespace {

constexpr int kNT = 256;  // threads per block = LDS tile of j-particles
constexpr int kNbW = 8;   // record row and block partials: PE, KE, Px, Py, Pz, Lx, Ly, Lz

struct NbStep {
  double h_cur;    // closing half kick of the substep that ends here
  double h_next;   // opening half kick of the next substep
  double dt_next;  // drift of the next substep
  int drift;       // 0: last substep of a run: no drift, the acceleration is kept
};

// Closing kick of particle i, its record terms, then the next opening kick and drift (into
// the other buffer) or, at the end of a run, the stored acceleration.
__device__ __forceinline__ void nb_finish(int i, const double4& x, double ax, double ay, double az, double (&rec)[kNbW],
                                          bool en, const NbStep& sp, const Box& box, double4* __restrict__ vel,
                                          double4* __restrict__ acc, double4* __restrict__ next) {
  double4 v = vel[i];
  {
#pragma clang fp contract(off)
    v.x = v.x + sp.h_cur * ax;
    v.y = v.y + sp.h_cur * ay;
    v.z = v.z + sp.h_cur * az;
  }
  if (en) {  // system.py:106-114: KE, P = sum m v, L = sum x cross (m v)
    const double m = x.w, px = m * v.x, py = m * v.y, pz = m * v.z;
    rec[1] += 0.5 * m * ((v.x * v.x + v.z * v.z) + v.y * v.y);
    rec[2] += px;
    rec[3] += py;
    rec[4] += pz;
    rec[5] += x.y * pz - x.z * py;
    rec[6] += x.z * px - x.x * pz;
    rec[7] += x.x * py - x.y * px;
  }
  if (sp.drift) {
    double4 y = x;
    {
#pragma clang fp contract(off)
      v.x = v.x + sp.h_next * ax;
      v.y = v.y + sp.h_next * ay;
      v.z = v.z + sp.h_next * az;
      y.x = y.x + sp.dt_next * v.x;
      y.y = y.y + sp.dt_next * v.y;
      y.z = y.z + sp.dt_next * v.z;
    }
    if (box.periodic) {
      y.x = np_mod2(y.x, box.L[0]);
      y.y = np_mod2(y.y, box.L[1]);
      y.z = np_mod2(y.z, box.L[2]);
    }
    next[i] = y;
  } else {
    acc[i] = make_double4(ax, ay, az, 0.0);
  }
  vel[i] = v;
}

template <int NV>
__device__ __forceinline__ void store_partials(double (&v)[NV], double* lds, double* __restrict__ partials) {
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)blockIdx.x * NV + k] = v[k];
}

// Direct gravity step. F32: float32 pair terms relative to `origin` (the initial centre of
// mass), tile sums in float64; else float64 throughout (periodic allowed).
template <bool F32, bool EN>
__global__ void __launch_bounds__(kNT) k_nb_direct(int n, double4* __restrict__ posA, double4* __restrict__ posB,
                                                   const int* __restrict__ par, double4* __restrict__ vel,
                                                   double4* __restrict__ acc, double G, double eps2, Box box,
                                                   double ox, double oy, double oz, NbStep sp,
                                                   double* __restrict__ partials) {
  __shared__ double4 tile64[F32 ? 1 : kNT];
  __shared__ float4 tile32[F32 ? kNT : 1];
  __shared__ double red[kNbW * (kNT / 64)];
  const double4* __restrict__ cur = par[0] ? posB : posA;
  double4* __restrict__ next = par[0] ? posA : posB;
  const int i = xcd_block(blockIdx.x, gridDim.x) * kNT + threadIdx.x;
  const bool active = i < n;
  const double4 xi = cur[active ? i : 0];
  const float xf = (float)(xi.x - ox), yf = (float)(xi.y - oy), zf = (float)(xi.z - oz), e2 = (float)eps2;
  double fx = 0, fy = 0, fz = 0, phi = 0;
  for (int base = 0; base < n; base += kNT) {
    const int j = base + threadIdx.x;
    if (j < n) {
      const double4 pj = cur[j];
      if (F32) tile32[threadIdx.x] = make_float4((float)(pj.x - ox), (float)(pj.y - oy), (float)(pj.z - oz), (float)pj.w);
      else tile64[threadIdx.x] = pj;
    }
    __syncthreads();
    const int cnt = (n - base) < kNT ? (n - base) : kNT;
    const bool self = __any(i >= base && i < base + cnt);
    if (F32) {
      float gx = 0.f, gy = 0.f, gz = 0.f, gp = 0.f;
      if (self) grav32_tile<true>(tile32, cnt, base, i, xf, yf, zf, e2, gx, gy, gz, gp);
      else grav32_tile<false>(tile32, cnt, base, i, xf, yf, zf, e2, gx, gy, gz, gp);
      fx += (double)gx;
      fy += (double)gy;
      fz += (double)gz;
      phi += (double)gp;
    } else {
      if (self) grav64_tile<true>(tile64, cnt, base, i, xi.x, xi.y, xi.z, eps2, box, fx, fy, fz, phi);
      else grav64_tile<false>(tile64, cnt, base, i, xi.x, xi.y, xi.z, eps2, box, fx, fy, fz, phi);
    }
    __syncthreads();
  }
  double rec[kNbW] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (active) {
    const double gm = G * xi.w;
    // forces.py:44-47: F_i = G m_i sum_j m_j d / r^3; simulation.py:68: a = F / m
    const double ax = gm * fx / xi.w, ay = gm * fy / xi.w, az = gm * fz / xi.w;
    if (EN) rec[0] = -0.5 * gm * phi;
    nb_finish(i, xi, ax, ay, az, rec, EN, sp, box, vel, acc, next);
  }
  if (EN) store_partials<kNbW>(rec, red, partials);
}

// Barnes-Hut: positions and masses in the tree build's layout
__global__ void k_nb_split(int n, const double4* __restrict__ posA, const double4* __restrict__ posB,
                           const int* __restrict__ par, double* __restrict__ pos3, double* __restrict__ m) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= n) return;
  const double4 x = (par[0] ? posB : posA)[i];
  pos3[3 * (size_t)i] = x.x;
  pos3[3 * (size_t)i + 1] = x.y;
  pos3[3 * (size_t)i + 2] = x.z;
  m[i] = x.w;
}

// Barnes-Hut: kicks, record terms and drift from the walk's forces F [n][3]
template <bool EN>
__global__ void __launch_bounds__(kNT) k_nb_kick(int n, double4* __restrict__ posA, double4* __restrict__ posB,
                                                 const int* __restrict__ par, double4* __restrict__ vel,
                                                 double4* __restrict__ acc, const double* __restrict__ F, Box box,
                                                 NbStep sp, double* __restrict__ partials) {
  __shared__ double red[kNbW * (kNT / 64)];
  const double4* __restrict__ cur = par[0] ? posB : posA;
  double4* __restrict__ next = par[0] ? posA : posB;
  const int i = blockIdx.x * kNT + threadIdx.x;
  double rec[kNbW] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (i < n) {
    const double4 x = cur[i];
    const double ax = F[3 * (size_t)i] / x.w, ay = F[3 * (size_t)i + 1] / x.w, az = F[3 * (size_t)i + 2] / x.w;
    nb_finish(i, x, ax, ay, az, rec, EN, sp, box, vel, acc, next);
  }
  if (EN) store_partials<kNbW>(rec, red, partials);
}

// First opening kick and drift of a run, from the stored acceleration
__global__ void k_nb_drift(int n, double4* __restrict__ posA, double4* __restrict__ posB, const int* __restrict__ par,
                           double4* __restrict__ vel, const double4* __restrict__ acc, double h, double dt, Box box) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= n) return;
  const double4* __restrict__ cur = par[0] ? posB : posA;
  double4* __restrict__ next = par[0] ? posA : posB;
  const double4 a = acc[i];
  double4 v = vel[i], y = cur[i];
  {
#pragma clang fp contract(off)
    v.x = v.x + h * a.x;
    v.y = v.y + h * a.y;
    v.z = v.z + h * a.z;
    y.x = y.x + dt * v.x;
    y.y = y.y + dt * v.y;
    y.z = y.z + dt * v.z;
  }
  if (box.periodic) {
    y.x = np_mod2(y.x, box.L[0]);
    y.y = np_mod2(y.y, box.L[1]);
    y.z = np_mod2(y.z, box.L[2]);
  }
  next[i] = y;
  vel[i] = v;
}

__global__ void k_nb_flip(int* par) {
  if (threadIdx.x == 0) par[0] ^= 1;
}

// One record row: the block partials summed in block order (a fixed order: the same
// bits on every run and graph replay); Uext (Barnes-Hut) replaces the potential term.
__global__ void __launch_bounds__(kNT) k_nb_record(const double* __restrict__ partials, int nb,
                                                   const double* __restrict__ Uext, double* __restrict__ log,
                                                   int log_cap, int* __restrict__ log_count) {
  __shared__ double red[kNbW * (kNT / 64)];
  double v[kNbW] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += kNT)
#pragma unroll
    for (int k = 0; k < kNbW; ++k) v[k] += partials[(size_t)b * kNbW + k];
  block_sum<kNbW>(v, red);
  if (threadIdx.x == 0) {
    const int idx = log_count[0];
    if (idx < log_cap) {
      double* row = log + (size_t)idx * kNbW;
      if (Uext) v[0] = Uext[0];
      for (int k = 0; k < kNbW; ++k) row[k] = v[k];
    }
    log_count[0] = idx + 1;
  }
}

}  // namespace
}  // namespace accu

using namespace accu;

struct accu_nbody {
  int n = 0, nb = 0;
  hipStream_t st = nullptr;
  Box box = {};
  double origin[3] = {0, 0, 0};
  int mode = 0;  // 1: direct gravity, 2: Barnes-Hut
  int precision = ACCU_FP64;
  double G = 0, eps2 = 0, theta = 0;
  int leaf = 8, quad = 0;
  double4* pos[2] = {nullptr, nullptr};
  double4 *vel = nullptr, *acc = nullptr;
  int* par = nullptr;
  double* partials = nullptr;
  double* log = nullptr;
  int log_cap = 0;
  int* log_count = nullptr;
  // Barnes-Hut
  double *pos3 = nullptr, *m = nullptr, *F = nullptr, *Ubh = nullptr;
  accu_bh* bh = nullptr;
  // graph of the last captured run
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<double> graph_key;
  bool ready = false;

  void release() {
    void* ps[] = {pos[0], pos[1], vel, acc, par, partials, log, log_count, pos3, m, F, Ubh};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    if (bh) accu_bh_destroy(bh);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (st) (void)hipStreamDestroy(st);
  }

  // the force evaluation that closes one substep and opens the next
  int force_step(const NbStep& sp, bool rec) {
    const dim3 g(nb), b(kNT);
    if (mode == 1) {
      const bool f32 = precision == ACCU_FP32_PAIR;
      auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, g, b, 0, st, n, pos[0], pos[1], par, vel, acc, G, eps2, box, origin[0], origin[1],
                           origin[2], sp, partials);
      };
      if (f32) rec ? go(k_nb_direct<true, true>) : go(k_nb_direct<true, false>);
      else rec ? go(k_nb_direct<false, true>) : go(k_nb_direct<false, false>);
      LAUNCH_CHECK();
    } else {
      hipLaunchKernelGGL(k_nb_split, g, b, 0, st, n, pos[0], pos[1], par, pos3, m);
      LAUNCH_CHECK();
      int rc = accu_bh_build(bh, pos3, m, n, leaf, quad, st);
      if (rc) return rc;
      rc = accu_bh_walk(bh, G, eps2, theta, F, rec ? Ubh : nullptr, st);
      if (rc) return rc;
      if (rec)
        hipLaunchKernelGGL(k_nb_kick<true>, g, b, 0, st, n, pos[0], pos[1], par, vel, acc, F, box, sp, partials);
      else
        hipLaunchKernelGGL(k_nb_kick<false>, g, b, 0, st, n, pos[0], pos[1], par, vel, acc, F, box, sp, partials);
      LAUNCH_CHECK();
    }
    if (sp.drift) hipLaunchKernelGGL(k_nb_flip, dim3(1), dim3(64), 0, st, par);
    if (rec)
      hipLaunchKernelGGL(k_nb_record, dim3(1), dim3(kNT), 0, st, partials, nb, mode == 2 ? Ubh : nullptr, log, log_cap,
                         log_count);
    LAUNCH_CHECK();
    return ACCU_OK;
  }

  // `steps` steps of dt split into substeps of weights w[0..nsub)
  int enqueue(int steps, double dt, const double* w, int nsub, int record_every) {
    const double dt0 = w[0] * dt, h0 = 0.5 * dt0;  // integrators.py:25-26 with dt_k = w dt
    hipLaunchKernelGGL(k_nb_drift, dim3(nb), dim3(kNT), 0, st, n, pos[0], pos[1], par, vel, acc, h0, dt0, box);
    hipLaunchKernelGGL(k_nb_flip, dim3(1), dim3(64), 0, st, par);
    LAUNCH_CHECK();
    for (int s = 0; s < steps; ++s)
      for (int k = 0; k < nsub; ++k) {
        const bool last = s == steps - 1 && k == nsub - 1;
        const int kn = (k + 1) % nsub;
        const bool rec = record_every > 0 && k == nsub - 1 && ((s + 1) % record_every == 0 || s == steps - 1);
        const double dtk = w[k] * dt, dtn = w[kn] * dt;
        NbStep sp = {0.5 * dtk, 0.5 * dtn, dtn, last ? 0 : 1};
        int rc = force_step(sp, rec);
        if (rc) return rc;
      }
    return ACCU_OK;
  }
};

extern "C" {

int accu_nbody_create(accu_nbody** out, int n, const double* pos_host, const double* vel_host, const double* mass_host,
                      const double* box_host, int periodic) {
  ARG_CHECK(out && pos_host && vel_host && mass_host, "null argument");
  ARG_CHECK(n >= 2, "an N-body run needs at least two particles");
  ARG_CHECK(!box_host || periodic, "walls (a box that is not periodic) are not on this path");
  accu_nbody* h = new accu_nbody();
  h->n = n;
  h->nb = (n + kNT - 1) / kNT;
  h->box = make_box(box_host, periodic);
  std::vector<double4> x(n), v(n);
  double M = 0, c[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i) {
    x[i] = make_double4(pos_host[3 * (size_t)i], pos_host[3 * (size_t)i + 1], pos_host[3 * (size_t)i + 2], mass_host[i]);
    v[i] = make_double4(vel_host[3 * (size_t)i], vel_host[3 * (size_t)i + 1], vel_host[3 * (size_t)i + 2], 0.0);
    M += mass_host[i];
    for (int d = 0; d < 3; ++d) c[d] += mass_host[i] * pos_host[3 * (size_t)i + d];
  }
  for (int d = 0; d < 3; ++d) h->origin[d] = M > 0 ? c[d] / M : 0.0;
  const size_t d4 = (size_t)n * sizeof(double4);
  int rc = ACCU_OK;
  auto alloc = [&](void** p, size_t bytes) {
    if (rc == ACCU_OK && hipMalloc(p, bytes) != hipSuccess) {
      set_error("accu_nbody_create: device allocation of %zu bytes failed", bytes);
      rc = ACCU_ERUNTIME;
    }
  };
  alloc((void**)&h->pos[0], d4);
  alloc((void**)&h->pos[1], d4);
  alloc((void**)&h->vel, d4);
  alloc((void**)&h->acc, d4);
  alloc((void**)&h->par, 64);
  alloc((void**)&h->partials, (size_t)h->nb * kNbW * sizeof(double));
  alloc((void**)&h->log_count, 64);
  if (rc == ACCU_OK &&
      (hipMemcpy(h->pos[0], x.data(), d4, hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(h->vel, v.data(), d4, hipMemcpyHostToDevice) != hipSuccess ||
       hipMemset(h->acc, 0, d4) != hipSuccess || hipMemset(h->par, 0, 64) != hipSuccess ||
       hipMemset(h->log_count, 0, 64) != hipSuccess ||
       hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess ||
       hipDeviceSynchronize() != hipSuccess)) {  // null-stream fills before the engine stream runs
    set_error("accu_nbody_create: upload failed");
    rc = ACCU_ERUNTIME;
  }
  if (rc) {
    h->release();
    delete h;
    return rc;
  }
  *out = h;
  return ACCU_OK;
}

int accu_nbody_destroy(accu_nbody* h) {
  if (h) {
    h->release();
    delete h;
  }
  return ACCU_OK;
}

int accu_nbody_set_direct(accu_nbody* h, double G, double eps2, int precision) {
  ARG_CHECK(h && !h->ready, "set the force before accu_nbody_init");
  ARG_CHECK(precision == ACCU_FP64 || precision == ACCU_FP32_PAIR, "unknown precision %d", precision);
  ARG_CHECK(!(precision == ACCU_FP32_PAIR && h->box.periodic), "float32 gravity supports open boundaries only");
  h->mode = 1;
  h->precision = precision;
  h->G = G;
  h->eps2 = eps2;
  return ACCU_OK;
}

int accu_nbody_set_bh(accu_nbody* h, double G, double eps2, double theta, int leaf_size, int quadrupole) {
  ARG_CHECK(h && !h->ready, "set the force before accu_nbody_init");
  ARG_CHECK(!h->box.periodic, "Barnes-Hut is open-space gravity (tree.py has no periodic images)");
  ARG_CHECK(leaf_size >= 1, "leaf_size must be >= 1");
  h->mode = 2;
  h->G = G;
  h->eps2 = eps2;
  h->theta = theta;
  h->leaf = leaf_size;
  h->quad = quadrupole != 0;
  const size_t n = h->n;
  HIP_TRY(hipMalloc(&h->pos3, 3 * n * sizeof(double)));
  HIP_TRY(hipMalloc(&h->m, n * sizeof(double)));
  HIP_TRY(hipMalloc(&h->F, 3 * n * sizeof(double)));
  HIP_TRY(hipMalloc(&h->Ubh, 64));
  return accu_bh_create(&h->bh);
}

int accu_nbody_init(accu_nbody* h, int log_cap) {
  ARG_CHECK(h && h->mode, "set a force first");
  ARG_CHECK(!h->ready, "accu_nbody_init was called already");
  h->log_cap = log_cap > 4 ? log_cap : 4;
  HIP_TRY(hipMalloc(&h->log, (size_t)h->log_cap * kNbW * sizeof(double)));
  // forces at t = 0 and the first record (simulation.py:53, 104)
  const NbStep sp = {0.0, 0.0, 0.0, 0};
  int rc = h->force_step(sp, true);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(h->st));
  h->ready = true;
  return ACCU_OK;
}

int accu_nbody_run(accu_nbody* h, int steps, double dt, const double* w_host, int nsub, int record_every,
                   int use_graph) {
  ARG_CHECK(h && w_host, "null argument");
  ARG_CHECK(h->ready, "call accu_nbody_init first");
  ARG_CHECK(steps >= 1 && nsub >= 1 && nsub <= 16, "bad step counts");
  ARG_CHECK(!(use_graph && h->mode == 2), "Barnes-Hut runs synchronise once per tree build and are not captured");
  int cnt = 0;
  HIP_TRY(hipMemcpyAsync(&cnt, h->log_count, sizeof(int), hipMemcpyDeviceToHost, h->st));
  HIP_TRY(hipStreamSynchronize(h->st));
  const int nrec = record_every > 0 ? steps / record_every + 2 : 0;
  ARG_CHECK(cnt + nrec <= h->log_cap, "the energy log is full (%d of %d records)", cnt, h->log_cap);
  if (!use_graph) {
    const int rc = h->enqueue(steps, dt, w_host, nsub, record_every);
    if (rc) return rc;
    LAUNCH_CHECK();
    return ACCU_OK;
  }
  std::vector<double> key = {(double)steps, dt, (double)nsub, (double)record_every};
  for (int k = 0; k < nsub; ++k) key.push_back(w_host[k]);
  if (!(h->gexec && key == h->graph_key)) {
    if (h->gexec) HIP_TRY(hipGraphExecDestroy(h->gexec));
    if (h->graph) HIP_TRY(hipGraphDestroy(h->graph));
    h->gexec = nullptr;
    h->graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
    const int rc = h->enqueue(steps, dt, w_host, nsub, record_every);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(h->st, &gr);
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) {
      set_error("graph capture failed: %s", hipGetErrorString(e));
      return ACCU_ERUNTIME;
    }
    h->graph = gr;
    HIP_TRY(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
    h->graph_key = key;
  }
  HIP_TRY(hipGraphLaunch(h->gexec, h->st));
  return ACCU_OK;
}

int accu_nbody_sync(accu_nbody* h) {
  ARG_CHECK(h, "null handle");
  HIP_TRY(hipStreamSynchronize(h->st));
  return ACCU_OK;
}

int accu_nbody_get(accu_nbody* h, double* pos_host, double* vel_host, double* acc_host, double* log_host, int log_max,
                   long long* nrec_out) {
  ARG_CHECK(h && h->ready, "call accu_nbody_init first");
  HIP_TRY(hipStreamSynchronize(h->st));
  int par = 0, cnt = 0;
  HIP_TRY(hipMemcpy(&par, h->par, sizeof(int), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(&cnt, h->log_count, sizeof(int), hipMemcpyDeviceToHost));
  const size_t n = h->n;
  std::vector<double4> t(n);
  auto unpack = [&](const double4* src, double* dst) -> int {
    HIP_TRY(hipMemcpy(t.data(), src, n * sizeof(double4), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
      dst[3 * i] = t[i].x;
      dst[3 * i + 1] = t[i].y;
      dst[3 * i + 2] = t[i].z;
    }
    return ACCU_OK;
  };
  int rc;
  if (pos_host && (rc = unpack(h->pos[par], pos_host))) return rc;
  if (vel_host && (rc = unpack(h->vel, vel_host))) return rc;
  if (acc_host && (rc = unpack(h->acc, acc_host))) return rc;
  const int nr = cnt < h->log_cap ? cnt : h->log_cap;
  if (log_host && log_max > 0) {
    const int k = nr < log_max ? nr : log_max;
    if (k) HIP_TRY(hipMemcpy(log_host, h->log, (size_t)k * kNbW * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (nrec_out) *nrec_out = nr;
  return ACCU_OK;
}

int accu_nbody_stream(accu_nbody* h, void** stream_out) {
  ARG_CHECK(h && stream_out, "null argument");
  *stream_out = h->st;
  return ACCU_OK;
}

}  // extern "C"
