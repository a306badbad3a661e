// Pair forces over the Verlet list (N >= pairlist_threshold) or all pairs (small N):
// Lennard-Jones, damped shifted force (DSF) Coulomb, Ewald real space.
// List kernels: a quad of lanes per atom (its sorted slot), lane s of the quad taking list
// entries s, s + 4, ...; the full list means each atom's force is summed by its own quad in
// float64 (the quad's sums combined in a fixed order) and never scattered, so results do
// not depend on scheduling. Four lanes per atom give four times the waves of one thread
// per atom, which the latency of the neighbour gathers needs. Dense kernels: one thread
// per atom.
#include "common.h"
#include "nlist.h"
#include "pairkern.h"
#include "../../include/accusim_hip.h"

#include <cmath>
#include <vector>

namespace accu {

constexpr int kT = 256;

// ---------------------------------------------------------------- Lennard-Jones

template <typename P>
__global__ void __launch_bounds__(kT) k_lj_list(LJParams prm, const double* __restrict__ pos, int ps,
                                                const int* __restrict__ sp, int n, const int* __restrict__ sorted,
                                                const int* __restrict__ nn, const int* __restrict__ nbr, int stride,
                                                Box box, double* __restrict__ F, double* partials,
                                                unsigned* counter, double* Uout) {
  __shared__ double lds[3 * kMaxSpecies * kMaxSpecies];
  __shared__ double red[16];
  LJ<P> it;
  it.sp = sp;
  it.stage(prm, lds);
  __syncthreads();
  const int ns = prm.ns;
  const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int t = g >> 2, sub = g & 3;
  const bool on = t < n;
  double fx = 0, fy = 0, fz = 0, u = 0;
  int i = 0;
  if (on) {
    i = sorted ? sorted[t] : t;
    const double xi = pos[(size_t)ps * i], yi = pos[(size_t)ps * i + 1], zi = pos[(size_t)ps * i + 2];
    const int row = sp[i] * ns;
    const int cnt = nn[t];
    for (int k = sub; k < cnt; k += 4) {
      const int j = nbr[(size_t)k * stride + t];
      double dx = pos[(size_t)ps * j] - xi, dy = pos[(size_t)ps * j + 1] - yi, dz = pos[(size_t)ps * j + 2] - zi;
      if (box.periodic) {
        dx = minimg_fast(dx, box.L[0], box.inv[0]);
        dy = minimg_fast(dy, box.L[1], box.inv[1]);
        dz = minimg_fast(dz, box.L[2], box.inv[2]);
      }
      const int pj = row + sp[j];
      if constexpr (sizeof(P) == 8) {
        lj_pair64(dx, dy, dz, r2_np(dx, dy, dz), it.eps[pj], it.sig2[pj], it.rc2[pj], it.shift, fx, fy, fz, u);
      } else {  // float32 pair math on the float64 minimum-image displacement; float64 sums
        const float fdx = (float)dx, fdy = (float)dy, fdz = (float)dz;
        float gx = 0.f, gy = 0.f, gz = 0.f, gu = 0.f;
        lj_pair32(fdx, fdy, fdz, r2_npf(fdx, fdy, fdz), it.eps[pj], it.sig2[pj], it.rc2[pj], it.shift, gx, gy, gz,
                  gu);
        fx += (double)gx;
        fy += (double)gy;
        fz += (double)gz;
        u += (double)gu;
      }
    }
  }
  fx = quad_sum(fx);
  fy = quad_sum(fy);
  fz = quad_sum(fz);
  u = quad_sum(u);
  double U[1] = {0.0};
  if (on && sub == 0) {
    F[3 * (size_t)i] = fx;
    F[3 * (size_t)i + 1] = fy;
    F[3 * (size_t)i + 2] = fz;
    U[0] = u;
  }
  if (Uout) grid_sum<1>(U, partials, counter, Uout, red);
}

// All pairs (forces.py:157-175); thread i loops over every j != i.
__global__ void __launch_bounds__(kT) k_lj_dense(LJParams prm, const double* __restrict__ pos,
                                                 const int* __restrict__ sp, int n, Box box, double* __restrict__ F,
                                                 double* partials, unsigned* counter, double* Uout) {
  __shared__ double lds[3 * kMaxSpecies * kMaxSpecies];
  __shared__ double red[16];
  LJ<double> it;
  it.sp = sp;
  it.stage(prm, lds);
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double U[1] = {0.0};
  if (i < n) {
    const double xi = pos[3 * (size_t)i], yi = pos[3 * (size_t)i + 1], zi = pos[3 * (size_t)i + 2];
    const int row = sp[i] * prm.ns;
    double fx = 0, fy = 0, fz = 0, u = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double dx = pos[3 * (size_t)j] - xi, dy = pos[3 * (size_t)j + 1] - yi, dz = pos[3 * (size_t)j + 2] - zi;
      if (box.periodic) {
        dx = minimg_exact(dx, box.L[0]);
        dy = minimg_exact(dy, box.L[1]);
        dz = minimg_exact(dz, box.L[2]);
      }
      const int pj = row + sp[j];
      lj_pair64(dx, dy, dz, r2_np(dx, dy, dz), it.eps[pj], it.sig2[pj], it.rc2[pj], it.shift, fx, fy, fz, u);
    }
    F[3 * (size_t)i] = fx;
    F[3 * (size_t)i + 1] = fy;
    F[3 * (size_t)i + 2] = fz;
    U[0] = u;
  }
  if (Uout) grid_sum<1>(U, partials, counter, Uout, red);
}

// ---------------------------------------------------------------- Coulomb (DSF / Ewald real)

template <typename P>
__global__ void __launch_bounds__(kT) k_coul_list(CoulParams p, const double* __restrict__ pos, int ps,
                                                  const double* __restrict__ q, int n, const int* __restrict__ sorted,
                                                  const int* __restrict__ nn, const int* __restrict__ nbr, int stride,
                                                  Box box, double* __restrict__ F, double* partials,
                                                  unsigned* counter, double* Uout) {
  __shared__ double red[16];
  const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int t = g >> 2, sub = g & 3;
  const bool on = t < n;
  double fx = 0, fy = 0, fz = 0, u = 0, qi = 0;
  int i = 0;
  if (on) {
    i = sorted ? sorted[t] : t;
    const double xi = pos[(size_t)ps * i], yi = pos[(size_t)ps * i + 1], zi = pos[(size_t)ps * i + 2];
    qi = q[i];
    const int cnt = nn[t];
    for (int k = sub; k < cnt; k += 4) {
      const int j = nbr[(size_t)k * stride + t];
      if (p.nex && excluded(p, i, j, n)) continue;
      double dx = pos[(size_t)ps * j] - xi, dy = pos[(size_t)ps * j + 1] - yi, dz = pos[(size_t)ps * j + 2] - zi;
      if (box.periodic) {
        dx = minimg_fast(dx, box.L[0], box.inv[0]);
        dy = minimg_fast(dy, box.L[1], box.inv[1]);
        dz = minimg_fast(dz, box.L[2], box.inv[2]);
      }
      const double qq = qi * q[j];
      if constexpr (sizeof(P) == 8) {
        coul_pair64(p, dx, dy, dz, r2_np(dx, dy, dz), qq, fx, fy, fz, u);
      } else {
        const float fdx = (float)dx, fdy = (float)dy, fdz = (float)dz;
        float gx = 0.f, gy = 0.f, gz = 0.f, gu = 0.f;
        coul_pair32((float)p.alpha, (float)p.rc, (float)p.c, (float)p.erc_rc, (float)p.fshift, p.shifted, fdx, fdy,
                    fdz, r2_npf(fdx, fdy, fdz), (float)qq, gx, gy, gz, gu);
        fx += (double)gx;
        fy += (double)gy;
        fz += (double)gz;
        u += (double)gu;
      }
    }
  }
  fx = quad_sum(fx);
  fy = quad_sum(fy);
  fz = quad_sum(fz);
  u = quad_sum(u);
  double U[1] = {0.0};
  if (on && sub == 0) {
    u -= qi * qi * p.self;
    F[3 * (size_t)i] = p.k * fx;
    F[3 * (size_t)i + 1] = p.k * fy;
    F[3 * (size_t)i + 2] = p.k * fz;
    U[0] = p.k * u;
  }
  if (Uout) grid_sum<1>(U, partials, counter, Uout, red);
}

__global__ void __launch_bounds__(kT) k_coul_dense(CoulParams p, const double* __restrict__ pos,
                                                   const double* __restrict__ q, int n, Box box,
                                                   double* __restrict__ F, double* partials, unsigned* counter,
                                                   double* Uout) {
  __shared__ double red[16];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double U[1] = {0.0};
  if (i < n) {
    const double xi = pos[3 * (size_t)i], yi = pos[3 * (size_t)i + 1], zi = pos[3 * (size_t)i + 2];
    const double qi = q[i];
    double fx = 0, fy = 0, fz = 0, u = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      if (p.nex && excluded(p, i, j, n)) continue;
      double dx = pos[3 * (size_t)j] - xi, dy = pos[3 * (size_t)j + 1] - yi, dz = pos[3 * (size_t)j + 2] - zi;
      if (box.periodic) {
        dx = minimg_exact(dx, box.L[0]);
        dy = minimg_exact(dy, box.L[1]);
        dz = minimg_exact(dz, box.L[2]);
      }
      coul_pair64(p, dx, dy, dz, r2_np(dx, dy, dz), qi * q[j], fx, fy, fz, u);
    }
    u -= qi * qi * p.self;
    F[3 * (size_t)i] = p.k * fx;
    F[3 * (size_t)i + 1] = p.k * fy;
    F[3 * (size_t)i + 2] = p.k * fz;
    U[0] = p.k * u;
  }
  if (Uout) grid_sum<1>(U, partials, counter, Uout, red);
}

Box make_box(const double* box, int periodic) {
  Box b{};
  b.periodic = periodic && box;
  for (int d = 0; d < 3; ++d) {
    b.L[d] = box ? box[d] : 1.0;
    b.inv[d] = 1.0 / b.L[d];
  }
  return b;
}

}  // namespace accu

using namespace accu;

static int fill_lj(LJParams& p, int ns, const double* eps, const double* sig2, const double* rc2, double shift) {
  ARG_CHECK(ns >= 1 && ns <= kMaxSpecies, "number of species %d outside [1, %d]", ns, kMaxSpecies);
  p.ns = ns;
  p.shift = shift;
  for (int k = 0; k < ns * ns; ++k) {
    p.eps[k] = eps[k];
    p.sig2[k] = sig2[k];
    p.rc2[k] = rc2[k];
  }
  return ACCU_OK;
}

extern "C" {

int accu_lj_forces(accu_nlist* hh, const double* pos, const int* sp, int n, int ns, const double* eps,
                   const double* sig2, const double* rc2, double shift, const double* box, int periodic,
                   int precision, double* F, double* U, void* stream) {
  ARG_CHECK(pos && sp && F && eps && sig2 && rc2, "null argument");
  ARG_CHECK(precision == ACCU_FP64 || precision == ACCU_FP32_PAIR, "unknown precision %d", precision);
  LJParams prm;
  int rc = fill_lj(prm, ns, eps, sig2, rc2, shift);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (n <= 0) return ACCU_OK;
  Workspace* ws = workspace_for(st);
  if (!ws) return ACCU_ERUNTIME;
  const int nb = (n + kT - 1) / kT, nbl = (int)((4LL * n + kT - 1) / kT);  // dense / list (a quad per atom)
  if ((rc = ws->ensure((size_t)nbl * 2))) return rc;
  Box b = make_box(box, periodic);
  if (hh) {
    NList& h = *reinterpret_cast<NList*>(hh);
    ARG_CHECK(h.n == n, "neighbour list was built for %d atoms, got %d", h.n, n);
    if (precision == ACCU_FP64)
      hipLaunchKernelGGL(k_lj_list<double>, dim3(nbl), dim3(kT), 0, st, prm, pos, 3, sp, n, h.sorted, h.nn, h.nbr,
                         h.stride, b, F, ws->partials, ws->counter, U);
    else
      hipLaunchKernelGGL(k_lj_list<float>, dim3(nbl), dim3(kT), 0, st, prm, pos, 3, sp, n, h.sorted, h.nn, h.nbr,
                         h.stride, b, F, ws->partials, ws->counter, U);
  } else {
    hipLaunchKernelGGL(k_lj_dense, dim3(nb), dim3(kT), 0, st, prm, pos, sp, n, b, F, ws->partials, ws->counter, U);
  }
  LAUNCH_CHECK();
  return ACCU_OK;
}

int accu_coulomb_forces(accu_nlist* hh, const double* pos, const double* q, int n, int shifted, double alpha,
                        double rcut, double erc, double k, const long long* ex_keys, int nex, const double* box,
                        int periodic, int precision, double* F, double* U, void* stream) {
  ARG_CHECK(pos && q && F, "null argument");
  ARG_CHECK(precision == ACCU_FP64 || precision == ACCU_FP32_PAIR, "unknown precision %d", precision);
  ARG_CHECK(nex == 0 || ex_keys, "exclusion keys missing");
  hipStream_t st = (hipStream_t)stream;
  if (n <= 0) return ACCU_OK;
  CoulParams p{};
  p.alpha = alpha;
  p.rc = rcut;
  p.c = 2 * alpha / std::sqrt(M_PI);
  p.erc_rc = shifted ? erc / rcut : 0.0;
  // forces.py:416 fshift = erc / rc^2 + c exp(-(a rc)^2) / rc
  p.fshift = shifted ? erc / (rcut * rcut) + p.c * std::exp(-(alpha * rcut) * (alpha * rcut)) / rcut : 0.0;
  // forces.py:434 and 387
  p.self = shifted ? (erc / (2 * rcut) + alpha / std::sqrt(M_PI)) : alpha / std::sqrt(M_PI);
  p.k = k;
  p.shifted = shifted;
  p.nex = nex;
  p.ex = ex_keys;
  Workspace* ws = workspace_for(st);
  if (!ws) return ACCU_ERUNTIME;
  const int nb = (n + kT - 1) / kT, nbl = (int)((4LL * n + kT - 1) / kT);  // dense / list (a quad per atom)
  int rc = ws->ensure((size_t)nbl * 2);
  if (rc) return rc;
  Box b = make_box(box, periodic);
  if (hh) {
    NList& h = *reinterpret_cast<NList*>(hh);
    ARG_CHECK(h.n == n, "neighbour list was built for %d atoms, got %d", h.n, n);
    if (precision == ACCU_FP64)
      hipLaunchKernelGGL(k_coul_list<double>, dim3(nbl), dim3(kT), 0, st, p, pos, 3, q, n, h.sorted, h.nn, h.nbr,
                         h.stride, b, F, ws->partials, ws->counter, U);
    else
      hipLaunchKernelGGL(k_coul_list<float>, dim3(nbl), dim3(kT), 0, st, p, pos, 3, q, n, h.sorted, h.nn, h.nbr,
                         h.stride, b, F, ws->partials, ws->counter, U);
  } else {
    hipLaunchKernelGGL(k_coul_dense, dim3(nb), dim3(kT), 0, st, p, pos, q, n, b, F, ws->partials, ws->counter, U);
  }
  LAUNCH_CHECK();
  return ACCU_OK;
}

}  // extern "C"
