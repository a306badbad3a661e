// Ewald reciprocal-space sum (forces.py:380-386):
//   S(k)  = sum_j q_j exp(i k.r_j)
//   U_rec = 2 pi / V sum_k w_k |S(k)|^2,       w_k = exp(-k^2 / 4 a^2) / k^2
//   F_i  += 4 pi / V q_i sum_k w_k Im(conj(S(k)) exp(i k.r_i)) k
// k = 2 pi n / L runs over one half-space of "columns" (nx, ny, nz_lo..nz_hi); the
// other half contributes the same amount (S(-k) = conj S(k)), hence the factors 2.
// Along a column exp(i k.r) advances by one complex multiply per nz step, so the
// only transcendental work is one sincos per (particle, column).
#include "common.h"
#include "../../include/accusim_hip.h"

#include <cmath>
#include <vector>

namespace accu {

constexpr int kET = 256;
constexpr int kEM = 8;        // particles per thread in the structure-factor kernel
constexpr int kMaxCol = 512;  // max nz values per column

struct EwCols {
  const int* col;       // [ncol][4] = nx, ny, nz_lo, nz_hi
  const long long* off; // [ncol] first k index of the column
  int ncol;
  double tw[3];         // 2 pi / L per axis
};

// Partial structure factors: block (cchunk, pblock) covers `cpb` columns and
// kET * kEM particles; writes part[pblock][k] (complex).
__global__ void __launch_bounds__(kET) k_sk_partial(const double* __restrict__ pos, const double* __restrict__ q,
                                                    int n, EwCols cols, int cpb, long long K,
                                                    double2* __restrict__ part) {
  __shared__ double2 acc_lds[kET / 64][kMaxCol];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int pblock = blockIdx.y;
  double x[kEM], y[kEM], z[kEM], qq[kEM], ezc[kEM], ezs[kEM];
#pragma unroll
  for (int m = 0; m < kEM; ++m) {
    const int j = (pblock * kEM + m) * kET + threadIdx.x;
    if (j < n) {
      x[m] = pos[3 * (size_t)j];
      y[m] = pos[3 * (size_t)j + 1];
      z[m] = pos[3 * (size_t)j + 2];
      qq[m] = q[j];
    } else {
      x[m] = y[m] = z[m] = 0.0;
      qq[m] = 0.0;
    }
    sincos(cols.tw[2] * z[m], &ezs[m], &ezc[m]);
  }
  const int c0 = blockIdx.x * cpb, c1 = min(c0 + cpb, cols.ncol);
  for (int c = c0; c < c1; ++c) {
    const int nx = cols.col[4 * c], ny = cols.col[4 * c + 1], lo = cols.col[4 * c + 2], hi = cols.col[4 * c + 3];
    const double kx = cols.tw[0] * nx, ky = cols.tw[1] * ny, kz = cols.tw[2] * lo;
    double er[kEM], ei[kEM];
#pragma unroll
    for (int m = 0; m < kEM; ++m) sincos(kx * x[m] + ky * y[m] + kz * z[m], &ei[m], &er[m]);
    const int len = hi - lo + 1;
    for (int s = 0; s < len; ++s) {
      double ar = 0.0, ai = 0.0;
#pragma unroll
      for (int m = 0; m < kEM; ++m) {
        ar = fma(qq[m], er[m], ar);
        ai = fma(qq[m], ei[m], ai);
        const double nr = fma(er[m], ezc[m], -ei[m] * ezs[m]);
        const double ni = fma(er[m], ezs[m], ei[m] * ezc[m]);
        er[m] = nr;
        ei[m] = ni;
      }
      ar = wave_sum(ar);
      ai = wave_sum(ai);
      if (lane == 0) acc_lds[w][s] = make_double2(ar, ai);
    }
    __syncthreads();
    const long long kb = cols.off[c];
    for (int s = threadIdx.x; s < len; s += kET) {
      double2 v = acc_lds[0][s];
#pragma unroll
      for (int q2 = 1; q2 < kET / 64; ++q2) {
        v.x += acc_lds[q2][s].x;
        v.y += acc_lds[q2][s].y;
      }
      part[(size_t)pblock * K + kb + s] = v;
    }
    __syncthreads();
  }
}

// S(k) = sum over particle blocks (fixed order); A(k) = w_k S(k); energy partial w|S|^2.
__global__ void __launch_bounds__(kET) k_sk_reduce(const double2* __restrict__ part, int npb, long long K,
                                                   const double* __restrict__ kw, double2* __restrict__ A,
                                                   double* partials, unsigned* counter, double* Uout) {
  __shared__ double red[16];
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double U[1] = {0.0};
  if (k < K) {
    double sr = 0.0, si = 0.0;
    for (int p = 0; p < npb; ++p) {
      const double2 v = part[(size_t)p * K + k];
      sr += v.x;
      si += v.y;
    }
    const double w = kw[k];
    A[k] = make_double2(w * sr, w * si);
    U[0] = w * (sr * sr + si * si);
  }
  grid_sum<1>(U, partials, counter, Uout, red);
}

// Per-column weights w_k for every k of the half space.
__global__ void k_kweights(EwCols cols, double inv4a2, double* __restrict__ kw) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols.ncol) return;
  const int nx = cols.col[4 * c], ny = cols.col[4 * c + 1], lo = cols.col[4 * c + 2], hi = cols.col[4 * c + 3];
  const double kx = cols.tw[0] * nx, ky = cols.tw[1] * ny;
  long long o = cols.off[c];
  for (int nz = lo; nz <= hi; ++nz, ++o) {
    const double kz = cols.tw[2] * nz;
    const double k2 = (kx * kx + ky * ky) + kz * kz;
    kw[o] = exp(-k2 * inv4a2) / k2;
  }
}

// F_i = scale q_i sum_k Im(conj(S_k) e^{i k.r_i}) w_k k. A block holds 64 particles, one
// per lane, and kRP waves that take every kRP-th column each, so that kRP waves per SIMD
// hide the latency of the per-column rotation chains; the waves' partial forces are
// added in wave order through the LDS.
constexpr int kRP = 4;
__global__ void __launch_bounds__(64 * kRP) k_recip_force(const double* __restrict__ pos, const double* __restrict__ q,
                                                          int n, EwCols cols, const double2* __restrict__ A,
                                                          double scale, double* __restrict__ F, int accumulate) {
  __shared__ double red[kRP - 1][3][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = xcd_block(blockIdx.x, gridDim.x) * 64 + lane;
  const bool on = i < n;
  const double x = on ? pos[3 * (size_t)i] : 0.0, y = on ? pos[3 * (size_t)i + 1] : 0.0,
               z = on ? pos[3 * (size_t)i + 2] : 0.0;
  double ezs, ezc;
  sincos(cols.tw[2] * z, &ezs, &ezc);
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int c = w; c < cols.ncol; c += kRP) {
    const int nx = cols.col[4 * c], ny = cols.col[4 * c + 1], lo = cols.col[4 * c + 2], hi = cols.col[4 * c + 3];
    const double kx = cols.tw[0] * nx, ky = cols.tw[1] * ny;
    double er, ei;
    sincos(kx * x + ky * y + cols.tw[2] * lo * z, &ei, &er);
    const double2* a = A + cols.off[c];
    double g = 0.0, gz = 0.0;
    for (int nz = lo; nz <= hi; ++nz) {
      const double2 s = *a++;
      // Im(conj(S) e) with S pre-multiplied by w: S.x * ei - S.y * er
      const double t = fma(s.x, ei, -s.y * er);
      g += t;
      gz = fma(t, (double)nz, gz);
      const double nr = fma(er, ezc, -ei * ezs);
      const double ni = fma(er, ezs, ei * ezc);
      er = nr;
      ei = ni;
    }
    fx = fma(kx, g, fx);
    fy = fma(ky, g, fy);
    fz = fma(cols.tw[2], gz, fz);
  }
  if (w > 0) {
    red[w - 1][0][lane] = fx;
    red[w - 1][1][lane] = fy;
    red[w - 1][2][lane] = fz;
  }
  __syncthreads();
  if (w != 0 || !on) return;
#pragma unroll
  for (int r = 0; r < kRP - 1; ++r) {
    fx += red[r][0][lane];
    fy += red[r][1][lane];
    fz += red[r][2][lane];
  }
  const double s = scale * q[i];
  if (accumulate) {
    F[3 * (size_t)i] += s * fx;
    F[3 * (size_t)i + 1] += s * fy;
    F[3 * (size_t)i + 2] += s * fz;
  } else {
    F[3 * (size_t)i] = s * fx;
    F[3 * (size_t)i + 1] = s * fy;
    F[3 * (size_t)i + 2] = s * fz;
  }
}

__global__ void k_scale1(double* v, double f) {
  if (threadIdx.x == 0 && blockIdx.x == 0) v[0] *= f;
}

}  // namespace accu

using namespace accu;

extern "C" {

// cols: host int32 [ncol][4] (nx, ny, nz_lo, nz_hi) over one half space.
int accu_ewald_recip(const double* pos, const double* q, int n, const double* box, double alpha, const int* cols_host,
                     int ncol, double k, int accumulate, double* F, double* U, void* stream) {
  ARG_CHECK(pos && q && box && F && cols_host, "null argument");
  ARG_CHECK(alpha > 0, "alpha must be positive");
  hipStream_t st = (hipStream_t)stream;
  if (n <= 0 || ncol <= 0) return ACCU_OK;
  // column offsets on the host
  long long K = 0;
  std::vector<long long> off(ncol);
  int maxlen = 0;
  for (int c = 0; c < ncol; ++c) {
    const int len = cols_host[4 * c + 3] - cols_host[4 * c + 2] + 1;
    ARG_CHECK(len >= 1, "empty k column %d", c);
    off[c] = K;
    K += len;
    maxlen = len > maxlen ? len : maxlen;
  }
  ARG_CHECK(maxlen <= kMaxCol, "k column of %d values exceeds %d", maxlen, kMaxCol);
  Workspace* ws = workspace_for(st);
  if (!ws) return ACCU_ERUNTIME;
  const int npb = (n + kET * kEM - 1) / (kET * kEM);
  if (ws->grow(ws->ew_col, ws->ew_col_cap, (size_t)ncol * 4) || ws->grow(ws->ew_off, ws->ew_off_cap, (size_t)ncol) ||
      ws->grow(ws->ew_A, ws->ew_A_cap, (size_t)K) || ws->grow(ws->ew_kw, ws->ew_kw_cap, (size_t)K) ||
      ws->grow(ws->ew_part, ws->ew_part_cap, (size_t)npb * K))
    return ACCU_ERUNTIME;
  struct {
    int* col;
    long long* off;
    double2 *A, *part;
    double* kw;
  } s{ws->ew_col, ws->ew_off, ws->ew_A, ws->ew_part, ws->ew_kw};
  HIP_TRY(hipMemcpyAsync(s.col, cols_host, (size_t)ncol * 4 * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(s.off, off.data(), (size_t)ncol * sizeof(long long), hipMemcpyHostToDevice, st));
  EwCols ec{s.col, s.off, ncol, {2 * M_PI / box[0], 2 * M_PI / box[1], 2 * M_PI / box[2]}};
  const double V = box[0] * box[1] * box[2];
  const int nbk = (int)((K + kET - 1) / kET);
  int rc = ws->ensure((size_t)nbk * 2 + 16);
  if (rc) return rc;
  hipLaunchKernelGGL(k_kweights, dim3((ncol + 255) / 256), dim3(256), 0, st, ec, 1.0 / (4 * alpha * alpha), s.kw);
  // enough column chunks to give ~4 blocks per CU
  int cpb = (int)((long long)ncol * npb / 1024);
  cpb = cpb < 1 ? 1 : (cpb > 64 ? 64 : cpb);
  const int nchunk = (ncol + cpb - 1) / cpb;
  hipLaunchKernelGGL(k_sk_partial, dim3(nchunk, npb), dim3(kET), 0, st, pos, q, n, ec, cpb, K, s.part);
  double* Ud = U ? U : ws->out;
  hipLaunchKernelGGL(k_sk_reduce, dim3(nbk), dim3(kET), 0, st, s.part, npb, K, s.kw, s.A, ws->partials, ws->counter,
                     Ud);
  // U_rec = 2 * (2 pi / V) * sum_half w |S|^2, times the Coulomb constant
  if (U) hipLaunchKernelGGL(k_scale1, dim3(1), dim3(64), 0, st, U, 4.0 * M_PI / V * k);
  const int nbf = (n + 63) / 64;
  hipLaunchKernelGGL(k_recip_force, dim3(nbf), dim3(64 * kRP), 0, st, pos, q, n, ec, s.A, 2.0 * 4.0 * M_PI / V * k,
                     F, accumulate);
  LAUNCH_CHECK();
  return ACCU_OK;
}

}  // extern "C"
