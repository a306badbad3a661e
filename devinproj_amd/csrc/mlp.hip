// Learned DSF -> Ewald pair correction: a 5 -> 64 -> 64 -> 1 MLP per directed pair,
// layer 2 (and the backward GEMMs) on gfx950 bf16 MFMA with float32 accumulation.
// This has no counterpart in the reference: it corrects DampedShiftedCoulomb
// (forces.py:391-452) toward Ewald (forces.py:333-388).
//
//   n_i  = sum_j sw(r_ij)                          (cosine switch r_on -> r_off)
//   x    = [r / rc, q_i q_j, (n_i + n_j) s_n, |n_i - n_j| s_n, q_i + q_j]   (symmetric in i <-> j;
//          q_i + q_j tells like pairs of different species apart, e.g. Na-Na from Cl-Cl)
//   h1   = silu(W1 x + b1);  h2 = silu(W2 h1 + b2)  (W2, h1 in bfloat16)
//   s    = (w3 . h2 + b3) (1 - (r/rc)^2)^2
//   F_i += F0 s d / r,  d = r_j - r_i
// Both directions of a pair see identical features, so s_ij == s_ji bit for bit and
// the correction forces cancel exactly in pairs (momentum is conserved).
//
// Work layout: one wave per atom. Its neighbours inside rc are compacted (ballot) into
// an LDS queue and processed 32 pairs at a time. For a tile, lane l handles pair
// p = l % 32 and the half hk = l / 32 of every 16-wide K block:
//   layer 2:  D[u][p] = sum_v W2[u][v] h1[p][v]  as two 32x32x16 MFMA chains
//             (A = W2 rows u, B = h1^T), so each lane ends up holding 32 hidden units
//             of one pair and layer 3 is a per-lane dot product plus one lane swap.
#include "common.h"
#include "nlist.h"
#include "../../include/accusim_hip.h"

#include <cmath>
#include <vector>

namespace accu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kH = 64;
constexpr int kIn = 5;  // input features
constexpr int kNParam = kH * kIn + kH + kH * kH + kH + kH + 1;  // 4609
constexpr int kOffW1 = 0, kOffB1 = kH * kIn, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH,
              kOffB3 = kOffW3 + kH;
constexpr int kQ = 96;  // per-wave queue of compacted pairs

struct MLPGeom {
  float rc, inv_rc, rc2, ns;  // ns = 1 / n_scale
  float F0;
  float r_on, r_off;
};

struct QEntry {
  int j;
  float dx, dy, dz, r;
};

__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
  return sg * (1.0f + z * (1.0f - sg));
}

// Coordination numbers n_i = sum_j sw(r_ij) over the list (float64 positions).
__global__ void k_mlp_coord(const double* __restrict__ pos, int ps, int n, const int* __restrict__ sorted,
                            const int* __restrict__ nn, const int* __restrict__ nbr, int stride, Box box, MLPGeom g,
                            float* __restrict__ coord) {
  // a quad of lanes per atom, lane s taking list entries s, s + 4, ... (four times the
  // waves of one thread per atom, for the latency of the gathers); quad sums in a fixed order
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = gt >> 2, sub = gt & 3;
  const bool on = t < n;
  float c = 0.f;
  int i = 0;
  if (on) {
    i = sorted ? sorted[t] : t;
    const double xi = pos[(size_t)ps * i], yi = pos[(size_t)ps * i + 1], zi = pos[(size_t)ps * i + 2];
    const float w = g.r_off - g.r_on;
    const int cnt = nn[t];
    for (int k = sub; k < cnt; k += 4) {
      const int j = nbr[(size_t)k * stride + t];
      double dx = pos[(size_t)ps * j] - xi, dy = pos[(size_t)ps * j + 1] - yi, dz = pos[(size_t)ps * j + 2] - zi;
      if (box.periodic) {
        dx = minimg_fast(dx, box.L[0], box.inv[0]);
        dy = minimg_fast(dy, box.L[1], box.inv[1]);
        dz = minimg_fast(dz, box.L[2], box.inv[2]);
      }
      const float r = (float)sqrt(dx * dx + dy * dy + dz * dz);
      if (r < g.r_off) {
        const float xr = fminf(fmaxf((r - g.r_on) / w, 0.f), 1.f);
        c += 0.5f * (1.f + cosf(3.14159265358979f * xr));
      }
    }
  }
  c = quad_sumf(c);
  if (on && sub == 0) coord[i] = c;
}

// Parameters staged in LDS (float32) and the per-lane W2 operands (bf16).
struct MLPShared {
  float W1[kH * kIn];
  float b1[kH];
  float b2[kH];
  float w3[kH];
  float b3;
};

__device__ __forceinline__ void stage_params(const float* __restrict__ prm, MLPShared& sh) {
  for (int k = threadIdx.x; k < kH * kIn; k += blockDim.x) sh.W1[k] = prm[kOffW1 + k];
  for (int k = threadIdx.x; k < kH; k += blockDim.x) {
    sh.b1[k] = prm[kOffB1 + k];
    sh.b2[k] = prm[kOffB2 + k];
    sh.w3[k] = prm[kOffW3 + k];
  }
  if (threadIdx.x == 0) sh.b3 = prm[kOffB3];
}

// Forward of one tile. Returns s for pair p (both halves) and leaves the layer-2 D
// values and the features in the caller's registers when asked (training).
struct TileState {
  float x[kIn];
  float r, env;
  bf16x8 bh1[4];  // h1 of pair p for units kb*16 + hk*8 + t
  f32x16 z2[2];   // layer-2 pre-activations (with bias), D layout
  float s_raw;
};

__device__ __forceinline__ void tile_forward(const MLPShared& sh, const bf16x8 (&aW2)[2][4], int hk, TileState& st) {
  // layer 1 (VALU, float32): 32 units per lane, packed to bf16 as the MFMA B operand
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int u = kb * 16 + hk * 8 + t;
      float z = sh.b1[u];
#pragma unroll
      for (int c = 0; c < kIn; ++c) z = fmaf(sh.W1[kIn * u + c], st.x[c], z);
      st.bh1[kb][t] = (__bf16)silu(z);
    }
  }
  // layer 2 (MFMA 32x32x16 bf16 -> f32)
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
    f32x16 acc = {};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aW2[mb][kb], st.bh1[kb], acc, 0, 0, 0);
    st.z2[mb] = acc;
  }
  // bias, activation, layer 3
  float part = 0.f;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int u = mb * 32 + 8 * (r >> 2) + 4 * hk + (r & 3);
      const float z = st.z2[mb][r] + sh.b2[u];
      st.z2[mb][r] = z;
      part = fmaf(sh.w3[u], silu(z), part);
    }
  }
  part += __shfl_xor(part, 32, 64);
  st.s_raw = part + sh.b3;
}

// Load one compacted pair of the wave's atom i and build its features.
__device__ __forceinline__ void tile_features(const QEntry& e, bool valid, float qi, float ni,
                                              const float* __restrict__ q, const float* __restrict__ coord,
                                              const MLPGeom& g, TileState& st) {
  const float qj = valid ? q[e.j] : 0.f;
  const float nj = valid ? coord[e.j] : ni;
  st.r = valid ? e.r : g.rc;
  st.x[0] = st.r * g.inv_rc;
  st.x[1] = qi * qj;
  st.x[2] = (ni + nj) * g.ns;
  st.x[3] = fabsf(ni - nj) * g.ns;
  st.x[4] = qi + qj;
  const float a = 1.f - st.r * st.r * g.inv_rc * g.inv_rc;
  st.env = valid ? a * a : 0.f;
}

__device__ __forceinline__ void load_w2_operands(const __bf16* __restrict__ W2b, int p, int hk, bf16x8 (&a)[2][4]) {
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
      a[mb][kb] = *reinterpret_cast<const bf16x8*>(W2b + (size_t)(mb * 32 + p) * kH + kb * 16 + hk * 8);
}

// Compact the neighbours of atom i inside rc into the wave's queue; returns the new count.
__device__ __forceinline__ int enqueue_chunk(int base, int cnt, int i, int t, const double* __restrict__ pos, int ps,
                                             const int* __restrict__ nbr, int stride, const Box& box, const MLPGeom& g,
                                             double xi, double yi, double zi, QEntry* qbuf, int qcount) {
  const int lane = threadIdx.x & 63;
  const int k = base + lane;
  bool inside = false;
  QEntry e{0, 0.f, 0.f, 0.f, 0.f};
  if (k < cnt) {
    const int j = nbr[(size_t)k * stride + t];
    double dx = pos[(size_t)ps * j] - xi, dy = pos[(size_t)ps * j + 1] - yi, dz = pos[(size_t)ps * j + 2] - zi;
    if (box.periodic) {
      dx = minimg_fast(dx, box.L[0], box.inv[0]);
      dy = minimg_fast(dy, box.L[1], box.inv[1]);
      dz = minimg_fast(dz, box.L[2], box.inv[2]);
    }
    const double r2 = dx * dx + dy * dy + dz * dz;
    inside = r2 < (double)g.rc * (double)g.rc;
    e = QEntry{j, (float)dx, (float)dy, (float)dz, (float)sqrt(r2)};
  }
  const unsigned long long m = __ballot(inside);
  const int before = __popcll(m & ((1ull << lane) - 1ull));
  if (inside) qbuf[qcount + before] = e;
  return qcount + __popcll(m);
}

// Correction forces for every atom (added to F when accumulate != 0).
__global__ void __launch_bounds__(256) k_mlp_force(const double* __restrict__ pos, int ps, int n,
                                                   const int* __restrict__ sorted, const int* __restrict__ nn,
                                                   const int* __restrict__ nbr, int stride, Box box, MLPGeom g,
                                                   const float* __restrict__ prm, const __bf16* __restrict__ W2b,
                                                   const float* __restrict__ q, const float* __restrict__ coord,
                                                   double* __restrict__ F, int accumulate) {
  __shared__ MLPShared sh;
  __shared__ QEntry qall[4][kQ];
  stage_params(prm, sh);
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, p = lane & 31, hk = lane >> 5;
  QEntry* qb = qall[w];
  bf16x8 aW2[2][4];
  load_w2_operands(W2b, p, hk, aW2);
  const int nwaves_total = gridDim.x * (blockDim.x >> 6);
  for (int t = blockIdx.x * (blockDim.x >> 6) + w; t < n; t += nwaves_total) {
    const int i = sorted ? sorted[t] : t;
    const double xi = pos[(size_t)ps * i], yi = pos[(size_t)ps * i + 1], zi = pos[(size_t)ps * i + 2];
    const float qi = q[i], ni = coord[i];
    const int cnt = nn[t];
    // float64 per-lane sums: s * dx of two floats is exact in float64, so pair terms
    // (i, j) and (j, i) cancel to float64 rounding and the total force stays ~0
    double fx = 0.0, fy = 0.0, fz = 0.0;
    int qc = 0;
    for (int base = 0; base < cnt || qc > 0; base += 64) {
      if (base < cnt) qc = enqueue_chunk(base, cnt, i, t, pos, ps, nbr, stride, box, g, xi, yi, zi, qb, qc);
      // process full tiles, and the remainder once the list is exhausted
      while (qc >= 32 || (base + 64 >= cnt && qc > 0)) {
        const int m = qc < 32 ? qc : 32;
        const bool valid = p < m;
        const QEntry e = qb[valid ? p : 0];
        TileState st;
        tile_features(e, valid, qi, ni, q, coord, g, st);
        tile_forward(sh, aW2, hk, st);
        const float s = st.s_raw * st.env * __builtin_amdgcn_rcpf(st.r);
        if (hk == 0 && valid) {
          fx += (double)s * (double)e.dx;
          fy += (double)s * (double)e.dy;
          fz += (double)s * (double)e.dz;
        }
        // shift the queue down by m
        const int rest = qc - m;
        QEntry mv{};
        if (lane < rest) mv = qb[m + lane];
        __builtin_amdgcn_wave_barrier();
        if (lane < rest) qb[lane] = mv;
        __builtin_amdgcn_wave_barrier();
        qc = rest;
      }
    }
    double sx = wave_sum(fx), sy = wave_sum(fy), sz = wave_sum(fz);
    if (lane == 0) {
      const double f0 = (double)g.F0;
      if (accumulate) {
        F[3 * (size_t)i] += f0 * sx;
        F[3 * (size_t)i + 1] += f0 * sy;
        F[3 * (size_t)i + 2] += f0 * sz;
      } else {
        F[3 * (size_t)i] = f0 * sx;
        F[3 * (size_t)i + 1] = f0 * sy;
        F[3 * (size_t)i + 2] = f0 * sz;
      }
    }
  }
}

// ---------------------------------------------------------------- training

// Per-wave LDS staging for the layout changes between the MFMA products (bf16).
// 16-byte aligned: the MFMA operands are read back as 16-byte bf16x8 vectors.
struct __attribute__((aligned(16))) GradScratch {
  __bf16 h1T[kH * 32];   // [v][p]
  __bf16 dz2[kH * 32];   // [u][p]
  __bf16 dz2T[32 * kH];  // [p][u]
  __bf16 dz1[kH * 32];   // [v][p]
  __bf16 xT[16 * 32];    // [c][p], c = kIn is the bias column
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Gradient of L = (1 / 3N) sum_i |(F_corr_i - target_i) / F0|^2 for one snapshot.
// Each wave accumulates its atoms' contributions in registers (MFMA accumulators for
// dW2 and dW1|db1) and writes one partial gradient vector; a second kernel sums the
// partials in wave order, so the result does not depend on scheduling.
//   ds      = dL/ds_ij = (2/3N) e_i . d_ij / r_ij * env(r_ij)
//   dW2    += dz2 (u x p) h1 (p x v)          MFMA 32x32x16, K = pairs
//   dh1     = W2^T (v x u) dz2 (u x p)         MFMA 32x32x16, K = hidden units
//   dW1|db1 += dz1 (v x p) [x | 1] (p x c)     MFMA 16x16x32, K = pairs
__global__ void __launch_bounds__(256) k_mlp_grad(const double* __restrict__ pos, int ps, int n,
                                                  const int* __restrict__ sorted, const int* __restrict__ nn,
                                                  const int* __restrict__ nbr, int stride, Box box, MLPGeom g,
                                                  const float* __restrict__ prm, const __bf16* __restrict__ W2b,
                                                  const __bf16* __restrict__ W2Tb, const float* __restrict__ q,
                                                  const float* __restrict__ coord, const double* __restrict__ Fcorr,
                                                  const double* __restrict__ target, double inv_m,
                                                  float* __restrict__ gpart, double* __restrict__ lpart) {
  __shared__ MLPShared sh;
  __shared__ QEntry qall[4][kQ];
  __shared__ GradScratch gs[4];
  stage_params(prm, sh);
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, p = lane & 31, hk = lane >> 5;
  QEntry* qb = qall[w];
  GradScratch& S = gs[w];
  bf16x8 aW2[2][4], aW2T[2][4];
  load_w2_operands(W2b, p, hk, aW2);
  load_w2_operands(W2Tb, p, hk, aW2T);
  f32x16 gW2[2][2];
  f32x4 gW1[4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) gW2[a][b] = f32x16{};
#pragma unroll
  for (int a = 0; a < 4; ++a) gW1[a] = f32x4{};
  float gw3[32], gb2[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) gw3[k] = gb2[k] = 0.f;
  float gb3 = 0.f;
  double loss = 0.0;
  const int wave_id = blockIdx.x * (blockDim.x >> 6) + w;
  const int nwaves_total = gridDim.x * (blockDim.x >> 6);
  const double invF0 = 1.0 / (double)g.F0;
  for (int t = wave_id; t < n; t += nwaves_total) {
    const int i = sorted ? sorted[t] : t;
    const double xi = pos[(size_t)ps * i], yi = pos[(size_t)ps * i + 1], zi = pos[(size_t)ps * i + 2];
    const float qi = q[i], ni = coord[i];
    const double ex = (Fcorr[3 * (size_t)i] - target[3 * (size_t)i]) * invF0;
    const double ey = (Fcorr[3 * (size_t)i + 1] - target[3 * (size_t)i + 1]) * invF0;
    const double ez = (Fcorr[3 * (size_t)i + 2] - target[3 * (size_t)i + 2]) * invF0;
    if (lane == 0) loss += (ex * ex + ey * ey + ez * ez) * inv_m;
    const float gx = (float)(2.0 * ex * inv_m), gy = (float)(2.0 * ey * inv_m), gz = (float)(2.0 * ez * inv_m);
    const int cnt = nn[t];
    int qc = 0;
    for (int base = 0; base < cnt || qc > 0; base += 64) {
      if (base < cnt) qc = enqueue_chunk(base, cnt, i, t, pos, ps, nbr, stride, box, g, xi, yi, zi, qb, qc);
      while (qc >= 32 || (base + 64 >= cnt && qc > 0)) {
        const int m = qc < 32 ? qc : 32;
        const bool valid = p < m;
        const QEntry e = qb[valid ? p : 0];
        TileState st;
        tile_features(e, valid, qi, ni, q, coord, g, st);
        tile_forward(sh, aW2, hk, st);
        const float ds = valid ? (gx * e.dx + gy * e.dy + gz * e.dz) * __builtin_amdgcn_rcpf(st.r) * st.env : 0.f;
        gb3 += (hk == 0) ? ds : 0.f;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int u = mb * 32 + 8 * (r >> 2) + 4 * hk + (r & 3);
            const float z = st.z2[mb][r];
            gw3[mb * 16 + r] = fmaf(ds, silu(z), gw3[mb * 16 + r]);
            const float dz = ds * sh.w3[u] * dsilu(z);
            gb2[mb * 16 + r] += dz;
            const __bf16 bz = (__bf16)dz;
            S.dz2[u * 32 + p] = bz;
            S.dz2T[p * kH + u] = bz;
          }
        }
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
#pragma unroll
          for (int tt = 0; tt < 8; ++tt) S.h1T[(kb * 16 + hk * 8 + tt) * 32 + p] = st.bh1[kb][tt];
        if (hk == 0) {
#pragma unroll
          for (int c = 0; c < 16; ++c) {
            const float xv = !valid ? 0.f : (c < kIn ? st.x[c] : (c == kIn ? 1.f : 0.f));
            S.xT[c * 32 + p] = (__bf16)xv;
          }
        }
        wave_sync();
        // dW2 += dz2 h1
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
          for (int mb = 0; mb < 2; ++mb) {
            const bf16x8 A = *reinterpret_cast<const bf16x8*>(&S.dz2[(mb * 32 + p) * 32 + kb * 16 + hk * 8]);
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
              const bf16x8 Bm = *reinterpret_cast<const bf16x8*>(&S.h1T[(nb * 32 + p) * 32 + kb * 16 + hk * 8]);
              gW2[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, Bm, gW2[mb][nb], 0, 0, 0);
            }
          }
        }
        // dh1 = W2^T dz2, then dz1 = dh1 * silu'(z1) (z1 recomputed for the D-layout units)
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
          f32x16 dh = {};
#pragma unroll
          for (int kb = 0; kb < 4; ++kb) {
            const bf16x8 Bm = *reinterpret_cast<const bf16x8*>(&S.dz2T[p * kH + kb * 16 + hk * 8]);
            dh = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aW2T[mb][kb], Bm, dh, 0, 0, 0);
          }
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int v = mb * 32 + 8 * (r >> 2) + 4 * hk + (r & 3);
            float z1 = sh.b1[v];
#pragma unroll
            for (int c = 0; c < kIn; ++c) z1 = fmaf(sh.W1[kIn * v + c], st.x[c], z1);
            S.dz1[v * 32 + p] = (__bf16)(dh[r] * dsilu(z1));
          }
        }
        wave_sync();
        // dW1 | db1 += dz1 [x | 1]
#pragma unroll
        for (int vb = 0; vb < 4; ++vb) {
          const bf16x8 A = *reinterpret_cast<const bf16x8*>(&S.dz1[(vb * 16 + (lane & 15)) * 32 + (lane >> 4) * 8]);
          const bf16x8 Bm = *reinterpret_cast<const bf16x8*>(&S.xT[(lane & 15) * 32 + (lane >> 4) * 8]);
          gW1[vb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bm, gW1[vb], 0, 0, 0);
        }
        wave_sync();
        const int rest = qc - m;
        QEntry mv{};
        if (lane < rest) mv = qb[m + lane];
        wave_sync();
        if (lane < rest) qb[lane] = mv;
        wave_sync();
        qc = rest;
      }
    }
  }
  // one partial gradient per wave
  float* gp = gpart + (size_t)wave_id * kNParam;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int u = mb * 32 + 8 * (r >> 2) + 4 * hk + (r & 3), v = nb * 32 + p;
        gp[kOffW2 + u * kH + v] = gW2[mb][nb][r];
      }
#pragma unroll
  for (int vb = 0; vb < 4; ++vb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int v = vb * 16 + (lane >> 4) * 4 + r, c = lane & 15;
      if (c < kIn) gp[kOffW1 + v * kIn + c] = gW1[vb][r];
      else if (c == kIn) gp[kOffB1 + v] = gW1[vb][r];
    }
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    float a = gw3[k], b = gb2[k];
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
      a += __shfl_xor(a, o, 64);
      b += __shfl_xor(b, o, 64);
    }
    if (p == 0) {
      const int mb = k >> 4, r = k & 15;
      const int u = mb * 32 + 8 * (r >> 2) + 4 * hk + (r & 3);
      gp[kOffW3 + u] = a;
      gp[kOffB2 + u] = b;
    }
  }
  const double b3 = wave_sum((double)gb3);
  if (lane == 0) {
    gp[kOffB3] = (float)b3;
    lpart[wave_id] = loss;
  }
}

// grad[k] = sum over waves (fixed order); loss likewise.
__global__ void k_mlp_reduce(const float* __restrict__ gpart, const double* __restrict__ lpart, int nw,
                             float* __restrict__ grad, double* __restrict__ loss) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < kNParam) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += gpart[(size_t)w * kNParam + k];
    grad[k] = (float)s;
  }
  if (k == kNParam) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += lpart[w];
    *loss = s;
  }
}

// Adam on the float32 master parameters; refreshes the bf16 copies of W2 and W2^T.
__global__ void k_mlp_adam(float* __restrict__ prm, float* __restrict__ m1, float* __restrict__ m2,
                           const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1, float c2,
                           __bf16* __restrict__ W2b, __bf16* __restrict__ W2Tb) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kNParam) return;
  const float gk = grad[k];
  const float a = b1 * m1[k] + (1.f - b1) * gk;
  const float b = b2 * m2[k] + (1.f - b2) * gk * gk;
  m1[k] = a;
  m2[k] = b;
  const float v = prm[k] - lr * (a * c1) / (sqrtf(b * c2) + eps);
  prm[k] = v;
  if (k >= kOffW2 && k < kOffW2 + kH * kH) {
    const int u = (k - kOffW2) / kH, c = (k - kOffW2) % kH;
    W2b[u * kH + c] = (__bf16)v;
    W2Tb[c * kH + u] = (__bf16)v;
  }
}

__global__ void k_mlp_sync_w2(const float* __restrict__ prm, __bf16* __restrict__ W2b, __bf16* __restrict__ W2Tb) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kH * kH) return;
  const int u = k / kH, c = k % kH;
  const __bf16 v = (__bf16)prm[kOffW2 + k];
  W2b[u * kH + c] = v;
  W2Tb[c * kH + u] = v;
}

__global__ void k_to_float(const double* __restrict__ a, double scale, int n, float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = (float)(a[k] * scale);
}

struct MLP {
  float* prm = nullptr;  // kNParam
  float* m1 = nullptr;
  float* m2 = nullptr;
  float* grad = nullptr;
  __bf16* W2b = nullptr;
  __bf16* W2Tb = nullptr;
  float* q = nullptr;
  float* coord = nullptr;
  double* Fcorr = nullptr;
  int cap_n = 0;
  float* gpart = nullptr;
  double* lpart = nullptr;
  double* loss = nullptr;
  int nwaves = 0;
  int step = 0;
  MLPGeom g{};

  int ensure(int n) {
    if (n > cap_n) {
      if (q) HIP_TRY(hipFree(q));
      if (coord) HIP_TRY(hipFree(coord));
      if (Fcorr) HIP_TRY(hipFree(Fcorr));
      HIP_TRY(hipMalloc(&q, (size_t)n * sizeof(float)));
      HIP_TRY(hipMalloc(&coord, (size_t)n * sizeof(float)));
      HIP_TRY(hipMalloc(&Fcorr, (size_t)n * 3 * sizeof(double)));
      cap_n = n;
    }
    return ACCU_OK;
  }
  void release() {
    void* ps[] = {prm, m1, m2, grad, W2b, W2Tb, q, coord, Fcorr, gpart, lpart, loss};
    for (void* x : ps)
      if (x) (void)hipFree(x);
  }
};

}  // namespace accu

using namespace accu;

static Box box_of(const double* box, int periodic) {
  Box b{};
  b.periodic = periodic && box;
  for (int d = 0; d < 3; ++d) {
    b.L[d] = box ? box[d] : 1.0;
    b.inv[d] = 1.0 / b.L[d];
  }
  return b;
}

extern "C" {

int accu_mlp_create(accu_mlp** out, const float* params_host) {
  ARG_CHECK(out && params_host, "null argument");
  MLP* m = new MLP();
  auto fail = [&](hipError_t e) {
    set_error("hipMalloc: %s", hipGetErrorString(e));
    m->release();
    delete m;
    return ACCU_ERUNTIME;
  };
  hipError_t e;
  if ((e = hipMalloc(&m->prm, kNParam * sizeof(float))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&m->m1, kNParam * sizeof(float))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&m->m2, kNParam * sizeof(float))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&m->grad, kNParam * sizeof(float))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&m->W2b, kH * kH * sizeof(__bf16))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&m->W2Tb, kH * kH * sizeof(__bf16))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&m->loss, 64)) != hipSuccess) return fail(e);
  HIP_TRY(hipMemcpy(m->prm, params_host, kNParam * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(m->m1, 0, kNParam * sizeof(float)));
  HIP_TRY(hipMemset(m->m2, 0, kNParam * sizeof(float)));
  hipLaunchKernelGGL(k_mlp_sync_w2, dim3(kH * kH / 256), dim3(256), 0, 0, m->prm, m->W2b, m->W2Tb);
  LAUNCH_CHECK();
  HIP_TRY(hipDeviceSynchronize());
  *out = reinterpret_cast<accu_mlp*>(m);
  return ACCU_OK;
}

int accu_mlp_destroy(accu_mlp* h) {
  if (h) {
    MLP* m = reinterpret_cast<MLP*>(h);
    m->release();
    delete m;
  }
  return ACCU_OK;
}

int accu_mlp_params(accu_mlp* h, float* get_host, const float* set_host) {
  ARG_CHECK(h, "null handle");
  MLP& m = *reinterpret_cast<MLP*>(h);
  if (set_host) {
    HIP_TRY(hipMemcpy(m.prm, set_host, kNParam * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_sync_w2, dim3(kH * kH / 256), dim3(256), 0, 0, m.prm, m.W2b, m.W2Tb);
    LAUNCH_CHECK();
    HIP_TRY(hipDeviceSynchronize());
  }
  if (get_host) HIP_TRY(hipMemcpy(get_host, m.prm, kNParam * sizeof(float), hipMemcpyDeviceToHost));
  return ACCU_OK;
}

int accu_mlp_configure(accu_mlp* h, double rc, double r_on, double r_off, double n_scale, double F0) {
  ARG_CHECK(h, "null handle");
  ARG_CHECK(rc > 0 && r_off > r_on && n_scale > 0 && F0 > 0, "bad MLP geometry");
  MLP& m = *reinterpret_cast<MLP*>(h);
  m.g.rc = (float)rc;
  m.g.inv_rc = (float)(1.0 / rc);
  m.g.rc2 = (float)(rc * rc);
  m.g.ns = (float)(1.0 / n_scale);
  m.g.F0 = (float)F0;
  m.g.r_on = (float)r_on;
  m.g.r_off = (float)r_off;
  return ACCU_OK;
}

// Per-atom inputs for a snapshot: charges (device float64, multiplied by qscale) and
// coordination numbers from the list (which must reach at least r_off).
int accu_mlp_prepare(accu_mlp* h, accu_nlist* nlh, const double* pos, const double* q, double qscale, int n,
                     const double* box, int periodic, void* stream) {
  ARG_CHECK(h && nlh && pos && q, "null argument");
  MLP& m = *reinterpret_cast<MLP*>(h);
  NList& nl = *reinterpret_cast<NList*>(nlh);
  ARG_CHECK(nl.n == n, "neighbour list built for %d atoms, got %d", nl.n, n);
  hipStream_t st = (hipStream_t)stream;
  int rc = m.ensure(n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_to_float, dim3((n + 255) / 256), dim3(256), 0, st, q, qscale, n, m.q);
  hipLaunchKernelGGL(k_mlp_coord, dim3((int)((4LL * n + 255) / 256)), dim3(256), 0, st, pos, 3, n, nl.sorted, nl.nn,
                     nl.nbr, nl.stride, box_of(box, periodic), m.g, m.coord);
  LAUNCH_CHECK();
  return ACCU_OK;
}

static int mlp_forward(MLP& m, NList& nl, const double* pos, int n, const Box& b, double* F, int accumulate,
                       hipStream_t st) {
  const int blocks = (n + 3) / 4 < 2048 ? (n + 3) / 4 : 2048;
  hipLaunchKernelGGL(k_mlp_force, dim3(blocks), dim3(256), 0, st, pos, 3, n, nl.sorted, nl.nn, nl.nbr, nl.stride, b,
                     m.g, m.prm, m.W2b, m.q, m.coord, F, accumulate);
  LAUNCH_CHECK();
  return ACCU_OK;
}

int accu_mlp_forces(accu_mlp* h, accu_nlist* nlh, const double* pos, int n, const double* box, int periodic,
                    int accumulate, double* F, void* stream) {
  ARG_CHECK(h && nlh && pos && F, "null argument");
  MLP& m = *reinterpret_cast<MLP*>(h);
  NList& nl = *reinterpret_cast<NList*>(nlh);
  ARG_CHECK(nl.n == n && m.cap_n >= n, "call accu_mlp_prepare for this snapshot first");
  return mlp_forward(m, nl, pos, n, box_of(box, periodic), F, accumulate, (hipStream_t)stream);
}

// Gradient of the snapshot loss against target (device float64 [n][3], newtons).
// grad_host (may be NULL) receives the gradient; loss_host the loss (synchronises).
int accu_mlp_grad(accu_mlp* h, accu_nlist* nlh, const double* pos, int n, const double* box, int periodic,
                  const double* target, float* grad_host, double* loss_host, void* stream) {
  ARG_CHECK(h && nlh && pos && target, "null argument");
  MLP& m = *reinterpret_cast<MLP*>(h);
  NList& nl = *reinterpret_cast<NList*>(nlh);
  ARG_CHECK(nl.n == n && m.cap_n >= n, "call accu_mlp_prepare for this snapshot first");
  hipStream_t st = (hipStream_t)stream;
  const Box b = box_of(box, periodic);
  int rc = mlp_forward(m, nl, pos, n, b, m.Fcorr, 0, st);
  if (rc) return rc;
  const int blocks = (n + 3) / 4 < 1024 ? (n + 3) / 4 : 1024;
  const int nw = blocks * 4;
  if (nw > m.nwaves) {
    if (m.gpart) HIP_TRY(hipFree(m.gpart));
    if (m.lpart) HIP_TRY(hipFree(m.lpart));
    HIP_TRY(hipMalloc(&m.gpart, (size_t)nw * kNParam * sizeof(float)));
    HIP_TRY(hipMalloc(&m.lpart, (size_t)nw * sizeof(double)));
    m.nwaves = nw;
  }
  hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(256), 0, st, pos, 3, n, nl.sorted, nl.nn, nl.nbr, nl.stride, b,
                     m.g, m.prm, m.W2b, m.W2Tb, m.q, m.coord, m.Fcorr, target, 1.0 / (3.0 * n), m.gpart, m.lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParam + 1 + 255) / 256), dim3(256), 0, st, m.gpart, m.lpart, nw, m.grad,
                     m.loss);
  LAUNCH_CHECK();
  if (grad_host) HIP_TRY(hipMemcpyAsync(grad_host, m.grad, kNParam * sizeof(float), hipMemcpyDeviceToHost, st));
  if (loss_host) HIP_TRY(hipMemcpyAsync(loss_host, m.loss, sizeof(double), hipMemcpyDeviceToHost, st));
  if (grad_host || loss_host) HIP_TRY(hipStreamSynchronize(st));
  return ACCU_OK;
}

// One Adam step with the gradient of the last accu_mlp_grad call.
int accu_mlp_adam(accu_mlp* h, double lr, double b1, double b2, double eps, void* stream) {
  ARG_CHECK(h, "null handle");
  MLP& m = *reinterpret_cast<MLP*>(h);
  m.step += 1;
  const float c1 = (float)(1.0 / (1.0 - std::pow(b1, m.step)));
  const float c2 = (float)(1.0 / (1.0 - std::pow(b2, m.step)));
  hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, (hipStream_t)stream, m.prm, m.m1, m.m2,
                     m.grad, (float)lr, (float)b1, (float)b2, (float)eps, c1, c2, m.W2b, m.W2Tb);
  LAUNCH_CHECK();
  return ACCU_OK;
}

}  // extern "C"
