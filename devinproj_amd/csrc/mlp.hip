// Learned DSF -> Ewald pair correction: a 4 -> 64 -> 64 -> 1 MLP per directed pair,
// layer 2 (and the backward GEMMs) on gfx950 bf16 MFMA with float32 accumulation.
// This has no counterpart in the reference: it corrects DampedShiftedCoulomb
// (forces.py:391-452) toward Ewald (forces.py:333-388).
//
//   n_i  = sum_j sw(r_ij)                          (cosine switch r_on -> r_off)
//   x    = [r / rc, q_i q_j, (n_i + n_j) s_n, |n_i - n_j| s_n]   (symmetric in i <-> j)
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
constexpr int kNParam = kH * 4 + kH + kH * kH + kH + kH + 1;  // 4545
constexpr int kOffW1 = 0, kOffB1 = 256, kOffW2 = 320, kOffB2 = 4416, kOffW3 = 4480, kOffB3 = 4544;
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
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = sorted ? sorted[t] : t;
  const double xi = pos[(size_t)ps * i], yi = pos[(size_t)ps * i + 1], zi = pos[(size_t)ps * i + 2];
  float c = 0.f;
  const float w = g.r_off - g.r_on;
  for (int k = 0; k < nn[t]; ++k) {
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
  coord[i] = c;
}

// Parameters staged in LDS (float32) and the per-lane W2 operands (bf16).
struct MLPShared {
  float W1[kH * 4];
  float b1[kH];
  float b2[kH];
  float w3[kH];
  float b3;
};

__device__ __forceinline__ void stage_params(const float* __restrict__ prm, MLPShared& sh) {
  for (int k = threadIdx.x; k < kH * 4; k += blockDim.x) sh.W1[k] = prm[kOffW1 + k];
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
  float x[4];
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
      z = fmaf(sh.W1[4 * u], st.x[0], z);
      z = fmaf(sh.W1[4 * u + 1], st.x[1], z);
      z = fmaf(sh.W1[4 * u + 2], st.x[2], z);
      z = fmaf(sh.W1[4 * u + 3], st.x[3], z);
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
    float fx = 0.f, fy = 0.f, fz = 0.f;
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
          fx = fmaf(s, e.dx, fx);
          fy = fmaf(s, e.dy, fy);
          fz = fmaf(s, e.dz, fz);
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
    double sx = wave_sum((double)fx), sy = wave_sum((double)fy), sz = wave_sum((double)fz);
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

}  // namespace accu
