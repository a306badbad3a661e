// Shared device/host helpers for the accusim force-path library (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#define ACCU_OK 0
#define ACCU_EINVAL (-1)     // invalid physics setup / arguments  -> ValueError
#define ACCU_ERUNTIME (-2)   // HIP / device failure                -> RuntimeError
#define ACCU_EOVERFLOW (-3)  // a fixed-capacity buffer overflowed  -> RuntimeError

namespace accu {
void set_error(const char* fmt, ...);
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      accu::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return ACCU_ERUNTIME;                                                                  \
    }                                                                                        \
  } while (0)

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

#define ARG_CHECK(cond, ...)            \
  do {                                  \
    if (!(cond)) {                      \
      accu::set_error(__VA_ARGS__);     \
      return ACCU_EINVAL;               \
    }                                   \
  } while (0)

namespace accu {

constexpr int kXcd = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin over them

// Map a dispatch-order block id to a logical block id so that each XCD works on one
// contiguous range of atoms (its own L2 then holds one spatial window).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int per = nb / kXcd, rem = nb % kXcd;
  const int x = b % kXcd, k = b / kXcd;
  return x * per + (x < rem ? x : rem) + k;
}

struct Box {
  double L[3];
  double inv[3];
  int periodic;
};

// numpy's round-half-even minimum image d -= L * round(d / L) with the true
// division (bit-for-bit the reference expression; no contraction).
__device__ __forceinline__ double minimg_exact(double d, double L) {
#pragma clang fp contract(off)
  return d - L * rint(d / L);
}

// Same image choice with a reciprocal multiply: identical for every |d_min| < L/2
// (only d/L within an ulp of a half-integer can round differently).
__device__ __forceinline__ double minimg_fast(double d, double L, double inv) {
  return fma(-L, rint(d * inv), d);
}

// numpy's x % L for float64 (the periodic wrap of system.py:127-129): fmod, moved into
// [0, L) by adding L when the signs differ, and +0.0 for an exact multiple.
__device__ __forceinline__ double np_mod2(double x, double L) {
  double r = fmod(x, L);
  if (r != 0.0) {
    if ((r < 0.0) != (L < 0.0)) r += L;
  } else {
    r = copysign(0.0, L);
  }
  return r;
}

// |d|^2 summed as numpy's einsum("ij,ij->i") sums a 3-vector: (x^2 + z^2) + y^2.
__device__ __forceinline__ double r2_np(double x, double y, double z) {
#pragma clang fp contract(off)
  return (x * x + z * z) + y * y;
}
__device__ __forceinline__ float r2_npf(float x, float y, float z) {
#pragma clang fp contract(off)
  return (x * x + z * z) + y * y;
}

// 1/x in float64 from v_rcp_f64 plus two Newton steps (within 1 ulp).
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// 1/sqrt(x) in float64 from v_rsq_f64 plus two Newton steps.
__device__ __forceinline__ double rsqrt64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}

// One DPP move of a double (each 32-bit half moved alike); lanes of rows outside
// ROW_MASK get 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROW_MASK, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Sums over the 4 lanes of each quad (every lane of the quad gets it) in a fixed order, by
// DPP quad permutes. All 4 lanes of the quad must be active.
__device__ __forceinline__ double quad_sum(double v) {
  v += dpp_f64<0xB1, 0xf>(v);  // quad_perm [1, 0, 3, 2]
  v += dpp_f64<0x4E, 0xf>(v);  // quad_perm [2, 3, 0, 1]
  return v;
}
__device__ __forceinline__ float quad_sumf(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xf, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xf, 0xf, false));
  return v;
}

// Sum over the 64 lanes of a wave, in a fixed order, returned to every lane. All lanes
// must be active. Cross-lane moves are DPP (quad permutes, row rotations, row
// broadcasts), so the sum takes no LDS instructions: quads, then 16-lane rows, then
// row 0 + row 1 and row 2 + row 3, then all four, which lane 63 holds.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_f64<0x4E, 0xf>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_f64<0x124, 0xf>(v);  // row_ror:4
  v += dpp_f64<0x128, 0xf>(v);  // row_ror:8
  v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Deterministic sum of `nv` values per thread over the block; result valid in wave 0.
// A wave's NV sums sit together (lds[w * NV + k]), so its stores take one base address
// and immediate offsets (a kernel that sums inside a loop keeps one address live, not NV).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* [NV * nwaves] */) {
  static_assert(NV <= 64, "one lane of wave 0 per value");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[w * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x < 64) {  // wave 0: lane k sums value k over the waves, in wave order
    double s = 0.0;
    if (lane < NV)
      for (int q = 0; q < nw; ++q) s += lds[q * NV + lane];
    const unsigned long long b = __builtin_bit_cast(unsigned long long, s);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, k);
      const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), k);
      v[k] = __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
    }
  }
  __syncthreads();
}

// Grid-wide deterministic reduction: every block stores its NV partials, the last
// block to arrive sums them in block order and writes out[0..NV). `counter` must be
// zero on entry and is left zero. Blocks must be 1-D; gridDim.x partial slots.
template <int NV>
__device__ void grid_sum(double (&v)[NV], double* partials, unsigned* counter, double* out,
                         double* lds) {
  block_sum<NV>(v, lds);
  __shared__ int am_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)blockIdx.x * NV + k] = v[k];
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    am_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return;
  __threadfence();
  const int nb = gridDim.x;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  // fixed-order strided partial sums, then the fixed-order block sum
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      acc[k] += __hip_atomic_load(&partials[(size_t)b * NV + k], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  block_sum<NV>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = acc[k];
    *counter = 0u;
  }
}

// Scratch of the handle-free entry points (reduction partials and counter, the Ewald
// reciprocal buffers), one per (device, stream): calls on different streams or devices never
// share it, and calls on one stream use it in stream order. Calls on one stream from several
// host threads at once must be serialised by the caller.
struct Workspace {
  int dev = 0;
  hipStream_t st = nullptr;
  double* partials = nullptr;
  size_t cap = 0;  // doubles
  unsigned* counter = nullptr;
  double* out = nullptr;  // 16 doubles
  // Ewald reciprocal sum: partial structure factors, weighted S(k), weights, columns
  double2* ew_part = nullptr;
  double2* ew_A = nullptr;
  double* ew_kw = nullptr;
  int* ew_col = nullptr;
  long long* ew_off = nullptr;
  size_t ew_part_cap = 0, ew_A_cap = 0, ew_kw_cap = 0, ew_col_cap = 0, ew_off_cap = 0;  // elements
  int ensure(size_t ndoubles);
  // device buffer p grown to at least n elements of T (its contents are not kept); the
  // old one is freed after the stream's earlier work
  template <typename T>
  int grow(T*& p, size_t& cap_n, size_t n) {
    if (n <= cap_n && p) return 0;
    if (p) {
      if (hipStreamSynchronize(st) != hipSuccess || hipFree(p) != hipSuccess) return grow_error();
      p = nullptr;
    }
    if (hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T)) != hipSuccess) return grow_error();
    cap_n = n;
    return 0;
  }
  int grow_error();
};

// The workspace of calls on `stream` on the current device (created at its first use).
Workspace* workspace_for(hipStream_t stream);

}  // namespace accu
