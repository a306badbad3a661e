// Barnes-Hut gravity (tree.py:19-78): the octree is built on the GPU and walked on
// the GPU; nothing of it passes through the host but the node count.
//
// The tree is the reference's (tree.py:30-53), not merely a similar one:
//   - root cube from the bounding box, half = max extent / 2 * (1 + 1e-9), at least
//     1e-300, center = (lo + hi) / 2 (tree.py:51-53);
//   - octant o = (x > cx) + 2 (y > cy) + 4 (z > cz) against the node's center, the
//     child center = center +- half / 2, the child half = half / 2 (tree.py:37-42), in
//     the same float64 operations, so every comparison sees the reference's center;
//   - a node splits while it holds more than leaf_size particles and depth < 48 (36);
//   - children in octant order 0..7, empty octants skipped (38-40).
//
// Build, with no recursion and no host round trip per level:
//   1. bounding box (min/max, exact) and the root cube, on the device;
//   2. each particle's octant path from the root: 21 levels in one 63-bit key (3 bits
//      per level, level 0 highest); when a group of leaf_size + 1 particles shares all
//      21 levels, levels 21-41 and 42-47 go into two more keys (the deep pass);
//   3. radix sort of the keys (stable; ties keep the particle index order), so every
//      node's particles are one contiguous run of the sorted order;
//   4. with lcp(a, b) = the number of leading levels two sorted keys share, the node of
//      depth d holding particle k has more than leaf_size particles iff some window
//      j .. j + leaf_size around k has lcp(j, j + leaf_size) >= d. So particle k's leaf
//      depth is L(k) = min(max window lcp + 1, 48), and the nodes that START at k are
//      the depths lcp(k-1, k) + 1 .. L(k). Ordered by (start, depth) the nodes are in
//      depth-first pre-order: a scan of the per-particle node counts numbers them;
//   5. per node: cube, particle run (binary search for its end), next = first node
//      after the subtree (the first node starting at the run's end), parent, leaf flag;
//   6. mass moments bottom-up in one launch: each leaf sums its particles, and the last
//      child to finish a parent (an arrival counter per node) sums the parent from its
//      children in octant order; quadrupoles move to the parent's center of mass by the
//      parallel-axis rule.
// The walk visits the nodes in pre-order with no stack: a node's first child is
// node + 1 and next[node] skips the subtree. It accepts the reference's nodes (target
// outside the node box and (2 half)^2 < theta^2 r^2 with softened r^2, tree.py:67-70)
// and sums leaves directly with the self pair excluded (58-66); only the summation
// order differs. Quadrupole terms (traceless Q about the center of mass) extend the
// reference's monopole expansion when asked for.
#include "common.h"
#include "../../include/accusim_hip.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cstdint>

namespace accu {
namespace {

typedef unsigned long long u64;
constexpr int kLv1 = 21, kLv2 = 21, kLv3 = 6, kMaxDepth = 48;
constexpr int kThreads = 256;

// tree.py:37: octant bits (pos > center) . [1, 2, 4]
__device__ __forceinline__ int octant(double x, double y, double z, double cx, double cy, double cz) {
  return (x > cx ? 1 : 0) | (y > cy ? 2 : 0) | (z > cz ? 4 : 0);
}

// tree.py:41-42: child center = center + sign * half / 2, child half = half / 2 (both
// exact but the one rounding of the addition, as numpy does it)
__device__ __forceinline__ void descend(int o, double& cx, double& cy, double& cz, double& h) {
  const double q = h * 0.5;
  cx = (o & 1) ? cx + q : cx - q;
  cy = (o & 2) ? cy + q : cy - q;
  cz = (o & 4) ? cz + q : cz - q;
  h = q;
}

__device__ __forceinline__ int octant_at(u64 a, u64 b, u64 c, int l) {
  if (l < kLv1) return (int)((a >> (3 * (kLv1 - 1 - l))) & 7);
  if (l < kLv1 + kLv2) return (int)((b >> (3 * (kLv1 + kLv2 - 1 - l))) & 7);
  return (int)((c >> (3 * (kMaxDepth - 1 - l))) & 7);
}

// number of leading levels the sorted keys i and j share (NW key words per particle)
template <int NW>
__device__ __forceinline__ int lcp(const u64* __restrict__ K1, const u64* __restrict__ K2,
                                   const u64* __restrict__ K3, int i, int j) {
  u64 x = K1[i] ^ K1[j];
  if (x) return (__clzll((long long)x) - 1) / 3;
  if (NW == 1) return kLv1;
  x = K2[i] ^ K2[j];
  if (x) return kLv1 + (__clzll((long long)x) - 1) / 3;
  x = K3[i] ^ K3[j];
  if (x) return kLv1 + kLv2 + (__clzll((long long)x) - (64 - 3 * kLv3)) / 3;
  return kMaxDepth;
}

// ---- 1. bounding box and root cube (tree.py:51-53) --------------------------------

__device__ __forceinline__ void minmax_block(double (&lo)[3], double (&hi)[3], double* lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double a = __shfl_xor(lo[d], o, 64), b = __shfl_xor(hi[d], o, 64);
      lo[d] = a < lo[d] ? a : lo[d];
      hi[d] = b > hi[d] ? b : hi[d];
    }
  if (lane == 0)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lds[6 * w + d] = lo[d];
      lds[6 * w + 3 + d] = hi[d];
    }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 1; q < kThreads / 64; ++q)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        lo[d] = lds[6 * q + d] < lo[d] ? lds[6 * q + d] : lo[d];
        hi[d] = lds[6 * q + 3 + d] > hi[d] ? lds[6 * q + 3 + d] : hi[d];
      }
  __syncthreads();
}

__global__ void __launch_bounds__(kThreads) k_bbox(int n, const double* __restrict__ pos, double* part,
                                                   unsigned* counter, double4* root, int* info) {
  __shared__ double lds[6 * (kThreads / 64)];
  __shared__ int am_last;
  double lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = pos[d];
    hi[d] = pos[d];
  }
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double v = pos[3 * (size_t)i + d];
      lo[d] = v < lo[d] ? v : lo[d];
      hi[d] = v > hi[d] ? v : hi[d];
    }
  minmax_block(lo, hi, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      part[6 * (size_t)blockIdx.x + d] = lo[d];
      part[6 * (size_t)blockIdx.x + 3 + d] = hi[d];
    }
    __threadfence();
    am_last = atomicAdd(counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!am_last) return;
  __threadfence();
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = pos[d];
    hi[d] = pos[d];
  }
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kThreads)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double a = __hip_atomic_load(&part[6 * (size_t)b + d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const double c = __hip_atomic_load(&part[6 * (size_t)b + 3 + d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      lo[d] = a < lo[d] ? a : lo[d];
      hi[d] = c > hi[d] ? c : hi[d];
    }
  minmax_block(lo, hi, lds);
  if (threadIdx.x == 0) {
#pragma clang fp contract(off)
    double ext = hi[0] - lo[0];
    ext = (hi[1] - lo[1]) > ext ? (hi[1] - lo[1]) : ext;
    ext = (hi[2] - lo[2]) > ext ? (hi[2] - lo[2]) : ext;
    double half = ext / 2 * (1 + 1e-9);
    half = half >= 1e-300 ? half : 1e-300;
    *root = make_double4((lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2, half);
    for (int q = 0; q < 8; ++q) info[q] = 0;
    *counter = 0u;
  }
}

// ---- 2. octant-path keys ----------------------------------------------------------

template <int NW>
__global__ void __launch_bounds__(kThreads) k_keys(int n, const double* __restrict__ pos,
                                                   const double4* __restrict__ root, u64* __restrict__ K1,
                                                   u64* __restrict__ K2, u64* __restrict__ K3, int* __restrict__ idx) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const double x = pos[3 * (size_t)i], y = pos[3 * (size_t)i + 1], z = pos[3 * (size_t)i + 2];
  const double4 r = *root;
  double cx = r.x, cy = r.y, cz = r.z, h = r.w;
  u64 a = 0;
  for (int l = 0; l < kLv1; ++l) {
    const int o = octant(x, y, z, cx, cy, cz);
    a = (a << 3) | (u64)o;
    descend(o, cx, cy, cz, h);
  }
  K1[i] = a;
  if (NW == 3) {
    u64 b = 0, c = 0;
    for (int l = 0; l < kLv2; ++l) {
      const int o = octant(x, y, z, cx, cy, cz);
      b = (b << 3) | (u64)o;
      descend(o, cx, cy, cz, h);
    }
    for (int l = 0; l < kLv3; ++l) {
      const int o = octant(x, y, z, cx, cy, cz);
      c = (c << 3) | (u64)o;
      descend(o, cx, cy, cz, h);
    }
    K2[i] = b;
    K3[i] = c;
  }
  idx[i] = i;
}

__global__ void k_gather_u64(int n, const u64* __restrict__ src, const int* __restrict__ perm, u64* __restrict__ dst) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

// ---- 4. nodes per particle and the particles in tree order --------------------------

template <int NW>
__global__ void __launch_bounds__(kThreads) k_count(int n, int leaf, const u64* __restrict__ K1,
                                                    const u64* __restrict__ K2, const u64* __restrict__ K3,
                                                    const int* __restrict__ perm, const double* __restrict__ pos,
                                                    const double* __restrict__ m, int* __restrict__ cnt,
                                                    unsigned char* __restrict__ d0s, unsigned char* __restrict__ Ls,
                                                    double4* __restrict__ ppos, int* info) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k == 0) cnt[n] = 0;  // the scan's extra slot gives the node total
  if (k >= n) return;
  const int lp = k == 0 ? -1 : lcp<NW>(K1, K2, K3, k - 1, k);
  int W = -1;
  for (int j = k - leaf < 0 ? 0 : k - leaf; j <= k && j + leaf < n; ++j) {
    const int v = lcp<NW>(K1, K2, K3, j, j + leaf);
    W = v > W ? v : W;
  }
  if (NW == 1 && W == kLv1) info[1] = 1;  // more than leaf_size particles share 21 levels
  const int L = W < 0 ? 0 : (W + 1 < kMaxDepth ? W + 1 : kMaxDepth);
  const int d0 = lp + 1;
  cnt[k] = L >= d0 ? L - d0 + 1 : 0;
  d0s[k] = (unsigned char)d0;
  Ls[k] = (unsigned char)L;
  if (L > 0) atomicMax(&info[2], L);
  const int p = perm[k];
  ppos[k] = make_double4(pos[3 * (size_t)p], pos[3 * (size_t)p + 1], pos[3 * (size_t)p + 2], m[p]);
}

__global__ void k_total(int n, const int* __restrict__ base, int* info) { info[0] = base[n]; }

__global__ void k_zero2(int nn, int* __restrict__ a, int* __restrict__ b) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < nn) {
    a[i] = 0;
    b[i] = 0;
  }
}

// ---- 5. node records --------------------------------------------------------------

template <int NW>
__global__ void __launch_bounds__(kThreads) k_nodes(
    int n, int nn, const u64* __restrict__ K1, const u64* __restrict__ K2, const u64* __restrict__ K3,
    const int* __restrict__ cnt, const int* __restrict__ base, const unsigned char* __restrict__ d0s,
    const unsigned char* __restrict__ Ls, const double4* __restrict__ root, double4* __restrict__ cube,
    int* __restrict__ next, int* __restrict__ leafv, int* __restrict__ lstart, int* __restrict__ lcount,
    int* __restrict__ parent, unsigned char* __restrict__ depth, int* nchild, int* info) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n || cnt[k] == 0) return;
  const int d0 = d0s[k], L = Ls[k], b = base[k];
  const u64 a1 = K1[k], a2 = NW == 3 ? K2[k] : 0, a3 = NW == 3 ? K3[k] : 0;
  const double4 r = *root;
  double cx = r.x, cy = r.y, cz = r.z, h = r.w;
  for (int l = 0; l < d0; ++l) descend(octant_at(a1, a2, a3, l), cx, cy, cz, h);
  int hi = n;
  for (int d = d0; d <= L; ++d) {
    const int id = b + d - d0;
    // end of the run sharing d levels with k: first e in (k, hi) with lcp(k, e) < d
    int lo = k + 1, up = hi;
    while (lo < up) {
      const int mid = (lo + up) >> 1;
      if (lcp<NW>(K1, K2, K3, k, mid) >= d) lo = mid + 1;
      else up = mid;
    }
    const int end = lo;
    hi = end;
    int par;
    if (d > d0) {
      par = id - 1;
    } else if (k == 0) {
      par = -1;
    } else {
      // the depth d-1 node holding k starts at the first s <= k sharing d-1 levels with k
      int s0 = 0, s1 = k;
      while (s0 < s1) {
        const int mid = (s0 + s1) >> 1;
        if (lcp<NW>(K1, K2, K3, mid, k) >= d - 1) s1 = mid;
        else s0 = mid + 1;
      }
      par = base[s0] + (d - 1) - d0s[s0];
    }
    cube[id] = make_double4(cx, cy, cz, h);
    next[id] = end < n ? base[end] : nn;
    leafv[id] = d == L;
    lstart[id] = k;
    lcount[id] = end - k;
    parent[id] = par;
    depth[id] = (unsigned char)d;
    if (par >= 0) atomicAdd(&nchild[par], 1);
    if (d == L) atomicAdd(&info[3], 1);
    if (d < L) descend(octant_at(a1, a2, a3, d), cx, cy, cz, h);
  }
}

// ---- 6. mass moments, bottom-up -----------------------------------------------------

template <bool QUAD>
__global__ void __launch_bounds__(kThreads) k_moments(int nn, const int* __restrict__ leafv,
                                                      const int* __restrict__ lstart, const int* __restrict__ lcount,
                                                      const int* __restrict__ next, const int* __restrict__ parent,
                                                      const int* __restrict__ nchild, const double4* __restrict__ cube,
                                                      const double4* __restrict__ ppos, double4* mom, double4* com,
                                                      double* quad, int* arrive) {
  const int id = blockIdx.x * kThreads + threadIdx.x;
  if (id >= nn || !leafv[id]) return;
  {
    const int a = lstart[id], c = lcount[id];
    double M = 0, Sx = 0, Sy = 0, Sz = 0;
    for (int s = a; s < a + c; ++s) {
      const double4 p = ppos[s];
      M += p.w;
      Sx += p.w * p.x;
      Sy += p.w * p.y;
      Sz += p.w * p.z;
    }
    const double4 cb = cube[id];
    const double4 cm = M > 0 ? make_double4(Sx / M, Sy / M, Sz / M, M) : make_double4(cb.x, cb.y, cb.z, M);
    mom[id] = make_double4(Sx, Sy, Sz, M);
    com[id] = cm;
    if (QUAD) {
      double q[6] = {0, 0, 0, 0, 0, 0};
      for (int s = a; s < a + c; ++s) {
        const double4 p = ppos[s];
        const double x = p.x - cm.x, y = p.y - cm.y, z = p.z - cm.z, r2 = x * x + y * y + z * z;
        q[0] += p.w * (3 * x * x - r2);
        q[1] += p.w * (3 * y * y - r2);
        q[2] += p.w * (3 * z * z - r2);
        q[3] += p.w * 3 * x * y;
        q[4] += p.w * 3 * x * z;
        q[5] += p.w * 3 * y * z;
      }
      for (int t = 0; t < 6; ++t) quad[6 * (size_t)id + t] = q[t];
    }
  }
  int node = id;
  for (;;) {
    const int p = parent[node];
    if (p < 0) break;
    __threadfence();
    const int prev = __hip_atomic_fetch_add(&arrive[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev != nchild[p] - 1) break;  // a sibling still runs: the last one to finish goes on
    const int stop = next[p];
    double M = 0, Sx = 0, Sy = 0, Sz = 0;
    for (int c = p + 1; c != stop; c = next[c]) {
      const double4 s = mom[c];
      Sx += s.x;
      Sy += s.y;
      Sz += s.z;
      M += s.w;
    }
    const double4 cb = cube[p];
    const double4 cm = M > 0 ? make_double4(Sx / M, Sy / M, Sz / M, M) : make_double4(cb.x, cb.y, cb.z, M);
    mom[p] = make_double4(Sx, Sy, Sz, M);
    com[p] = cm;
    if (QUAD) {
      double q[6] = {0, 0, 0, 0, 0, 0};
      for (int c = p + 1; c != stop; c = next[c]) {
        const double4 cc = com[c];
        const double x = cc.x - cm.x, y = cc.y - cm.y, z = cc.z - cm.z, r2 = x * x + y * y + z * z;
        const double* Q = quad + 6 * (size_t)c;
        q[0] += Q[0] + cc.w * (3 * x * x - r2);
        q[1] += Q[1] + cc.w * (3 * y * y - r2);
        q[2] += Q[2] + cc.w * (3 * z * z - r2);
        q[3] += Q[3] + cc.w * 3 * x * y;
        q[4] += Q[4] + cc.w * 3 * x * z;
        q[5] += Q[5] + cc.w * 3 * y * z;
      }
      for (int t = 0; t < 6; ++t) quad[6 * (size_t)p + t] = q[t];
    }
    node = p;
  }
}

// ---- the walk (tree.py:54-78) -------------------------------------------------------
//
// Each lane walks for one target (particles in tree order, so a wave's 64 targets are
// neighbours). The wave steps through the tree together: the node is the same for every
// lane, so its data loads are shared, and a lane that accepted a node (or finished a
// leaf) skips that node's subtree, [node, next[node]), while the others descend. The wave
// moves to the first child when some lane still needs it, else past the subtree. Every
// lane therefore meets exactly the nodes and leaves of its own stackless walk, in the same
// pre-order, and sums the same terms in the same order as a walk of its own would.

template <bool QUAD>
__global__ void __launch_bounds__(kThreads) k_bh_walk(int n, int nnodes, const double4* __restrict__ com,
                                                      const double4* __restrict__ cube, const int* __restrict__ next,
                                                      const int* __restrict__ leaf, const int* __restrict__ lstart,
                                                      const int* __restrict__ lcount, const double* __restrict__ quad,
                                                      const int* __restrict__ perm, const double4* __restrict__ ppos,
                                                      double G, double eps2, double th2, double* __restrict__ F,
                                                      double* partials, unsigned* counter, double* Uout) {
  __shared__ double red[16];
  const int k = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  double U[1] = {0.0};
  const bool on = k < n;
  const double4 p = on ? ppos[k] : make_double4(0.0, 0.0, 0.0, 0.0);
  double ax = 0, ay = 0, az = 0, phi = 0;
  int skip = on ? 0 : nnodes;  // this lane sits out the nodes below skip (an accepted subtree)
  int node = 0;                // the wave's node (uniform)
  while (node < nnodes) {
    const bool act = node >= skip;
    const int nx = next[node];
    if (leaf[node]) {
      if (act) {
        const int a = lstart[node], b = a + lcount[node];
        for (int s = a; s < b; ++s) {
          if (s == k) continue;
          const double4 q = ppos[s];
          const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
          const double r2 = r2_np(dx, dy, dz) + eps2;
          const double inv = rsqrt64(r2);
          const double mi3 = q.w * inv * inv * inv;
          ax = fma(mi3, dx, ax);
          ay = fma(mi3, dy, ay);
          az = fma(mi3, dz, az);
          phi = fma(-q.w, inv, phi);
        }
      }
      node = nx;
      continue;
    }
    bool descend = false;
    if (act) {
      const double4 c = com[node];
      const double4 cb = cube[node];
      const double dx = c.x - p.x, dy = c.y - p.y, dz = c.z - p.z;
      const double r2 = r2_np(dx, dy, dz) + eps2;
      const bool outside = fabs(p.x - cb.x) > cb.w || fabs(p.y - cb.y) > cb.w || fabs(p.z - cb.z) > cb.w;
      const double s2 = (2 * cb.w) * (2 * cb.w);
      if (outside && s2 < th2 * r2) {
        const double inv = rsqrt64(r2);
        const double mi3 = c.w * inv * inv * inv;
        ax = fma(mi3, dx, ax);
        ay = fma(mi3, dy, ay);
        az = fma(mi3, dz, az);
        phi = fma(-c.w, inv, phi);
        if (QUAD) {
          // R = target - com = -d; a += Q R / r^5 - 5/2 (R Q R) R / r^7; phi -= 1/2 (R Q R) / r^5
          const double* Q = quad + 6 * (size_t)node;
          const double Rx = -dx, Ry = -dy, Rz = -dz;
          const double QRx = Q[0] * Rx + Q[3] * Ry + Q[4] * Rz;
          const double QRy = Q[3] * Rx + Q[1] * Ry + Q[5] * Rz;
          const double QRz = Q[4] * Rx + Q[5] * Ry + Q[2] * Rz;
          const double RQR = Rx * QRx + Ry * QRy + Rz * QRz;
          const double inv2 = inv * inv, inv5 = inv2 * inv2 * inv;
          const double inv7 = inv5 * inv2;
          ax += QRx * inv5 - 2.5 * RQR * Rx * inv7;
          ay += QRy * inv5 - 2.5 * RQR * Ry * inv7;
          az += QRz * inv5 - 2.5 * RQR * Rz * inv7;
          phi -= 0.5 * RQR * inv5;
        }
        skip = nx;
      } else {
        descend = true;
      }
    }
    node = __any(descend) ? node + 1 : nx;  // first child, or past the subtree
  }
  if (on) {
    const int i = perm[k];
    F[3 * (size_t)i] = G * p.w * ax;
    F[3 * (size_t)i + 1] = G * p.w * ay;
    F[3 * (size_t)i + 2] = G * p.w * az;
    U[0] = 0.5 * G * p.w * phi;
  }
  if (Uout) grid_sum<1>(U, partials, counter, Uout, red);
}

template <typename T>
int grow(T** p, size_t need, size_t& cap) {
  if (need <= cap && *p) return ACCU_OK;
  if (*p) HIP_TRY(hipFree(*p));
  *p = nullptr;
  const size_t c = need + need / 4 + 64;
  HIP_TRY(hipMalloc(p, c * sizeof(T)));
  cap = c;
  return ACCU_OK;
}

}  // namespace
}  // namespace accu

using namespace accu;

struct accu_bh {
  int n = 0, nn = 0, leaf_size = 8, quad = 0, deep = 0, max_depth = 0, built = 0;
  // particles (capacity pcap)
  size_t pcap = 0;
  u64 *kin = nullptr, *ks1 = nullptr;
  int *idx = nullptr, *perm = nullptr, *cnt = nullptr, *base = nullptr;
  unsigned char *d0s = nullptr, *Ls = nullptr;
  double4* ppos = nullptr;
  // deep pass (capacity dcap)
  size_t dcap = 0;
  u64 *K1o = nullptr, *K2o = nullptr, *K3o = nullptr, *ks2 = nullptr, *ks3 = nullptr, *tk = nullptr;
  int *pa = nullptr, *pb = nullptr;
  // nodes (capacity ncap)
  size_t ncap = 0;
  double4 *cube = nullptr, *com = nullptr, *mom = nullptr;
  double* quadv = nullptr;
  int *next = nullptr, *leaf = nullptr, *lstart = nullptr, *lcount = nullptr, *parent = nullptr, *nchild = nullptr,
      *arrive = nullptr;
  unsigned char* depth = nullptr;
  // reductions, root, info, scratch of the sorts and scans
  double* part = nullptr;  // bounding-box partials [1024][6], then walk energy partials
  size_t part_cap = 0;
  unsigned* counter = nullptr;  // [0]: bounding box, [1]: walk energy
  double4* root = nullptr;
  int* info = nullptr;       // device [8]: nodes, deep, max depth, leaves
  int* info_h = nullptr;     // pinned host [8]
  void* tmp = nullptr;
  size_t tmp_cap = 0;
};

namespace {

int bh_particles(accu_bh* h, int n) {
  if ((size_t)n + 1 <= h->pcap) return ACCU_OK;
  const size_t c = (size_t)n + n / 4 + 64;
  void** bufs[] = {(void**)&h->kin, (void**)&h->ks1, (void**)&h->idx, (void**)&h->perm, (void**)&h->cnt,
                   (void**)&h->base, (void**)&h->d0s, (void**)&h->Ls, (void**)&h->ppos};
  const size_t sz[] = {8, 8, 4, 4, 4, 4, 1, 1, 32};
  for (int b = 0; b < 9; ++b) {
    if (*bufs[b]) HIP_TRY(hipFree(*bufs[b]));
    *bufs[b] = nullptr;
    HIP_TRY(hipMalloc(bufs[b], c * sz[b]));
  }
  h->pcap = c;
  return ACCU_OK;
}

int bh_deep(accu_bh* h, int n) {
  if ((size_t)n <= h->dcap) return ACCU_OK;
  const size_t c = (size_t)n + n / 4 + 64;
  void** bufs[] = {(void**)&h->K1o, (void**)&h->K2o, (void**)&h->K3o, (void**)&h->ks2,
                   (void**)&h->ks3, (void**)&h->tk,  (void**)&h->pa,  (void**)&h->pb};
  const size_t sz[] = {8, 8, 8, 8, 8, 8, 4, 4};
  for (int b = 0; b < 8; ++b) {
    if (*bufs[b]) HIP_TRY(hipFree(*bufs[b]));
    *bufs[b] = nullptr;
    HIP_TRY(hipMalloc(bufs[b], c * sz[b]));
  }
  h->dcap = c;
  return ACCU_OK;
}

int bh_nodes(accu_bh* h, int nn) {
  if ((size_t)nn <= h->ncap) return ACCU_OK;
  const size_t c = (size_t)nn + nn / 4 + 64;
  void** bufs[] = {(void**)&h->cube,   (void**)&h->com,    (void**)&h->mom,    (void**)&h->quadv,
                   (void**)&h->next,   (void**)&h->leaf,   (void**)&h->lstart, (void**)&h->lcount,
                   (void**)&h->parent, (void**)&h->nchild, (void**)&h->arrive, (void**)&h->depth};
  const size_t sz[] = {32, 32, 32, 48, 4, 4, 4, 4, 4, 4, 4, 1};
  for (int b = 0; b < 12; ++b) {
    if (*bufs[b]) HIP_TRY(hipFree(*bufs[b]));
    *bufs[b] = nullptr;
    HIP_TRY(hipMalloc(bufs[b], c * sz[b]));
  }
  h->ncap = c;
  return ACCU_OK;
}

int bh_tmp(accu_bh* h, size_t bytes) {
  if (bytes <= h->tmp_cap && h->tmp) return ACCU_OK;
  if (h->tmp) HIP_TRY(hipFree(h->tmp));
  h->tmp = nullptr;
  HIP_TRY(hipMalloc(&h->tmp, bytes + 4096));
  h->tmp_cap = bytes + 4096;
  return ACCU_OK;
}

int sort_pairs(accu_bh* h, const u64* kin, u64* kout, const int* vin, int* vout, int n, int bits, hipStream_t st) {
  size_t bytes = 0;
  HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (unsigned)n, 0, bits, st));
  int rc = bh_tmp(h, bytes);
  if (rc) return rc;
  bytes = h->tmp_cap;
  HIP_TRY(rocprim::radix_sort_pairs(h->tmp, bytes, kin, kout, vin, vout, (unsigned)n, 0, bits, st));
  return ACCU_OK;
}

// node counts per particle, their scan and the total, read back to the host
template <int NW>
int count_nodes(accu_bh* h, const double* pos, const double* m, int n, const u64* K1, const u64* K2, const u64* K3,
                hipStream_t st) {
  const int nb = (n + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_count<NW>, dim3(nb), dim3(kThreads), 0, st, n, h->leaf_size, K1, K2, K3, h->perm, pos, m,
                     h->cnt, h->d0s, h->Ls, h->ppos, h->info);
  LAUNCH_CHECK();
  size_t bytes = 0;
  HIP_TRY(rocprim::exclusive_scan(nullptr, bytes, h->cnt, h->base, 0, (size_t)n + 1, rocprim::plus<int>(), st));
  int rc = bh_tmp(h, bytes);
  if (rc) return rc;
  bytes = h->tmp_cap;
  HIP_TRY(rocprim::exclusive_scan(h->tmp, bytes, h->cnt, h->base, 0, (size_t)n + 1, rocprim::plus<int>(), st));
  hipLaunchKernelGGL(k_total, dim3(1), dim3(1), 0, st, n, h->base, h->info);
  LAUNCH_CHECK();
  HIP_TRY(hipMemcpyAsync(h->info_h, h->info, 8 * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return ACCU_OK;
}

template <int NW>
int make_nodes(accu_bh* h, int n, const u64* K1, const u64* K2, const u64* K3, hipStream_t st) {
  const int nn = h->nn;
  int rc = bh_nodes(h, nn);
  if (rc) return rc;
  const int nbn = (nn + kThreads - 1) / kThreads, nb = (n + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_zero2, dim3(nbn), dim3(kThreads), 0, st, nn, h->nchild, h->arrive);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_nodes<NW>, dim3(nb), dim3(kThreads), 0, st, n, nn, K1, K2, K3, h->cnt, h->base, h->d0s,
                     h->Ls, h->root, h->cube, h->next, h->leaf, h->lstart, h->lcount, h->parent, h->depth, h->nchild,
                     h->info);
  LAUNCH_CHECK();
  if (h->quad)
    hipLaunchKernelGGL(k_moments<true>, dim3(nbn), dim3(kThreads), 0, st, nn, h->leaf, h->lstart, h->lcount, h->next,
                       h->parent, h->nchild, h->cube, h->ppos, h->mom, h->com, h->quadv, h->arrive);
  else
    hipLaunchKernelGGL(k_moments<false>, dim3(nbn), dim3(kThreads), 0, st, nn, h->leaf, h->lstart, h->lcount, h->next,
                       h->parent, h->nchild, h->cube, h->ppos, h->mom, h->com, h->quadv, h->arrive);
  LAUNCH_CHECK();
  return ACCU_OK;
}

}  // namespace

extern "C" {

int accu_bh_create(accu_bh** out) {
  ARG_CHECK(out, "null handle pointer");
  accu_bh* h = new accu_bh();
  if (hipMalloc(&h->counter, 64) != hipSuccess || hipMemset(h->counter, 0, 64) != hipSuccess ||
      hipMalloc(&h->root, sizeof(double4)) != hipSuccess || hipMalloc(&h->info, 8 * sizeof(int)) != hipSuccess ||
      hipHostMalloc(&h->info_h, 8 * sizeof(int)) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {  // the counter fill completes before any caller stream runs
    accu_bh_destroy(h);
    set_error("accu_bh_create: device allocation failed");
    return ACCU_ERUNTIME;
  }
  *out = h;
  return ACCU_OK;
}

int accu_bh_destroy(accu_bh* h) {
  if (!h) return ACCU_OK;
  void* dev[] = {h->kin,  h->ks1,    h->idx,  h->perm,   h->cnt,    h->base,   h->d0s,    h->Ls,   h->ppos,
                 h->K1o,  h->K2o,    h->K3o,  h->ks2,    h->ks3,    h->tk,     h->pa,     h->pb,   h->cube,
                 h->com,  h->mom,    h->quadv, h->next,  h->leaf,   h->lstart, h->lcount, h->parent,
                 h->nchild, h->arrive, h->depth, h->part, h->counter, h->root, h->info,   h->tmp};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (h->info_h) (void)hipHostFree(h->info_h);
  delete h;
  return ACCU_OK;
}

int accu_bh_build(accu_bh* h, const double* pos, const double* m, int n, int leaf_size, int quadrupole, void* stream) {
  ARG_CHECK(h && pos && m, "null argument");
  ARG_CHECK(n >= 1, "Barnes-Hut needs at least one particle");
  // node ids are int32 and a tree holds at most 49 nodes per particle (a chain to depth 48)
  ARG_CHECK(n <= (1 << 25), "Barnes-Hut takes at most 33,554,432 particles per tree");
  ARG_CHECK(leaf_size >= 1 && leaf_size < 1024, "leaf_size must be in [1, 1024)");
  hipStream_t st = (hipStream_t)stream;
  h->built = 0;
  h->n = n;
  h->leaf_size = leaf_size;
  h->quad = quadrupole != 0;
  int rc;
  if ((rc = bh_particles(h, n))) return rc;
  const int nb = (n + kThreads - 1) / kThreads;
  const int nbb = nb < 1024 ? nb : 1024;
  if ((rc = grow(&h->part, (size_t)6 * 1024 + 2 * (size_t)nb, h->part_cap))) return rc;
  hipLaunchKernelGGL(k_bbox, dim3(nbb), dim3(kThreads), 0, st, n, pos, h->part, h->counter, h->root, h->info);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_keys<1>, dim3(nb), dim3(kThreads), 0, st, n, pos, h->root, h->kin, nullptr, nullptr, h->idx);
  LAUNCH_CHECK();
  if ((rc = sort_pairs(h, h->kin, h->ks1, h->idx, h->perm, n, 3 * kLv1, st))) return rc;
  if ((rc = count_nodes<1>(h, pos, m, n, h->ks1, nullptr, nullptr, st))) return rc;
  h->deep = h->info_h[1];
  if (h->deep) {
    // more than leaf_size particles share 21 levels: sort by all 48 levels (least
    // significant key first; each pass is stable, so ties keep the index order)
    if ((rc = bh_deep(h, n))) return rc;
    hipLaunchKernelGGL(k_keys<3>, dim3(nb), dim3(kThreads), 0, st, n, pos, h->root, h->K1o, h->K2o, h->K3o, h->idx);
    LAUNCH_CHECK();
    if ((rc = sort_pairs(h, h->K3o, h->tk, h->idx, h->pa, n, 3 * kLv3, st))) return rc;
    hipLaunchKernelGGL(k_gather_u64, dim3(nb), dim3(kThreads), 0, st, n, h->K2o, h->pa, h->kin);
    LAUNCH_CHECK();
    if ((rc = sort_pairs(h, h->kin, h->tk, h->pa, h->pb, n, 3 * kLv2, st))) return rc;
    hipLaunchKernelGGL(k_gather_u64, dim3(nb), dim3(kThreads), 0, st, n, h->K1o, h->pb, h->kin);
    LAUNCH_CHECK();
    if ((rc = sort_pairs(h, h->kin, h->ks1, h->pb, h->perm, n, 3 * kLv1, st))) return rc;
    hipLaunchKernelGGL(k_gather_u64, dim3(nb), dim3(kThreads), 0, st, n, h->K2o, h->perm, h->ks2);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gather_u64, dim3(nb), dim3(kThreads), 0, st, n, h->K3o, h->perm, h->ks3);
    LAUNCH_CHECK();
    if ((rc = count_nodes<3>(h, pos, m, n, h->ks1, h->ks2, h->ks3, st))) return rc;
  }
  h->nn = h->info_h[0];
  h->max_depth = h->info_h[2];
  if (h->deep) rc = make_nodes<3>(h, n, h->ks1, h->ks2, h->ks3, st);
  else rc = make_nodes<1>(h, n, h->ks1, nullptr, nullptr, st);
  if (rc) return rc;
  h->built = 1;
  return ACCU_OK;
}

int accu_bh_walk(accu_bh* h, double G, double eps2, double theta, double* F, double* U, void* stream) {
  ARG_CHECK(h && F, "null argument");
  ARG_CHECK(h->built, "accu_bh_walk: no tree (call accu_bh_build first)");
  hipStream_t st = (hipStream_t)stream;
  const int n = h->n;
  const int nb = (n + kThreads - 1) / kThreads;
  double* epart = h->part + 6 * 1024;
  unsigned* ecount = h->counter + 1;
  if (h->quad)
    hipLaunchKernelGGL(k_bh_walk<true>, dim3(nb), dim3(kThreads), 0, st, n, h->nn, h->com, h->cube, h->next, h->leaf,
                       h->lstart, h->lcount, h->quadv, h->perm, h->ppos, G, eps2, theta * theta, F, epart, ecount, U);
  else
    hipLaunchKernelGGL(k_bh_walk<false>, dim3(nb), dim3(kThreads), 0, st, n, h->nn, h->com, h->cube, h->next, h->leaf,
                       h->lstart, h->lcount, h->quadv, h->perm, h->ppos, G, eps2, theta * theta, F, epart, ecount, U);
  LAUNCH_CHECK();
  return ACCU_OK;
}

int accu_bh_info(accu_bh* h, long long* out_host) {
  ARG_CHECK(h && out_host, "null argument");
  int leaves = 0;
  if (h->built) {
    HIP_TRY(hipMemcpy(&leaves, h->info + 3, sizeof(int), hipMemcpyDeviceToHost));
  }
  out_host[0] = h->n;
  out_host[1] = h->built ? h->nn : 0;
  out_host[2] = leaves;
  out_host[3] = h->max_depth;
  out_host[4] = h->deep;
  out_host[5] = h->leaf_size;
  return ACCU_OK;
}

int accu_bh_export(accu_bh* h, double* cube, double* com, double* quad, int* next, int* leaf, int* start, int* count,
                   int* perm, void* stream) {
  ARG_CHECK(h, "null argument");
  ARG_CHECK(h->built, "accu_bh_export: no tree (call accu_bh_build first)");
  ARG_CHECK(!quad || h->quad, "accu_bh_export: the tree was built without quadrupoles");
  hipStream_t st = (hipStream_t)stream;
  const size_t nn = h->nn;
  if (cube) HIP_TRY(hipMemcpyAsync(cube, h->cube, nn * 32, hipMemcpyDeviceToDevice, st));
  if (com) HIP_TRY(hipMemcpyAsync(com, h->com, nn * 32, hipMemcpyDeviceToDevice, st));
  if (quad) HIP_TRY(hipMemcpyAsync(quad, h->quadv, nn * 48, hipMemcpyDeviceToDevice, st));
  if (next) HIP_TRY(hipMemcpyAsync(next, h->next, nn * 4, hipMemcpyDeviceToDevice, st));
  if (leaf) HIP_TRY(hipMemcpyAsync(leaf, h->leaf, nn * 4, hipMemcpyDeviceToDevice, st));
  if (start) HIP_TRY(hipMemcpyAsync(start, h->lstart, nn * 4, hipMemcpyDeviceToDevice, st));
  if (count) HIP_TRY(hipMemcpyAsync(count, h->lcount, nn * 4, hipMemcpyDeviceToDevice, st));
  if (perm) HIP_TRY(hipMemcpyAsync(perm, h->perm, (size_t)h->n * 4, hipMemcpyDeviceToDevice, st));
  return ACCU_OK;
}

}  // extern "C"
