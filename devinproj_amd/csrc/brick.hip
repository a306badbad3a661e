// Brick-tiled, LDS-staged MD engine for one or many periodic systems.
//
// Same physics as the single-system engine (integrators.py:19-47 with the Verlet rule
// of forces.py:86-98 over LJ forces.py:119-142 and DSF forces.py:408-438), laid out for
// the LDS: cells of each system are grouped into small bricks (bx x by x bz cells),
// cell ids are numbered brick by brick, and atoms are kept sorted by cell id, so a
// brick's home atoms are one contiguous range. A workgroup owns one brick:
//   1. it stages the atoms of the (bx+2)(by+2)(bz+2) region cells around the brick
//      into LDS, shifted by their periodic image and taken relative to the brick
//      centre (float64, or float32 in mixed mode), and
//   2. its home atoms read neighbours from LDS through a Verlet list of 16-bit local
//      indices into that region (built the same way at every rebuild).
// Neighbour positions therefore come from LDS instead of L2/HBM gathers, the list is
// half the bytes, and no minimum-image arithmetic is left in the pair loop.
//
// Several systems (e.g. Richardson replicas at dt, dt/2, dt/4 of many sizes) share one
// arena: systems are ordered so that the systems advanced at each "tick" form a prefix
// of the brick array, each system carries its own box, interactions, dt and rebuild
// flag, and every kernel of the rebuild chain skips systems whose flag is clear.
#include "common.h"
#include "pairkern.h"
#include "../../include/accusim_hip.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

namespace accu {

constexpr int kBT = 512;  // launch bound; the launch size is the runtime `threads`
constexpr int kMaxRegionCells = 256;
constexpr int kMaxHomeCells = 128;
constexpr int kRecW = 12;  // PE, KE, Px, Py, Pz, rebuilds, Lx, Ly, Lz, 0, 0, 0
constexpr int kRecP = 8;   // brick partials: PE, KE, Px, Py, Pz, Lx, Ly, Lz
constexpr int kSplit = 8;  // fused step: lanes per atom in a short last pass

struct BSys {
  double L[3], inv[3];
  double cinv[3];  // ncell / L
  int nc[3];
  int cell_off, ncells;
  int atom_off, natoms;
  int brick_off, nbricks;
  int cellmap_off;
  int ns;
  double mass[kMaxSpecies], charge[kMaxSpecies];
  double diff[kMaxSpecies];  // Brownian diffusion coefficient per species
  int has_lj, has_coul;
  double lj_eps[kMaxSpecies * kMaxSpecies], lj_sig2[kMaxSpecies * kMaxSpecies], lj_rc2[kMaxSpecies * kMaxSpecies];
  double lj_shift;
  CoulParams coul;
  double rc_list2;
  double half_skin2;
  double half_skin2_q;  // the same in fixed-point quanta squared (mixed mode)
  double dt;
  // mixed mode: the fixed-point position copy (offsets from each atom's cell corner in
  // quanta of unit * 2^-fx_k, `unit` being the staging unit of the fused launch)
  double hc[3];    // cell edge L / nc
  double fx_q;     // 2^fx_k / unit: metres -> quanta
  // the list launch's integer staging: coordinates in quanta bq = 2^bq_sh / fx_q (about 1/8192
  // of a cell), x and y as int16, z as int32; a pair is listed when its squared quantised
  // distance is below bq_T, which covers rc_list plus the rounding of both atoms
  double bq_inv;   // 1 / bq (metres -> build quanta)
  int bq_sh, bq_T, bq_W;  // shift from the fixed-point quanta; r^2 threshold; z window half-width
  float bq_m;      // bq in metres
  float bq_rcp2;   // squared radius of the column prune (metres), rc_list widened by the rounding
  float fx_s;      // 2^-fx_k: quanta -> staging units of the fused launch (exact)
  float fx_sm;     // unit * 2^-fx_k: quanta -> metres (the list launch)
  int fx_k;
};

struct BBrick {
  int sys;
  int cell0, nhome;  // home cells: global ids [cell0, cell0 + nhome)
  int hd[3], rd[3];
  int reg_off;
  double org[3];
};

struct BRegion {
  int gid;
  int shift;  // (sx+1) + 3 (sy+1) + 9 (sz+1), image offset s * L
};

template <typename P>
struct P4 {};
template <>
struct P4<double> {
  typedef double4 T;
};
template <>
struct P4<float> {
  typedef float4 T;
};

// ------------------------------------------------------------------ region staging

// A staged float32 slot without the species word (one species, LJ only): 12-byte slots
// make a wave's random neighbour reads hit 3 s mod 64 banks instead of one of 16 bank
// groups, which costs an extra LDS instruction per read and still runs faster.
struct Pk3 {
  float x, y, z;
};

template <bool PK3, typename T4>
__device__ __forceinline__ T4 ld_slot(const T4* lds, unsigned q) {
  if constexpr (PK3) {
    const Pk3 v = reinterpret_cast<const Pk3*>(lds)[q];
    T4 r;
    r.x = v.x;
    r.y = v.y;
    r.z = v.z;
    r.w = 0.f;
    return r;
  } else {
    return lds[q];
  }
}

template <bool PK3, typename T4>
__device__ __forceinline__ void st_slot(T4* lds, unsigned q, const T4& v) {
  if constexpr (PK3) {
    Pk3 w;
    w.x = v.x;
    w.y = v.y;
    w.z = v.z;
    reinterpret_cast<Pk3*>(lds)[q] = w;
  } else {
    lds[q] = v;
  }
}

// Entry k of a 16-byte list group (k a compile-time constant after unrolling): W12, 12-bit
// slot indices at bits 8 + 12 k (10 per group); else 16-bit at bits 16 k (8 per group).
// Decoded where the slot is read, so only the group itself stays live.
template <bool W12>
__device__ __forceinline__ unsigned list_entry(const uint4& g, int k) {
  if constexpr (W12) {
    switch (k) {
      case 0: return (g.x >> 8) & 0xfffu;
      case 1: return g.x >> 20;
      case 2: return g.y & 0xfffu;
      case 3: return (g.y >> 12) & 0xfffu;
      case 4: return __builtin_amdgcn_alignbit(g.z, g.y, 24) & 0xfffu;
      case 5: return (g.z >> 4) & 0xfffu;
      case 6: return (g.z >> 16) & 0xfffu;
      case 7: return __builtin_amdgcn_alignbit(g.w, g.z, 28) & 0xfffu;
      case 8: return (g.w >> 8) & 0xfffu;
      default: return g.w >> 20;
    }
  } else {
    const unsigned w = k < 2 ? g.x : k < 4 ? g.y : k < 6 ? g.z : g.w;
    return (k & 1) ? (w >> 16) : (w & 0xffffu);
  }
}

// Exclusive prefix of v over the block; total receives the block sum.
__device__ __forceinline__ int block_excl_scan(int v, int* wsum, int& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  __syncthreads();
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int woff = 0, tot = 0;
  for (int k = 0; k < nw; ++k) {
    woff += k < w ? wsum[k] : 0;
    tot += wsum[k];
  }
  total = tot;
  return woff + x - v;
}

__device__ __forceinline__ int home_region(const BBrick& b, int local) {
  // home cell local index (z fastest) -> region cell of the same cell
  const int lz = local % b.hd[2], ly = (local / b.hd[2]) % b.hd[1], lx = local / (b.hd[2] * b.hd[1]);
  return ((lx + 1) * b.rd[1] + (ly + 1)) * b.rd[2] + (lz + 1);
}

// Region set-up of one brick: cell starts rs[0..R] (block scan of the cell counts) and
// the region slot of each home cell's first atom (home atom a of home cell `local` sits
// in slot hself[local] + a). Returns the number of region atoms.
__device__ int region_setup(const BBrick& b, const BRegion* __restrict__ reg, const int* __restrict__ start, int* rs,
                            int* wsum, int* hself) {
  const int R = b.rd[0] * b.rd[1] * b.rd[2];
  int carry = 0;
  for (int r0 = 0; r0 < R; r0 += blockDim.x) {  // uniform trip count
    const int r = r0 + threadIdx.x;
    int cnt = 0;
    if (r < R) {
      const int c = reg[b.reg_off + r].gid;
      cnt = start[c + 1] - start[c];
    }
    int tot;
    const int ex = block_excl_scan(cnt, wsum, tot);
    if (r < R) rs[r] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) rs[R] = carry;
  __syncthreads();
  for (int l = threadIdx.x; l < b.nhome; l += blockDim.x) hself[l] = rs[home_region(b, l)] - start[b.cell0 + l];
  __syncthreads();
  return carry;
}

// Stage region atoms: lds[q] = (pos - org) * scale with species in w, where pos is taken
// at the periodic image nearest the centre of its region cell. That image is the one the
// region's cell shift names; positions are also wrapped into the box every step (as the
// reference does), so an atom that crosses a box face between rebuilds keeps its cell
// while its wrapped coordinate jumps by L; it is at most half a skin outside its cell and
// cells are at most L/3 wide, so the nearest image is unambiguous. Each region cell is
// staged by a group of blockDim/R threads (consecutive slots, coalesced reads), which
// needs neither a slot search nor any per-slot integer division.
template <typename P, bool PK3 = false>
__device__ void stage_region(const BBrick& b, const BSys& s, const BRegion* __restrict__ reg,
                             const int* __restrict__ start, const int* rs, const double4* __restrict__ pos,
                             typename P4<P>::T* lds, double scale = 1.0) {
  const int R = b.rd[0] * b.rd[1] * b.rd[2];
  const int G = max(1, (int)blockDim.x / R);  // threads per region cell
  const int per = (int)blockDim.x / G;         // region cells per pass
  const int sub = threadIdx.x % G;
  for (int r = threadIdx.x / G; r < R && threadIdx.x < per * G; r += per) {
    const int q0 = rs[r], q1 = rs[r + 1];
    const int base = start[reg[b.reg_off + r].gid] - q0;
    const int rr[3] = {r / (b.rd[2] * b.rd[1]), (r / b.rd[2]) % b.rd[1], r % b.rd[2]};
    double cen[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)  // region cell 0 starts 1 + hd/2 cells below the brick centre
      cen[d] = (rr[d] - 0.5 - 0.5 * b.hd[d]) * (s.L[d] / s.nc[d]);
    // two slots per thread and step, their loads in flight together
    for (int q = q0 + sub; q < q1; q += 2 * G) {
      double4 xx[2];
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (q + u * G < q1) xx[u] = pos[base + q + u * G];
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (q + u * G < q1) {
          const double xs[3] = {xx[u].x, xx[u].y, xx[u].z};
          double X[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double e = xs[d] - b.org[d];
            X[d] = fma(-s.L[d], rint((e - cen[d]) * s.inv[d]), e);
          }
          typename P4<P>::T v;
          v.x = (P)(X[0] * scale);
          v.y = (P)(X[1] * scale);
          v.z = (P)(X[2] * scale);
          v.w = (P)xx[u].w;
          st_slot<PK3>(lds, q + u * G, v);
        }
    }
  }
}

// ------------------------------------------------------------------ rebuild (one launch)
//
// Cell binning, counting sort and permutation into brick order of every flagged system
// run as the phases of ONE persistent launch separated by grid barriers (the grid is
// sized from the occupancy so that every block is resident); the Verlet lists follow
// in a second launch with one block per brick. Steps without a rebuild therefore cost
// two launches that exit at once. The rebuild launch also carries the step's
// bookkeeping: it flips the position-buffer parity of the systems that just drifted
// and hands the rebuild flags to the list launch.

constexpr int kRB = 1024;  // threads per block of the rebuild launch: one block per CU, so that
                           // few blocks pay the cache write-back of each barrier
constexpr int kBarSpin = 1 << 20;  // bounded barrier wait (0.1 s or more): a stuck barrier reports an error

// Per-atom engine state except positions, double-buffered like the positions: system s
// uses buffer spar[s]; a rebuild gathers into the other buffer and flips spar[s].
struct BState {
  double* vel[2];  // velocities, 3 doubles per atom (ld_vel / st_vel)
  double4* F[2];
  double4* ref[2];  // positions at the last list build
  int* perm[2];     // user index of each slot
  int* cell[2];     // cell id of each slot
  // mixed mode only (else null): the positions as fixed-point offsets from the corner of
  // each atom's cell, fxw ints per atom: x, y, z and, with several species, the species;
  // posf[0] pairs with posA and posf[1] with posB (the position parity, not the state
  // parity)
  int* posf[2];
  // mixed mode only (else null): the fixed-point positions at the last list build, in the
  // layout of posf and the state parity of ref (which mixed mode leaves unused)
  int* refq[2];
  const double4* ccorner;  // corner of every cell in the box frame (both modes)
  int fxw;                 // ints per atom of posf and refq: 3 with one species, else 4
};

// Velocities are kept as 24-byte records (x, y, z), not padded double4: 16 bytes less per
// atom and step for the fused launch, which reads and writes every velocity.
struct __attribute__((aligned(8))) Vel3 {
  double x, y, z;
};
__device__ __forceinline__ double4 ld_vel(const double* __restrict__ v, int i) {
  const Vel3 t = *reinterpret_cast<const Vel3*>(v + 3 * (size_t)i);
  return make_double4(t.x, t.y, t.z, 0.0);
}
__device__ __forceinline__ void st_vel(double* __restrict__ v, int i, const double4& u) {
  const Vel3 t = {u.x, u.y, u.z};
  *reinterpret_cast<Vel3*>(v + 3 * (size_t)i) = t;
}

// Fixed-point copy of position x of an atom of the cell with corner cc: per axis the offset
// from the corner at the periodic image nearest the cell's centre, in quanta of
// unit * 2^-fx_k (|offset| stays within a cell and half a skin, far inside int32).
__device__ __forceinline__ int4 fix_pos(const BSys& s, const double4& x, const double4& cc) {
  const double o[3] = {x.x - cc.x, x.y - cc.y, x.z - cc.z};
  int q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double e = fma(-s.L[d], rint((o[d] - 0.5 * s.hc[d]) * s.inv[d]), o[d]);
    q[d] = (int)rint(e * s.fx_q);
  }
  return make_int4(q[0], q[1], q[2], (int)x.w);
}

// One atom's record of posf / refq: 12 bytes (x, y, z) with one species, else 16 (the
// species in w). The 12-byte record is read and written as one dwordx3 access.
struct __attribute__((aligned(4))) Fix3 {
  int x, y, z;
};
template <bool W3>
__device__ __forceinline__ int4 ld_fix(const int* __restrict__ p, int i) {
  if constexpr (W3) {
    const Fix3 v = *reinterpret_cast<const Fix3*>(p + 3 * (size_t)i);
    return make_int4(v.x, v.y, v.z, 0);
  } else {
    return *reinterpret_cast<const int4*>(p + 4 * (size_t)i);
  }
}
__device__ __forceinline__ int4 ld_fix(const int* __restrict__ p, int i, int w) {
  return w == 3 ? ld_fix<true>(p, i) : ld_fix<false>(p, i);
}
__device__ __forceinline__ void st_fix(int* __restrict__ p, int i, int w, const int4& v) {
  if (w == 3) {
    const Fix3 t = {v.x, v.y, v.z};
    *reinterpret_cast<Fix3*>(p + 3 * (size_t)i) = t;
  } else {
    *reinterpret_cast<int4*>(p + 4 * (size_t)i) = v;
  }
}

// Mixed mode: whether an atom has moved more than half the skin since the last list
// build, from its fixed-point position now (f) and then (r). Both are offsets from the
// corner of the same cell (cells change only at rebuilds), so the displacement is an
// exact integer difference, compared in quanta squared.
__device__ __forceinline__ bool moved_far_q(const BSys& s, const int4& f, const int4& r) {
  const double dx = (double)(f.x - r.x), dy = (double)(f.y - r.y), dz = (double)(f.z - r.z);
  return !(fma(dz, dz, fma(dy, dy, dx * dx)) < s.half_skin2_q);
}

// Mixed mode: stage the region from the fixed-point copy. Region cell r (an image of cell
// gid shifted by s L) has its corner at B = corner(gid) + s L - org relative to the brick
// centre, so an atom's staged coordinate is (B + offset) in quanta: an exact integer sum,
// then one rounding to float and an exact power-of-two scale (`scale` = 2^-fx_k), which is
// the value the float64 staging rounds to, up to one quantum. 16 bytes per atom read
// instead of 32, and no float64 image arithmetic per atom.
template <bool PK3, bool W3>
__device__ void stage_region_fix(const BBrick& b, const BSys& s, const BRegion* __restrict__ reg,
                                 const int* __restrict__ start, const int* rs, const int* __restrict__ posf,
                                 const double4* __restrict__ ccorner, float4* lds, float scale) {
  const int R = b.rd[0] * b.rd[1] * b.rd[2];
  const int G = max(1, (int)blockDim.x / R);  // threads per region cell
  const int per = (int)blockDim.x / G;         // region cells per pass
  const int sub = threadIdx.x % G;
  for (int r = threadIdx.x / G; r < R && threadIdx.x < per * G; r += per) {
    const int q0 = rs[r], q1 = rs[r + 1];
    const BRegion rg = reg[b.reg_off + r];
    const int base = start[rg.gid] - q0;
    const double4 cc = ccorner[rg.gid];
    const double c3[3] = {cc.x, cc.y, cc.z};
    const int sh[3] = {rg.shift % 3 - 1, (rg.shift / 3) % 3 - 1, rg.shift / 9 - 1};
    int B[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) B[d] = (int)rint((c3[d] + sh[d] * s.L[d] - b.org[d]) * s.fx_q);
    // four slots per thread and step, their loads in flight together
    for (int q = q0 + sub; q < q1; q += 4 * G) {
      int4 o[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (q + u * G < q1) o[u] = ld_fix<W3>(posf, base + q + u * G);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (q + u * G < q1) {
          float4 v;
          v.x = (float)(B[0] + o[u].x) * scale;
          v.y = (float)(B[1] + o[u].y) * scale;
          v.z = (float)(B[2] + o[u].z) * scale;
          v.w = W3 ? 0.0f : (float)o[u].w;  // one species: species 0
          st_slot<PK3>(lds, q + u * G, v);
        }
    }
  }
}

// The list launch's region staging: 8-byte integer slots (x, y as int16 in the low and
// high halves of the first word, z as int32 in the second), in build quanta relative to the
// brick centre. FIX (mixed mode): from the fixed-point copy, the exact integer coordinate
// rounded by an arithmetic shift; else from the float64 positions with the image of
// stage_region. Rounding is monotone, so each column's run stays sorted by z.
__device__ __forceinline__ uint2 pack_q(int x, int y, int z) {
  return make_uint2(((unsigned)x & 0xffffu) | ((unsigned)y << 16), (unsigned)z);
}
template <bool FIX>
__device__ void stage_region_q(const BBrick& b, const BSys& s, const BRegion* __restrict__ reg,
                               const int* __restrict__ start, const int* rs, const int* __restrict__ posf,
                               int fxw, const double4* __restrict__ pos, const double4* __restrict__ ccorner,
                               uint2* lds) {
  const int R = b.rd[0] * b.rd[1] * b.rd[2];
  const int G = max(1, (int)blockDim.x / R);  // threads per region cell
  const int per = (int)blockDim.x / G;         // region cells per pass
  const int sub = threadIdx.x % G;
  const int sh = s.bq_sh, half = sh > 0 ? 1 << (sh - 1) : 0;
  for (int r = threadIdx.x / G; r < R && threadIdx.x < per * G; r += per) {
    const int q0 = rs[r], q1 = rs[r + 1];
    const BRegion rg = reg[b.reg_off + r];
    const int base = start[rg.gid] - q0;
    if constexpr (FIX) {
      const double4 cc = ccorner[rg.gid];
      const double c3[3] = {cc.x, cc.y, cc.z};
      const int shv[3] = {rg.shift % 3 - 1, (rg.shift / 3) % 3 - 1, rg.shift / 9 - 1};
      int B[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) B[d] = (int)rint((c3[d] + shv[d] * s.L[d] - b.org[d]) * s.fx_q);
      for (int q = q0 + sub; q < q1; q += 2 * G) {  // two loads in flight per thread
        int4 o[2];
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (q + u * G < q1) o[u] = ld_fix(posf, base + q + u * G, fxw);
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (q + u * G < q1)
            lds[q + u * G] = pack_q((B[0] + o[u].x + half) >> sh, (B[1] + o[u].y + half) >> sh,
                                    (B[2] + o[u].z + half) >> sh);
      }
    } else {
      const int rr[3] = {r / (b.rd[2] * b.rd[1]), (r / b.rd[2]) % b.rd[1], r % b.rd[2]};
      double cen[3];
#pragma unroll
      for (int d = 0; d < 3; ++d)  // region cell 0 starts 1 + hd/2 cells below the brick centre
        cen[d] = (rr[d] - 0.5 - 0.5 * b.hd[d]) * (s.L[d] / s.nc[d]);
      for (int q = q0 + sub; q < q1; q += 2 * G) {  // two loads in flight per thread
        double4 xx[2];
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (q + u * G < q1) xx[u] = pos[base + q + u * G];
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (q + u * G < q1) {
            const double xs[3] = {xx[u].x, xx[u].y, xx[u].z};
            int X[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const double e = xs[d] - b.org[d];
              X[d] = (int)rint(fma(-s.L[d], rint((e - cen[d]) * s.inv[d]), e) * s.bq_inv);
            }
            lds[q + u * G] = pack_q(X[0], X[1], X[2]);
          }
      }
    }
  }
}

struct RB {
  const BSys* sys;
  const BBrick* bricks;
  const BRegion* reg;
  const int* cellmap;
  const int* cellsys;
  const int* sysid;
  double4* posA;
  double4* posB;
  int* par;
  BState st;
  int* spar;
  int* count;   // per cell; all zero between rebuilds
  int* start;   // [ncells + 1]
  int* cursor;  // [ncells + 1]
  int* tmp;     // atoms grouped by cell (when a bucket overflowed)
  int* slot;    // each atom's arrival slot in its cell
  int* bucket;  // [ncells][bcap]: atoms of each cell in arrival order, written by the binning
  int bcap;     // bucket capacity per cell (0: no buckets, group through tmp)
  int* bsum;    // [grid]
  unsigned long long* look;  // [grid] the scan's per-block sums, tagged with the barrier generation
  int* nn;
  uint4* nbr;   // 16-byte groups of local slot indices (12-bit x 10 or 16-bit x 8): group g of atom a at [g * stride + a]
  int stride, ngroups, cap;
  int* ints;    // [0] max neighbours, [1] region overflow, [2] barrier timeout, [3] blocks seen then,
                // [4] a cell outgrew its bucket in this rebuild, [5] rebuilds in which one did
  long long* tphase;  // [0] start; ends of binning, scan, grouping ([4] = [3]), gather (device wall clock)
  unsigned* bar;
  int* flags;   // [S] per system + [S] any, set by the fused step
  int* bflags;  // the same, as handed to the list launch
  int* nreb;
  long long* ndrift;  // per system: drifts done (the stochastic integrators' noise counter)
  int* islot;         // slot of every user index (rigid bonds only, else null)
  int S, N, ncells, nbricks;
  int flip_n;    // systems [0, flip_n) drifted: their parity flips with this launch
  int bin_only;  // 1: stop after the scan (set-up measures the regions)
  // record of the step before (few systems): block 0 sums the brick partials
  int rec;
  const double* partials;
  double* log;
  int log_cap;
  int* log_count;
};

constexpr int kRecInline = 4;  // up to this many systems the rebuild launch records; else k_b_record

// Per-system sums of the brick partials in brick order (fixed thread assignment and a
// fixed-order block sum), appended to the system's log with its rebuild count.
__device__ void record_systems(const BSys* __restrict__ sys, int S, const double* __restrict__ partials,
                               double* __restrict__ log, int log_cap, int* __restrict__ log_count,
                               const int* __restrict__ nreb, double* red) {
  for (int si = 0; si < S; ++si) {
    const BSys& s = sys[si];
    double acc[kRecP] = {0, 0, 0, 0, 0, 0, 0, 0};
    // two bricks per thread and pass: their loads (which miss the L2, the partials being
    // written by the previous launch) are in flight together; the sum order stays fixed
    const double* __restrict__ ps = partials + (size_t)s.brick_off * kRecP;
    for (int b = threadIdx.x; b < s.nbricks; b += 2 * blockDim.x) {
      const int b2 = b + blockDim.x < s.nbricks ? b + blockDim.x : b;
      const double w2 = b + blockDim.x < s.nbricks ? 1.0 : 0.0;
      double v[kRecP], u[kRecP];
#pragma unroll
      for (int k = 0; k < kRecP; ++k) {
        v[k] = ps[(size_t)b * kRecP + k];
        u[k] = ps[(size_t)b2 * kRecP + k];
      }
#pragma unroll
      for (int k = 0; k < kRecP; ++k) acc[k] += v[k];
#pragma unroll
      for (int k = 0; k < kRecP; ++k) acc[k] = fma(w2, u[k], acc[k]);
    }
    block_sum<kRecP>(acc, red);
    if (threadIdx.x == 0) {
      const int idx = log_count[si];
      if (idx < log_cap) {
        double* o = log + ((size_t)si * log_cap + idx) * kRecW;
        for (int k = 0; k < 5; ++k) o[k] = acc[k];
        o[5] = (double)nreb[si];
        for (int k = 5; k < kRecP; ++k) o[k + 1] = acc[k];
        for (int k = kRecP + 1; k < kRecW; ++k) o[k] = 0.0;
      }
      log_count[si] = idx + 1;
    }
  }
}

// Grid-wide barrier. Returns false when the wait times out: err[0] is then set (every
// block leaves and the engine reports the error at the next get) and err[1] keeps the
// largest number of blocks seen at the barrier.
__device__ bool grid_sync(unsigned* __restrict__ bar, int* __restrict__ err) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int good = 1;
    // release this block's writes once (the fence makes the relaxed arrival below a release:
    // each further release would be one more L2 write-back); the polling loads are relaxed
    // so that the waiting blocks do not keep invalidating the caches of the working ones
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned g = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned arrived = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == gridDim.x - 1) {
      __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      int spins = 0;
      while (__hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (++spins > kBarSpin) {
          good = 0;
          atomicOr(err, 1);
          atomicMax(err + 1, (int)__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the other blocks' writes in
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

// Grid-wide barrier for data exchanged only through agent-scope atomics: every wave waits
// for its own memory operations (the atomics it issued are then performed), and the
// arrival and the wait are relaxed, with no release or acquire fence, so no block writes
// back or invalidates its L2. Plain stores made before it are NOT visible to other blocks
// after it (a later grid_sync publishes them). Timeouts are reported as in grid_sync.
__device__ bool grid_sync_atomics(unsigned* __restrict__ bar, int* __restrict__ err) {
  __shared__ int ok;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    int good = 1;
    const unsigned g = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned arrived = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == gridDim.x - 1) {
      __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      int spins = 0;
      while (__hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (++spins > kBarSpin) {
          good = 0;
          atomicOr(err, 1);
          atomicMax(err + 1, (int)__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

// Sum of v over the block (every thread gets it).
__device__ __forceinline__ int block_sum_int(int v, int* wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if (lane == 0) wsum[w] = v;
  __syncthreads();
  int t = 0;
  for (int k = 0; k < nw; ++k) t += wsum[k];
  return t;
}

// Each system's current positions live in posA or posB according to par[system]
// (double buffering per system, so systems can advance at different rates).
__device__ __forceinline__ const double4* cur_buf(const double4* A, const double4* B, const int* par, int s) {
  return par[s] ? B : A;
}

// Verlet list of one brick: every home atom against the region cells around its own
// cell, local slot indices in 16-byte groups, the last group padded with the index of
// the far sentinel slot (`total`). W12: 12-bit slot indices, 10 per group, entry k at
// bits 8 + 12 k (regions of at most 4,095 slots); else 16-bit, 8 per group, entry k at
// bits 16 k. The list only has to contain every pair within rc_list (a superset is
// harmless: forces apply their own cutoffs), so the test runs on integer coordinates
// (stage_region_q): a packed 16-bit difference of x and y, an int32 difference of z,
// one dot product, against a threshold that covers rc_list plus the rounding of both
// atoms; region cells whose box is beyond reach of the atom are skipped.
typedef short short2v __attribute__((ext_vector_type(2)));
template <bool W12>
__device__ void build_brick(const RB& p, const BBrick& b, uint2* lds, int* rs, int* hself, int* wsum, int& mymax) {
  constexpr int kE = W12 ? 10 : 8;    // entries per group
  constexpr int kW = W12 ? 12 : 16;   // bits per entry
  const BSys& s = p.sys[b.sys];
  const int total = region_setup(b, p.reg, p.start, rs, wsum, hself);
  if (total + 1 > p.cap) {
    if (threadIdx.x == 0) atomicMax(&p.ints[1], total + 1);
    return;
  }
  if (p.st.posf[0])
    stage_region_q<true>(b, s, p.reg, p.start, rs, p.st.posf[p.par[b.sys]], p.st.fxw, nullptr, p.st.ccorner, lds);
  else
    stage_region_q<false>(b, s, p.reg, p.start, rs, nullptr, 0, cur_buf(p.posA, p.posB, p.par, b.sys), nullptr,
                          lds);
  __syncthreads();
  const float h[3] = {(float)(s.L[0] / s.nc[0]), (float)(s.L[1] / s.nc[1]), (float)(s.L[2] / s.nc[2])};
  const float c0[3] = {-(1.0f + 0.5f * b.hd[0]) * h[0], -(1.0f + 0.5f * b.hd[1]) * h[1],
                       -(1.0f + 0.5f * b.hd[2]) * h[2]};  // low corner of region cell 0
  const float rcp2 = s.bq_rcp2, bqm = s.bq_m;
  // a candidate with |dz| >= W (quanta) has dz^2 >= T and fails the test anyway
  const int T = s.bq_T, W = s.bq_W;
  const unsigned sentinel = (unsigned)total;
  const int* __restrict__ cell = p.st.cell[p.spar[b.sys]];
  const int a0 = p.start[b.cell0], a1 = p.start[b.cell0 + b.nhome];
  for (int a = a0 + threadIdx.x; a < a1; a += blockDim.x) {
    const int local = cell[a] - b.cell0;
    const int self = hself[local] + a;
    const uint2 si = lds[self];
    const short2v xyi = __builtin_bit_cast(short2v, si.x);
    const int zi = (int)si.y;
    const float xm = (float)xyi.x * bqm, ym = (float)xyi.y * bqm;  // metres, for the column prune
    const int lz = local % b.hd[2], ly = (local / b.hd[2]) % b.hd[1], lx = local / (b.hd[2] * b.hd[1]);
    // each hit shifts the 128-bit register down by kW bits and enters at the top, so the
    // first of a group's kE entries ends at bit 128 - kE kW (0 for 16-bit, 8 for 12-bit)
    int cnt = 0, full = kE - 1;  // full: the count at which the current group is complete
    int grp = 0;                 // groups stored
    unsigned w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    for (int ox = 0; ox < 3; ++ox) {
      const int rx = lx + ox;
      const float ex = fmaxf(fmaxf(c0[0] + rx * h[0] - xm, xm - (c0[0] + (rx + 1) * h[0])), 0.f);
      for (int oy = 0; oy < 3; ++oy) {
        const int ry = ly + oy;
        const float ey = fmaxf(fmaxf(c0[1] + ry * h[1] - ym, ym - (c0[1] + (ry + 1) * h[1])), 0.f);
        if (ex * ex + ey * ey > rcp2) continue;  // the column's box is out of reach
        // the column's three cells (rz = lz .. lz + 2) are one run of slots sorted by z:
        // scan only from near the first candidate with dz > -W (skipping 8 slots at a
        // time past those below) to the first with dz >= W
        const int r0 = (rx * b.rd[1] + ry) * b.rd[2] + lz;
        int lo = rs[r0];
        const int qe = rs[r0 + 3];
        while (lo + 8 < qe && (int)lds[lo + 8].y - zi <= -W) lo += 8;
        // the atom itself is a candidate only in its own column
        const int skip = (ox == 1 && oy == 1) ? self : -1;
        // four candidates per step, at one base address plus fixed offsets; the next
        // step's four are read while this step's are tested, so no read waits on the loop
        // test (slots past the run's end, at most 7 of them and inside the LDS padding,
        // are read and either rejected by q + k < qe or never used)
        uint2 nx[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) nx[k] = lds[lo + k];
        for (int q = lo; q < qe; q += 4) {
          uint2 xj[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            xj[k] = nx[k];
            nx[k] = lds[q + 4 + k];
          }
          if ((int)xj[0].y - zi >= W) break;  // this and every later candidate is above the window
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const short2v dxy = __builtin_bit_cast(short2v, xj[k].x) - xyi;
            const int dz = (int)xj[k].y - zi;
            const int r2 = __builtin_amdgcn_sdot2(dxy, dxy, __mul24(dz, dz), false);
            if (r2 < T && q + k < qe && q + k != skip) {
              w0 = __builtin_amdgcn_alignbit(w1, w0, kW);
              w1 = __builtin_amdgcn_alignbit(w2, w1, kW);
              w2 = __builtin_amdgcn_alignbit(w3, w2, kW);
              w3 = __builtin_amdgcn_alignbit((unsigned)(q + k), w3, kW);
              if (cnt == full) {  // the last entry of a group: store the group
                if (grp < p.ngroups) p.nbr[(size_t)grp * p.stride + a] = make_uint4(w0, w1, w2, w3);
                ++grp;
                full += kE;
              }
              ++cnt;
            }
          }
        }
      }
    }
    if (cnt > grp * kE) {  // pad the last group with the sentinel
      for (int e = cnt - grp * kE; e < kE; ++e) {
        w0 = __builtin_amdgcn_alignbit(w1, w0, kW);
        w1 = __builtin_amdgcn_alignbit(w2, w1, kW);
        w2 = __builtin_amdgcn_alignbit(w3, w2, kW);
        w3 = __builtin_amdgcn_alignbit(sentinel, w3, kW);
      }
      if (grp < p.ngroups) p.nbr[(size_t)grp * p.stride + a] = make_uint4(w0, w1, w2, w3);
    }
    const int capn = p.ngroups * kE;
    p.nn[a] = cnt < capn ? cnt : capn;
    mymax = cnt > mymax ? cnt : mymax;
  }
  __syncthreads();  // the LDS is restaged by the next brick of this block
}

// Parity of system si as seen inside the rebuild launch (the flip commits at its end).
__device__ __forceinline__ int rb_par(const RB& p, int si) { return p.par[si] ^ (si < p.flip_n ? 1 : 0); }

// Sort key of atom slot v of system si inside a cell: its current z.
__device__ __forceinline__ double zkey(const RB& p, int v, int si) { return (rb_par(p, si) ? p.posB : p.posA)[v].z; }

__global__ void __launch_bounds__(kRB) k_b_rebuild(RB p) {
  const int tid = threadIdx.x, bd = blockDim.x;
  if (p.rec && blockIdx.x == 0) {
    __shared__ double red[kRecP * (kRB / 64)];
    record_systems(p.sys, p.S, p.partials, p.log, p.log_cap, p.log_count, p.nreb, red);
  }
  if (p.flags[p.S] == 0 || p.ints[2] != 0) {
    if (blockIdx.x == 0) {  // nothing to rebuild: commit the flip, the list launch exits
      for (int s = tid; s < p.flip_n; s += bd) {
        p.par[s] ^= 1;
        p.ndrift[s] += 1;
      }
      if (tid == 0) p.bflags[p.S] = 0;
    }
    return;
  }
  __shared__ int wsum[kRB / 64];
  if (blockIdx.x == 0 && tid == 0) p.tphase[0] = wall_clock64();
  const int gtid = blockIdx.x * bd + tid, gsz = gridDim.x * bd;
  // 1. cell of every atom of a flagged system and its arrival slot in that cell; cells of
  //    the other systems keep their counts. Atoms are still nearly in cell order, so a
  //    wave's lanes share a few cells: one atomic per distinct cell per wave (the lanes
  //    of a cell take consecutive slots in lane order).
  //    A wave takes kBinU chunks of 64 atoms at once and issues the atomics of all their
  //    distinct cells before it waits for any return value.
  {
    constexpr int kBinU = 4;
    const int lane = tid & 63;
    for (int a00 = gtid - lane; a00 < p.N; a00 += kBinU * gsz) {
      int gid[kBinU], ret[kBinU];
      // the cell of atom a of system si (flagged for a rebuild); gid -1 otherwise
      // (every table value is loaded before the first is used: one round trip, then the
      // position, then the cell map)
      auto bin_cell = [&](int a, int si) -> int {
        const BSys& s = p.sys[si];
        const int fl = p.flags[si], pr = rb_par(p, si), spr = p.spar[si];
        const double ci[3] = {s.cinv[0], s.cinv[1], s.cinv[2]};
        const int nc[3] = {s.nc[0], s.nc[1], s.nc[2]}, moff = s.cellmap_off;
        if (!fl) return -1;
        const double4 x = (pr ? p.posB : p.posA)[a];
        const double xs[3] = {x.x, x.y, x.z};
        int c[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double t = xs[d] * ci[d];
          const int m = nc[d] - 1;
          const int tr = (int)fmin(fmax(t, 0.0), (double)m);  // t in [0, m): its truncation
          c[d] = !(t >= 0.0) ? 0 : (t >= (double)m ? m : tr);
        }
        const int g = p.cellmap[moff + (c[0] * nc[1] + c[1]) * nc[2] + c[2]];
        p.st.cell[spr][a] = g;
        return g;
      };
#pragma unroll
      for (int u = 0; u < kBinU; ++u) {
        const int a = a00 + u * gsz + lane;
        gid[u] = -1;
        ret[u] = 0;
        if (a00 + u * gsz >= p.N) continue;  // the whole chunk is past the end (uniform)
        const int si_l = a < p.N ? p.sysid[a] : -1;
        // a chunk inside one system (the common case) reads that system's tables through
        // scalar loads; a chunk across systems per lane
        const int si0 = __builtin_amdgcn_readfirstlane(si_l);
        if (__all(a >= p.N || si_l == si0)) {
          if (a < p.N) gid[u] = bin_cell(a, si0);
        } else if (a < p.N) {
          gid[u] = bin_cell(a, si_l);
        }
      }
      // the leader lane of each distinct cell reserves that cell's slots (returns unused yet)
#pragma unroll
      for (int u = 0; u < kBinU; ++u) {
        unsigned long long todo = __ballot(gid[u] >= 0);
        while (todo) {
          const int leader = __ffsll((long long)todo) - 1;
          const int g = __shfl(gid[u], leader, 64);
          const unsigned long long same = __ballot(gid[u] == g);
          if (lane == leader) ret[u] = atomicAdd(&p.count[g], __popcll(same));
          todo &= ~same;
        }
      }
      // the same walk again: each lane's slot from its leader's reservation
#pragma unroll
      for (int u = 0; u < kBinU; ++u) {
        const int a = a00 + u * gsz + lane;
        unsigned long long todo = __ballot(gid[u] >= 0);
        while (todo) {
          const int leader = __ffsll((long long)todo) - 1;
          const int g = __shfl(gid[u], leader, 64);
          const unsigned long long same = __ballot(gid[u] == g);
          const int base = __shfl(ret[u], leader, 64);
          if (gid[u] == g) {
            const int sl = base + __popcll(same & ((1ull << lane) - 1ull));
            p.slot[a] = sl;
            if (sl < p.bcap) p.bucket[(size_t)g * p.bcap + sl] = a;
            else if (p.bcap) atomicOr(&p.ints[4], 1);
          }
          todo &= ~same;
        }
      }
    }
  }
  // cells of systems not rebuilt keep their counts (added to the zero the last scan left,
  // as an atomic: the scan reads the counts across a barrier without fences)
  for (int c = gtid; c < p.ncells; c += gsz)
    if (!p.flags[p.cellsys[c]]) atomicAdd(&p.count[c], p.start[c + 1] - p.start[c]);
  // the scan needs only the counts, all of them atomics: the binning's other stores (cells,
  // slots, buckets) are published by the release of the barrier after the scan
  if (!grid_sync_atomics(p.bar, p.ints + 2)) return;
  if (blockIdx.x == 0 && tid == 0) p.tphase[1] = wall_clock64();
  // 2. exclusive scan of the counts to the cell starts, counts back to 0, in one phase: each
  //    block sums its chunk of cells and publishes the sum in one 64-bit word tagged with
  //    this phase's barrier generation (so no word needs clearing), adds the sums of the
  //    blocks before it as they appear, then scans its chunk
  const int chunk = (p.ncells + gridDim.x - 1) / gridDim.x;
  const int c0 = min((int)blockIdx.x * chunk, p.ncells), c1 = min(c0 + chunk, p.ncells);
  {
    const unsigned long long tag =
        ((unsigned long long)(__hip_atomic_load(p.bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 0x7fffffffu) +
         1ull) << 32;
    int v = 0;
    for (int c = c0 + tid; c < c1; c += bd) v += __hip_atomic_load(&p.count[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v = block_sum_int(v, wsum);
    if (tid == 0)
      __hip_atomic_store(&p.look[blockIdx.x], tag | (unsigned)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __shared__ int s_off;
    if (tid < 64) {
      int off = 0;
      for (int b0 = 0; b0 < (int)blockIdx.x; b0 += 64) {
        const int b = b0 + tid;
        if (b < (int)blockIdx.x) {
          unsigned long long w = __hip_atomic_load(&p.look[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          int spins = 0;
          for (; (w >> 32) != (tag >> 32); w = __hip_atomic_load(&p.look[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            if (++spins > kBarSpin) {  // reported like a barrier timeout
              atomicOr(p.ints + 2, 1);
              w = 0;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
          off += (int)(unsigned)w;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) off += __shfl_xor(off, o, 64);
      if (tid == 0) s_off = off;
    }
    __syncthreads();
    int off = s_off;
    for (int base = c0; base < c1; base += bd) {
      const int c = base + tid;
      const int x = c < c1 ? __hip_atomic_load(&p.count[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
      int tot;
      const int ex = block_excl_scan(x, wsum, tot);
      if (c < c1) {
        p.start[c] = off + ex;
        p.count[c] = 0;
      }
      off += tot;
    }
    if (c1 == p.ncells && c0 < c1 && tid == 0) p.start[p.ncells] = off;
  }
  if (p.bin_only) return;
  if (!grid_sync(p.bar, p.ints + 2)) return;
  if (blockIdx.x == 0 && tid == 0) p.tphase[2] = wall_clock64();
  // 3. only when some cell outgrew its bucket: atoms grouped by cell through their slots
  const bool use_bucket =
      p.bcap > 0 && __hip_atomic_load(&p.ints[4], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
  if (!use_bucket) {
    for (int a = gtid; a < p.N; a += gsz) {
      const int si = p.sysid[a];
      if (p.flags[si]) p.tmp[p.start[p.st.cell[p.spar[si]][a]] + p.slot[a]] = a;
    }
    if (!grid_sync(p.bar, p.ints + 2)) return;
  }
  if (blockIdx.x == 0 && tid == 0) p.tphase[3] = p.tphase[4] = wall_clock64();
  // the atoms of cell c in arrival order
  auto atoms_of = [&](int c) -> const int* {
    return use_bucket ? p.bucket + (size_t)c * p.bcap : p.tmp + p.start[c];
  };
  // 4. rank each cell's atoms by z (ties by original index) and gather each atom's state
  //    to its sorted slot in the other buffers; the list reference positions are the
  //    gathered positions. Cells stacked along z then form z-sorted runs of slots, which
  //    the list build scans as one window per column. Forces are not gathered: a rebuild always sits between a drift and
  //    the fused step that recomputes every force of the system before any use. When both cells of a pair have at most 32 atoms each half-wave takes one; otherwise
  //    the whole wave takes the two in turn (shuffle rank up to 64 atoms, a rank against
  //    the cell's atoms in memory beyond that). Control flow stays wave-uniform.
  {
    __shared__ double rz[kRB];
    __shared__ int rv[kRB];
    const int lane = tid & 63, half = lane >> 5, hl = lane & 31;
    const int gw = gtid >> 6, nw = gsz >> 6;
    auto gather = [&](int c, int t, int v, int si) {
      const int pc = rb_par(p, si), sc = p.spar[si];
      const double4 x = (pc ? p.posB : p.posA)[v];
      (pc ? p.posA : p.posB)[t] = x;
      if (p.st.posf[0]) {
        const int4 f = fix_pos(p.sys[si], x, p.st.ccorner[c]);
        st_fix(p.st.posf[pc ^ 1], t, p.st.fxw, f);
        st_fix(p.st.refq[sc ^ 1], t, p.st.fxw, f);
      } else {
        p.st.ref[sc ^ 1][t] = x;
      }
      st_vel(p.st.vel[sc ^ 1], t, ld_vel(p.st.vel[sc], v));
      const int u = p.st.perm[sc][v];
      p.st.perm[sc ^ 1][t] = u;
      p.st.cell[sc ^ 1][t] = c;
      if (p.islot) p.islot[u] = t;
    };
    // a half-wave's cell of pair cp: whether it is rebuilt, its start and length, and (for a
    // cell of at most 32 atoms) this lane's atom, its position (z is the key), velocity and
    // user index; the next pair's are loaded while the current one is ranked and gathered,
    // so the gather's loads are in flight during the ranking
    struct Pair {
      int si, a, len, v, u;
      double4 x, vv;
    };
    auto load_pair = [&](int cp) {
      Pair q;
      q.si = 0;
      q.a = 0;
      q.len = 0;
      q.v = 0x7fffffff;
      q.u = 0;
      q.x = make_double4(0.0, 0.0, INFINITY, 0.0);
      q.vv = make_double4(0.0, 0.0, 0.0, 0.0);
      const int cm = cp + half;
      if (cm < p.ncells) {
        q.si = p.cellsys[cm];
        const int a0 = p.start[cm], a1 = p.start[cm + 1];
        if (p.flags[q.si]) {
          q.a = a0;
          q.len = a1 - a0;
          if (hl < q.len && q.len <= 32) {
            q.v = atoms_of(cm)[hl];
            const int pc = rb_par(p, q.si), sc = p.spar[q.si];
            q.x = (pc ? p.posB : p.posA)[q.v];
            q.vv = ld_vel(p.st.vel[sc], q.v);
            q.u = p.st.perm[sc][q.v];
          }
        }
      }
      return q;
    };
    // the gather of a preloaded atom (the small-cell path)
    auto gather_pre = [&](int c, int t, const Pair& q) {
      const int pc = rb_par(p, q.si), sc = p.spar[q.si];
      (pc ? p.posA : p.posB)[t] = q.x;
      if (p.st.posf[0]) {
        const int4 f = fix_pos(p.sys[q.si], q.x, p.st.ccorner[c]);
        st_fix(p.st.posf[pc ^ 1], t, p.st.fxw, f);
        st_fix(p.st.refq[sc ^ 1], t, p.st.fxw, f);
      } else {
        p.st.ref[sc ^ 1][t] = q.x;
      }
      st_vel(p.st.vel[sc ^ 1], t, q.vv);
      p.st.perm[sc ^ 1][t] = q.u;
      p.st.cell[sc ^ 1][t] = c;
      if (p.islot) p.islot[q.u] = t;
    };
    Pair nxt = load_pair(2 * gw);
    for (int cp = 2 * gw; cp < p.ncells; cp += 2 * nw) {
      const int cm = cp + half;  // this half-wave's cell
      const Pair cur = nxt;
      if (cp + 2 * nw < p.ncells) nxt = load_pair(cp + 2 * nw);
      const int len = cur.len, a = cur.a;
      const bool on = len > 0;
      if (!__any(len > 32)) {
        const bool mine = on && hl < len;
        const int v = cur.v;
        const double zv = cur.x.z;
        // the half-wave's keys through the LDS (lane slots; the unused ones hold keys that
        // rank after every atom), read back as broadcasts up to the longer cell of the pair
        rz[tid] = zv;
        rv[tid] = v;
        __builtin_amdgcn_wave_barrier();
        const int lmax = max(__builtin_amdgcn_readlane(len, 0), __builtin_amdgcn_readlane(len, 32));
        const int hb = tid & ~31;
        int r = 0;
        for (int x = 0; x < lmax; ++x) {
          const double zx = rz[hb + x];
          const int vx = rv[hb + x];
          r += (zx < zv || (zx == zv && vx < v)) ? 1 : 0;
        }
        __builtin_amdgcn_wave_barrier();
        if (mine) gather_pre(cm, a + r, cur);
      } else {
        for (int k = 0; k < 2; ++k) {
          const int c = cp + k;
          if (c >= p.ncells || !p.flags[p.cellsys[c]]) continue;
          const int si = p.cellsys[c];
          const int ca = p.start[c], cl = p.start[c + 1] - ca;
          const int* ac = atoms_of(c);
          for (int e0 = 0; e0 < cl; e0 += 64) {
            const int e = e0 + lane;
            const bool mine = e < cl;
            const int v = mine ? ac[e] : 0x7fffffff;
            const double zv = mine ? zkey(p, v, si) : INFINITY;
            int r = 0;
            if (cl <= 64) {
              for (int x = 0; x < cl; ++x) {
                const double zx = __shfl(zv, x, 64);
                const int vx = __shfl(v, x, 64);
                r += (zx < zv || (zx == zv && vx < v)) ? 1 : 0;
              }
            } else if (mine) {
              for (int x = 0; x < cl; ++x) {
                const int vx = ac[x];
                const double zx = zkey(p, vx, si);
                r += (zx < zv || (zx == zv && vx < v)) ? 1 : 0;
              }
            }
            if (mine) gather(c, ca + r, v, si);
          }
        }
      }
    }
  }
  // 5. bookkeeping, by the last block to finish the gather (no grid barrier: no block reads
  //    what the bookkeeping writes after its own gather, and the list launch sees it across
  //    the launch boundary): parity flips commit (drift, and the gather for rebuilt
  //    systems), flags move to the list launch, rebuilt systems count one rebuild. Every
  //    parity and flag a block read has been consumed by its gather before it counts itself
  //    done, so the flips cannot overtake a read.
  __shared__ int last;
  __syncthreads();
  if (tid == 0) {
    const unsigned done = __hip_atomic_fetch_add(p.bar + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = done == gridDim.x - 1;
    if (last) __hip_atomic_store(p.bar + 2, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  if (last) {
    if (tid == 0) p.tphase[5] = wall_clock64();
    for (int s = tid; s < p.S; s += bd) {
      const int f = p.flags[s];
      p.par[s] ^= (s < p.flip_n ? 1 : 0) ^ (f ? 1 : 0);
      if (s < p.flip_n) p.ndrift[s] += 1;
      if (f) p.spar[s] ^= 1;
      p.bflags[s] = f;
      if (f) {
        p.nreb[s] += 1;
        p.flags[s] = 0;
      }
    }
    __syncthreads();
    if (tid == 0) {
      p.bflags[p.S] = 1;
      p.flags[p.S] = 0;
      p.ints[4] = 0;
      if (!use_bucket && p.bcap > 0) p.ints[5] += 1;
    }
  }
}

// Verlet lists of the bricks of the systems the rebuild launch handed over.
// eight waves per SIMD (64 VGPRs): four 512-thread bricks share a CU
template <bool W12>
__global__ void __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(8))) k_b_build(RB p) {
  if (p.bflags[p.S] == 0 || p.ints[2] != 0) return;
  const BBrick b = p.bricks[blockIdx.x];
  if (!p.bflags[b.sys]) return;
  __shared__ int rs[kMaxRegionCells + 1];
  __shared__ int hself[kMaxHomeCells];
  __shared__ int wsum[kBT / 64];
  extern __shared__ __align__(16) uint2 ldsq[];
  int mymax = 0;
  build_brick<W12>(p, b, ldsq, rs, hself, wsum, mymax);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mymax = max(mymax, __shfl_down(mymax, o, 64));
  if ((threadIdx.x & 63) == 0 && mymax > 0) atomicMax(&p.ints[0], mymax);
}

// Per-system record, one block per system (many systems).
__global__ void k_b_record(const BSys* __restrict__ sys, const double* __restrict__ partials,
                           double* __restrict__ log, int log_cap, int* __restrict__ log_count,
                           const int* __restrict__ nreb) {
  __shared__ double red[kRecP * 8];
  const int si = blockIdx.x;
  record_systems(sys + si, 1, partials, log + (size_t)si * log_cap * kRecW, log_cap, log_count + si, nreb + si, red);
}

// After a snapshot frame: its record index (that of system 0, which every system shares:
// records are taken at ticks that advance every system), then one more frame.
__global__ void k_b_snap_bump(int* __restrict__ snap_count, int* __restrict__ snap_rec, int cap,
                              const int* __restrict__ log_count) {
  if (threadIdx.x == 0) {
    const int f = snap_count[0];
    if (f < cap) snap_rec[f] = log_count[0];
    snap_count[0] = f + 1;
  }
}

// a[0..n) = v (flags of every system at set-up, rebuild counts)
__global__ void k_b_fill(int* a, int n, int v) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) a[k] = v;
}

// ------------------------------------------------------------------ fused step


// LJ pair of the engine (forces.py:127-141; full list, so half the energy per visit).
// The cutoff is a select on 1/r2 rather than a branch, so the eight pairs of a group run
// as straight-line code: w = 1/r2 for r2 < rc2 and w = 0 otherwise (also for the far
// sentinel and, with rc2 = -1, for species pairs with eps = 0), so every term below is an
// exact zero for pairs outside and nothing overflows for them. With x = sig2 w,
// sr6 = x^3 and b = sr6^2 - sr6:
//   f += (sr6^2 + b) w d            (times -24 eps),
//   u += b, ni += 1 for a pair inside (the energy, 2 eps (u - shift ni), is formed per atom;
//   the count is an integer, one add with carry per pair).
// EPS: several species, so the pair's eps multiplies f and u, and n sums 2 eps over the
// pairs inside; with one species the factors are applied once per atom.
// UNIT_SIG: the staged coordinates are in units of sigma, so x = w.
template <typename P, bool EN, bool EPS, bool UNIT_SIG>
__device__ __forceinline__ void lj_engine(P dx, P dy, P dz, P r2, P e2, P m24, P sig2, P rc2, P& fx, P& fy, P& fz,
                                          P& u, P& n, unsigned& ni) {
  P ir2;
  if constexpr (sizeof(P) == 8) ir2 = rcp64(r2);
  else ir2 = __builtin_amdgcn_rcpf(r2);
  const bool in = r2 < rc2;
  const P w = in ? ir2 : P(0);
  const P x = UNIT_SIG ? w : sig2 * w;
  const P sr6 = x * x * x;
  const P b = fma(sr6, sr6, -sr6);
  P coef = fma(sr6, sr6, b) * w;
  if constexpr (EPS) coef *= m24;
  fx = fma(coef, dx, fx);
  fy = fma(coef, dy, fy);
  fz = fma(coef, dz, fz);
  if constexpr (EN) {
    if constexpr (EPS) {
      const P e = in ? e2 : P(0);
      u = fma(e, b, u);
      n += e;
    } else {
      u += b;
      ni += in ? 1u : 0u;
    }
  }
}

// (cp: the system's DSF constants copied into registers before the pair loop; qt: the
// species charges staged in the LDS, so no pair reads global memory)
template <typename P, bool LJ_ON, bool COUL_ON, bool ONE_SP, bool EN>
__device__ __forceinline__ void b_pair(const CoulParams& cp, const P* ljt, const P* qt, int row, P qi,
                                       const typename P4<P>::T& xi, const typename P4<P>::T& xj, P& fx, P& fy,
                                       P& fz, P& u, P& n, unsigned& ni, P& cx, P& cy, P& cz, P& cu) {
  const P dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
  const int tj = ONE_SP ? 0 : (int)xj.w;
  // list membership is decided at the build, so r^2 can use fused multiply-adds here
  const P r2 = fma(dz, dz, fma(dy, dy, dx * dx));
  if constexpr (LJ_ON) {
    const int pj = row + tj;
    lj_engine<P, EN, !ONE_SP, ONE_SP && !COUL_ON>(dx, dy, dz, r2, ljt[pj], ljt[64 + pj], ljt[128 + pj],
                                                  ljt[192 + pj], fx, fy, fz, u, n, ni);
  }
  if constexpr (COUL_ON) {
    if constexpr (sizeof(P) == 8)
      coul_pair64(cp, dx, dy, dz, r2, qi * qt[tj], cx, cy, cz, cu);
    else
      coul_pair32((float)cp.alpha, (float)cp.rc, (float)cp.c, (float)cp.erc_rc, (float)cp.fshift, cp.shifted, dx, dy,
                  dz, r2, qi * qt[tj], cx, cy, cz, cu);
  }
}

// One fused (sub)step of one brick: closing half kick with the new forces, then (for
// bricks below drift_from) the next opening kick and drift into the other position
// buffer, and the Verlet-distance test; with EN (a recorded step), the brick's energy,
// momentum and angular momentum partials (the pair energies are computed only then) and,
// when snap is set, the positions and velocities of the recorded state in user order into
// frame snap_count[0] of the snapshot ring (frames of N positions then N velocities).
// Waves per SIMD the compiler must leave room for: the float32 one-species LJ step and the
// float32 Coulomb steps that merge their sums (several species, or no LJ) fit 80 VGPRs
// without spilling, so three 512-thread bricks share a CU (as their LDS regions do); the
// one-species LJ + Coulomb float32 step is held to 128 VGPRs (it spills at 80); the
// float64 forms keep the compiler's own allocation (held to 128 they spill 36-60 values).
template <typename P, bool LJ_ON, bool COUL_ON, bool ONE_SP>
constexpr int fused_waves() {
  constexpr bool merge = COUL_ON && (!LJ_ON || !ONE_SP);  // the kernel's MERGE
  if constexpr (sizeof(P) == 4) return (LJ_ON && !COUL_ON && ONE_SP) || merge ? 6 : (COUL_ON ? 4 : 1);
  return 1;
}

template <typename P, bool LJ_ON, bool COUL_ON, bool ONE_SP, bool EN, bool W12>
__global__ void __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(fused_waves<P, LJ_ON, COUL_ON, ONE_SP>())))
k_b_fused(const BSys* __restrict__ sys, const BBrick* __restrict__ bricks,
                                                 const BRegion* __restrict__ reg, const int* __restrict__ start,
                                                 double4* __restrict__ posA, double4* __restrict__ posB,
                                                 const int* __restrict__ par, BState sta,
                                                 const int* __restrict__ spar, const int* __restrict__ nn,
                                                 const uint4* __restrict__ nbr, int stride, int cap, double wc,
                                                 double wn, int drift_from, int* __restrict__ flags, int S,
                                                 double* __restrict__ partials, int* __restrict__ ints,
                                                 double* __restrict__ snap, int snap_cap,
                                                 const int* __restrict__ snap_count, int N) {
  typedef typename P4<P>::T T4;
  // one species and no Coulomb: positions are staged in units of sigma (x = 1/r2)
  constexpr bool UNIT_SIG = LJ_ON && ONE_SP && !COUL_ON;
  // Coulomb pair terms go into the LJ sums (the Coulomb constant folded into q_i) unless
  // the one-species LJ factors are applied per atom after the loop: four float and four
  // double accumulators fewer
  constexpr bool MERGE = COUL_ON && (!LJ_ON || !ONE_SP);
  constexpr bool PK3 = UNIT_SIG && sizeof(P) == 4;  // 12-byte float32 slots (no species word)
  const BBrick b = bricks[blockIdx.x];
  const BSys& s = sys[b.sys];
  __shared__ int rs[kMaxRegionCells + 1];
  __shared__ int hself[kMaxHomeCells];
  __shared__ int wsum[kBT / 64];
  __shared__ P ljt[4 * 64];
  __shared__ P qtab[64];
  __shared__ double red[kRecP * 8];
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T4* lds = reinterpret_cast<T4*>(lds_raw);
  const int si = b.sys;
  const double4* __restrict__ pos = par[si] ? posB : posA;
  double4* __restrict__ pos_next = par[si] ? posA : posB;
  const int sp = spar[si];
  double* __restrict__ vel = sta.vel[sp];
  double4* __restrict__ Fst = sta.F[sp];
  const double4* __restrict__ ref = sta.ref[sp];
  const int* __restrict__ cell = sta.cell[sp];
  // species pair tables: 2 eps, -24 eps, sigma^2, and rc^2 (-1 where eps = 0: no pair is inside)
  const bool lj0 = LJ_ON && s.lj_eps[0] > 0.0 && s.lj_rc2[0] > 0.0;
  const double unit = UNIT_SIG && lj0 ? sqrt(s.lj_sig2[0]) : 1.0;
  if constexpr (LJ_ON) {
    const int m = s.ns * s.ns;
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
      const double e = s.lj_eps[k];
      const bool on = e > 0.0 && s.lj_rc2[k] > 0.0;
      ljt[k] = (P)(2.0 * e);
      ljt[64 + k] = (P)(-24.0 * e);
      ljt[128 + k] = (P)s.lj_sig2[k];
      ljt[192 + k] = on ? (P)(UNIT_SIG ? s.lj_rc2[k] / s.lj_sig2[k] : s.lj_rc2[k]) : (P)-1;
    }
  }
  CoulParams cpar = {};
  if constexpr (COUL_ON) {
    cpar = s.coul;
    for (int k = threadIdx.x; k < s.ns; k += blockDim.x) qtab[k] = (P)s.charge[k];
  }
  const int total = region_setup(b, reg, start, rs, wsum, hself);
  // uniform: an overflowing region (or a failed rebuild) skips the work; the host reports it
  const bool ok = total + 1 <= cap && ints[2] == 0;
  if (!ok && threadIdx.x == 0 && total + 1 > cap) atomicMax(&ints[1], total + 1);
  // EN: the brick's partials, summed by thread 0 over the passes of the atom loop, so
  // that no per-thread sums stay live across the pair loop
  __shared__ double tot[kRecP];
  if (EN && threadIdx.x == 0)
    for (int k2 = 0; k2 < kRecP; ++k2) tot[k2] = 0.0;
  bool moved_far = false;
  // snapshot frame of this recorded step (uniform; none when the ring is full)
  double* frame = nullptr;
  if (EN && snap) {
    const int f = snap_count[0];
    if (f < snap_cap) frame = snap + (size_t)f * N * 6;
  }
  if (ok) {
    if constexpr (sizeof(P) == 4)
      stage_region_fix<PK3, ONE_SP>(b, s, reg, start, rs, sta.posf[par[si]], sta.ccorner, lds, s.fx_s);
    else
      stage_region<P, PK3>(b, s, reg, start, rs, pos, lds, 1.0 / unit);
    if (threadIdx.x == 0) {  // far sentinel: the padding entries of the lists point here
      T4 far;
      far.x = (P)1e15;
      far.y = (P)1e15;
      far.z = (P)1e15;
      far.w = (P)0;
      st_slot<PK3>(lds, total, far);
    }
    __syncthreads();
    const double h_cur = 0.5 * wc * s.dt, h_next = 0.5 * wn * s.dt, dt_next = wn * s.dt;
    const bool do_drift = (int)blockIdx.x < drift_from;
    const int a0 = start[b.cell0], a1 = start[b.cell0 + b.nhome];
    const int nwave = (int)blockDim.x >> 6;
    for (int pass = a0; pass < a1;) {  // uniform trip count
      // a pass whose atoms fit the block with L > 1 lanes each (the last pass of a brick
      // with more atoms than threads, or a small brick) gives each atom L lanes, up to
      // kSplit: the largest L for which 64 / L atoms per wave still cover the pass. The
      // lanes of an atom are adjacent in one wave (lanes past (64 / L) L idle) and take its
      // neighbour groups in turn, then sum their partial forces into the first of them in
      // a fixed tree order, so that the pass costs about 1/L of a full one
      // (LJ alone: L a power of two, so that (64 / L) L = 64; the general packing made the
      // argon step 4 % slower)
      int L = 1;
      if constexpr (COUL_ON) {
        while (L < kSplit && nwave * (64 / (L + 1)) >= a1 - pass) ++L;
      } else {
        while (2 * L <= kSplit && (a1 - pass) * 2 * L <= (int)blockDim.x) L *= 2;
      }
      const bool split = L > 1;
      const int apw = 64 / L, lane = (int)threadIdx.x & 63;
      const int a = !COUL_ON ? pass + (int)threadIdx.x / L
                             : (lane < apw * L ? pass + ((int)threadIdx.x >> 6) * apw + lane / L : a1);
      const int sub = lane % L, gstep = L;
      pass += COUL_ON ? nwave * apw : (int)blockDim.x;
      double acc[kRecP] = {0, 0, 0, 0, 0, 0, 0, 0};
      double Fx = 0, Fy = 0, Fz = 0, U = 0, Nin = 0, Cx = 0, Cy = 0, Cz = 0, CU = 0;
      int ti = 0;
      if (a < a1) {
      const int self = hself[cell[a] - b.cell0] + a;
      const T4 xi = ld_slot<PK3>(lds, self);
      ti = ONE_SP ? 0 : (int)xi.w;
      const int row = ti * s.ns;
      const P qi = MERGE ? (P)(s.charge[ti] * s.coul.k) : (P)s.charge[ti];
      constexpr int kE = W12 ? 10 : 8;  // entries per 16-byte group of the list
      const int ng = (nn[a] + kE - 1) / kE;
      P fx = 0, fy = 0, fz = 0, u = 0, n = 0, cx = 0, cy = 0, cz = 0, cu = 0;
      unsigned ni = 0;  // pairs inside the cutoff (one species)
      // groups of kE neighbours; the groups kAhead ahead load while this one computes (the
      // LJ step waits on its global loads, so two in flight per lane beat one; a group of
      // Coulomb pairs is long enough to cover the next one's load, and the registers of a
      // second group would spill)
      constexpr int kAhead = COUL_ON ? 1 : 2;
      uint4 cur = make_uint4(0u, 0u, 0u, 0u), nx1 = cur;
      if (sub < ng) cur = nbr[(size_t)sub * stride + a];
      if (kAhead == 2 && sub + gstep < ng) nx1 = nbr[(size_t)(sub + gstep) * stride + a];
      for (int g = sub; g < ng; g += gstep) {
        uint4 nx2 = nx1;
        if (g + kAhead * gstep < ng) nx2 = nbr[(size_t)(g + kAhead * gstep) * stride + a];
        // half a group at a time: its slot reads are in flight together
        constexpr int kC = COUL_ON ? 2 : kE / 2;
#pragma unroll
        for (int h2 = 0; h2 < kE; h2 += kC) {
          T4 xj[kC];
#pragma unroll
          for (int k = 0; k < kC; ++k) xj[k] = ld_slot<PK3>(lds, list_entry<W12>(cur, h2 + k));
#pragma unroll
          for (int k = 0; k < kC; ++k)
            if constexpr (MERGE)
              b_pair<P, LJ_ON, COUL_ON, ONE_SP, EN>(cpar, ljt, qtab, row, qi, xi, xj[k], fx, fy, fz, u, n, ni, fx, fy,
                                                    fz, u);
            else
              b_pair<P, LJ_ON, COUL_ON, ONE_SP, EN>(cpar, ljt, qtab, row, qi, xi, xj[k], fx, fy, fz, u, n, ni, cx, cy,
                                                    cz, cu);
        }
        if constexpr (sizeof(P) == 4) {  // float32 pair terms, float64 sums every group
          Fx += fx; Fy += fy; Fz += fz; U += u; Nin += n; Cx += cx; Cy += cy; Cz += cz; CU += cu;
          fx = fy = fz = u = n = cx = cy = cz = cu = 0;
        }
        if constexpr (kAhead == 2) {
          cur = nx1;
          nx1 = nx2;
        } else {
          cur = nx2;
        }
      }
      Fx += fx; Fy += fy; Fz += fz; U += u; Nin += n; Cx += cx; Cy += cy; Cz += cz; CU += cu;
      Nin += (double)ni;
      }
      if (split) {
        // Coulomb: a tree toward the atom's first lane: at distance o, lanes whose index
        // within the atom is a multiple of 2 o add the lane o above them when it belongs to
        // the atom. LJ alone (L a power of two): a butterfly, every lane of the atom ends
        // with the sum.
#pragma unroll
        for (int o = 1; o < kSplit; o <<= 1) {
          if (o >= L) continue;  // L is the same for the whole block
          const bool take = (sub & (2 * o - 1)) == 0 && sub + o < L;
          auto down = [&](double& v) {
            if constexpr (COUL_ON) {
              const double t = __shfl_down(v, o, 64);
              if (take) v += t;
            } else {
              v += __shfl_xor(v, o, 64);
            }
          };
          down(Fx);
          down(Fy);
          down(Fz);
          if constexpr (EN) {
            down(U);
            down(Nin);
          }
          if constexpr (COUL_ON && !MERGE) {
            down(Cx);
            down(Cy);
            down(Cz);
            if constexpr (EN) down(CU);
          }
        }
      }
      if (a < a1 && sub == 0) {
      if constexpr (LJ_ON) {
        if constexpr (ONE_SP) {  // one species: -24 eps (per sigma for unit-sigma coordinates), 2 eps
          const double f = -24.0 * s.lj_eps[0] / unit;
          Fx *= f;
          Fy *= f;
          Fz *= f;
          U = 2.0 * s.lj_eps[0] * (U - s.lj_shift * Nin);
        } else {
          U -= s.lj_shift * Nin;
        }
      }
      if constexpr (MERGE) {
        const double q = s.charge[ti];
        U -= s.coul.k * (q * q * s.coul.self);
      } else if constexpr (COUL_ON) {
        const double q = s.charge[ti];
        CU -= q * q * s.coul.self;
        Fx += s.coul.k * Cx;
        Fy += s.coul.k * Cy;
        Fz += s.coul.k * Cz;
        U += s.coul.k * CU;
      }
      const double m = s.mass[ti];
      const double ax = Fx / m, ay = Fy / m, az = Fz / m;
      double4 v = ld_vel(vel, a);
      {
#pragma clang fp contract(off)
        v.x = v.x + h_cur * ax;
        v.y = v.y + h_cur * ay;
        v.z = v.z + h_cur * az;
      }
      if constexpr (EN) {  // system.py:106-114: KE, P = sum m v, L = sum x cross (m v)
        const double4 x = pos[a];
        const double px = m * v.x, py = m * v.y, pz = m * v.z;
        acc[0] += U;
        acc[1] += 0.5 * m * ((v.x * v.x + v.z * v.z) + v.y * v.y);
        acc[2] += px;
        acc[3] += py;
        acc[4] += pz;
        acc[5] += x.y * pz - x.z * py;
        acc[6] += x.z * px - x.x * pz;
        acc[7] += x.x * py - x.y * px;
        if (frame) {
          const size_t u = (size_t)sta.perm[sp][a];
          frame[3 * u] = x.x;
          frame[3 * u + 1] = x.y;
          frame[3 * u + 2] = x.z;
          frame[3 * ((size_t)N + u)] = v.x;
          frame[3 * ((size_t)N + u) + 1] = v.y;
          frame[3 * ((size_t)N + u) + 2] = v.z;
        }
      }
      // the force is kept only where the next kick runs in another launch (no drift here:
      // the last step of a run, the thermostatted integrators); a drifting brick has
      // already used it for both half kicks
      if (!do_drift) Fst[a] = make_double4(Fx, Fy, Fz, 0.0);
      if (do_drift) {
        double4 x = pos[a];
        {
#pragma clang fp contract(off)
          v.x = v.x + h_next * ax;
          v.y = v.y + h_next * ay;
          v.z = v.z + h_next * az;
          x.x = x.x + dt_next * v.x;
          x.y = x.y + dt_next * v.y;
          x.z = x.z + dt_next * v.z;
        }
        x.x = np_mod2(x.x, s.L[0]);
        x.y = np_mod2(x.y, s.L[1]);
        x.z = np_mod2(x.z, s.L[2]);
        pos_next[a] = x;
        if constexpr (sizeof(P) == 4) {
          const int4 f = fix_pos(s, x, sta.ccorner[cell[a]]);
          st_fix(sta.posf[par[si] ^ 1], a, ONE_SP ? 3 : 4, f);
          moved_far |= moved_far_q(s, f, ld_fix<ONE_SP>(sta.refq[sp], a));
        } else {
          const double4 r = ref[a];
          const double d0 = minimg_fast(x.x - r.x, s.L[0], s.inv[0]);
          const double d1 = minimg_fast(x.y - r.y, s.L[1], s.inv[1]);
          const double d2 = minimg_fast(x.z - r.z, s.L[2], s.inv[2]);
          moved_far |= !(r2_np(d0, d1, d2) < s.half_skin2);
        }
      }
      st_vel(vel, a, v);
      }
      if constexpr (EN) {
        block_sum<kRecP>(acc, red);
        if (threadIdx.x == 0)
          for (int k2 = 0; k2 < kRecP; ++k2) tot[k2] += acc[k2];
      }
    }
  }
  if (__any(moved_far) && (threadIdx.x & 63) == 0) {
    atomicOr(&flags[si], 1);
    atomicOr(&flags[S], 1);
  }
  if constexpr (EN) {
    if (threadIdx.x == 0)
      for (int k2 = 0; k2 < kRecP; ++k2) partials[(size_t)blockIdx.x * kRecP + k2] = tot[k2];
  }
}

// Integrator of the engine (every system of an arena uses the same one).
enum { kVV = 0, kLangevin = 1, kBrownian = 2, kNoseHoover = 3, kRattle = 4 };
struct BInteg {
  int mode;
  double gamma;  // Langevin friction (1/s)
  double T;      // bath temperature (K)
  double tau;    // Nose-Hoover period (s)
  unsigned long long seed;
};
constexpr double kBoltzmann = 1.380649e-23;

// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC11): 10 rounds of two 32x32->64
// multiplies with Weyl key increments.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, unsigned k0, unsigned k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const unsigned long long p0 = 0xD2511F53ull * c.x, p1 = 0xCD9E8D57ull * c.z;
    c = make_uint4((unsigned)(p1 >> 32) ^ c.y ^ k0, (unsigned)p1, (unsigned)(p0 >> 32) ^ c.w ^ k1, (unsigned)p0);
  }
  return c;
}

// Uniform in (0, 1) from two words: ((a >> 5) 2^26 + (b >> 6) + 0.5) 2^-53.
__device__ __forceinline__ double u53(unsigned a, unsigned b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6) + 0.5) * 1.1102230246251565e-16;
}

// Three standard normals of atom `id` at drift number `ctr` (Box-Muller on 4 uniforms).
__device__ __forceinline__ void noise3(unsigned long long seed, long long ctr, unsigned id, double& n0, double& n1,
                                       double& n2) {
  const unsigned k0 = (unsigned)seed, k1 = (unsigned)(seed >> 32);
  const unsigned c0 = (unsigned)ctr, c1 = (unsigned)((unsigned long long)ctr >> 32);
  const uint4 w = philox4x32_10(make_uint4(c0, c1, id, 0u), k0, k1);
  const uint4 z = philox4x32_10(make_uint4(c0, c1, id, 1u), k0, k1);
  const double ua = u53(w.x, w.y), ub = u53(w.z, w.w), uc = u53(z.x, z.y);
  const double r0 = sqrt(-2.0 * log(ua)), r1 = sqrt(-2.0 * log(uc));
  const double t0 = 2.0 * M_PI * ub;
  n0 = r0 * cos(t0);
  n1 = r0 * sin(t0);
  n2 = r1 * cos(2.0 * M_PI * u53(z.z, z.w));
}

// The drift of one step for the atoms of the active systems (the first step of a run,
// and every step of the thermostatted integrators, whose fused launch stops after the
// closing kick), in the reference's operation order:
//   velocity Verlet (integrators.py:19-31): v += dt/2 a; x += dt v
//   Langevin BAOAB (integrators.py:60-71): v += dt/2 a; x += dt/2 v;
//       v = c1 v + sqrt((1 - c1^2) kB T / m) xi with c1 = exp(-gamma dt); x += dt/2 v
//   Brownian (integrators.py:250-257): dx = D F / (kB T) dt + sqrt(2 D dt) xi; x += dx;
//       v = dx / dt
//   Nose-Hoover (integrators.py:104-112): v *= scale (the thermostat half steps, from
//       k_b_nh); v += dt/2 a; x += dt v
// then the periodic wrap and the Verlet-distance test. With drift == 0 only the
// Nose-Hoover scale is applied (the last half step of a run). xi: Philox normals of the
// atom's user index at the system's drift number.
__global__ void k_b_drift(const BSys* __restrict__ sys, const int* __restrict__ sysid, int natoms,
                          double4* __restrict__ posA, double4* __restrict__ posB, const int* __restrict__ par,
                          BState sta, const int* __restrict__ spar, double wn, int* __restrict__ flags, int S,
                          BInteg ig, const long long* __restrict__ ndrift, const double* __restrict__ nh,
                          int drift) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  bool moved_far = false;
  int si = 0;
  if (a < natoms) {
    si = sysid[a];
    const BSys& s = sys[si];
    const double4* pos = par[si] ? posB : posA;
    double4* pos_next = par[si] ? posA : posB;
    const int sp = spar[si];
    double* vel = sta.vel[sp];
    const double4* ref = sta.ref[sp];
    const double4 x0 = pos[a];
    const int ti = (int)x0.w;
    const double m = s.mass[ti];
    const double4 f = sta.F[sp][a];
    const double h = 0.5 * wn * s.dt, dtn = wn * s.dt;
    double4 v = ld_vel(vel, a), x = x0;
    double n0 = 0, n1 = 0, n2 = 0;
    if (ig.mode == kLangevin || ig.mode == kBrownian) noise3(ig.seed, ndrift[si], (unsigned)sta.perm[sp][a], n0, n1, n2);
    {
#pragma clang fp contract(off)
      if (ig.mode == kNoseHoover) {
        const double s1 = nh[6 * si + 3], s2 = nh[6 * si + 4];
        v.x = v.x * s1 * s2;
        v.y = v.y * s1 * s2;
        v.z = v.z * s1 * s2;
      }
      if (drift) {
        if (ig.mode == kBrownian) {
          const double D = s.diff[ti], kT = kBoltzmann * ig.T, sq = sqrt(2.0 * D * dtn);
          const double dx = D * f.x / kT * dtn + sq * n0;
          const double dy = D * f.y / kT * dtn + sq * n1;
          const double dz = D * f.z / kT * dtn + sq * n2;
          x.x = x.x + dx;
          x.y = x.y + dy;
          x.z = x.z + dz;
          v.x = dx / dtn;
          v.y = dy / dtn;
          v.z = dz / dtn;
        } else {
          v.x = v.x + h * (f.x / m);
          v.y = v.y + h * (f.y / m);
          v.z = v.z + h * (f.z / m);
          if (ig.mode == kLangevin) {
            const double hd = 0.5 * dtn;
            x.x = x.x + hd * v.x;
            x.y = x.y + hd * v.y;
            x.z = x.z + hd * v.z;
            const double c1 = exp(-ig.gamma * dtn);
            const double sg = sqrt((1.0 - c1 * c1) * kBoltzmann * ig.T / m);
            v.x = c1 * v.x + sg * n0;
            v.y = c1 * v.y + sg * n1;
            v.z = c1 * v.z + sg * n2;
            x.x = x.x + hd * v.x;
            x.y = x.y + hd * v.y;
            x.z = x.z + hd * v.z;
          } else {
            x.x = x.x + dtn * v.x;
            x.y = x.y + dtn * v.y;
            x.z = x.z + dtn * v.z;
          }
        }
      }
    }
    st_vel(vel, a, v);
    if (drift) {
      x.x = np_mod2(x.x, s.L[0]);
      x.y = np_mod2(x.y, s.L[1]);
      x.z = np_mod2(x.z, s.L[2]);
      pos_next[a] = x;
      if (sta.posf[0]) {
        const int4 f = fix_pos(s, x, sta.ccorner[sta.cell[sp][a]]);
        st_fix(sta.posf[par[si] ^ 1], a, sta.fxw, f);
        moved_far = moved_far_q(s, f, ld_fix(sta.refq[sp], a, sta.fxw));
      } else {
        const double4 r = ref[a];
        const double d0 = minimg_fast(x.x - r.x, s.L[0], s.inv[0]);
        const double d1 = minimg_fast(x.y - r.y, s.L[1], s.inv[1]);
        const double d2 = minimg_fast(x.z - r.z, s.L[2], s.inv[2]);
        moved_far = !(r2_np(d0, d1, d2) < s.half_skin2);
      }
      if (moved_far) {
        atomicOr(&flags[si], 1);
        atomicOr(&flags[S], 1);
      }
    }
  }
}

// The Nose-Hoover half steps of one system (k_b_nh), in the reference's operation order.
__device__ void nh_update(const BSys& s, const BInteg& ig, const double* acc, double* st, int end, int start, int rec,
                          double* log, int log_cap, int* log_count, int nreb) {
#pragma clang fp contract(off)
  double xi = st[0], eta = st[1], Q = st[2];
  const double kT = kBoltzmann * ig.T, nf = 3.0 * (s.natoms - 1 > 1 ? s.natoms - 1 : 1), dt = s.dt;
  double k2 = !end && st[5] >= 0.0 ? st[5] : 2.0 * acc[1];
  double s1 = 1.0, s2 = 1.0;
  auto half = [&]() {
    if (Q == 0.0) Q = nf * kT * (ig.tau * ig.tau);
    xi = xi + 0.25 * dt * (k2 - nf * kT) / Q;
    const double sc = exp(-0.5 * dt * xi);
    eta = eta + 0.5 * dt * xi;
    xi = xi + 0.25 * dt * (k2 * (sc * sc) - nf * kT) / Q;
    k2 = k2 * (sc * sc);
    return sc;
  };
  if (end) {
    s1 = half();
    if (rec) {
      const int idx = log_count[0];
      if (idx < log_cap) {
        double* o = log + (size_t)idx * kRecW;
        o[0] = acc[0];
        o[1] = acc[1] * (s1 * s1);
        for (int k = 2; k < 5; ++k) o[k] = acc[k] * s1;
        o[5] = (double)nreb;
        for (int k = 5; k < kRecP; ++k) o[k + 1] = acc[k] * s1;
        o[9] = 0.5 * Q * (xi * xi) + nf * kT * eta;
        o[10] = 0.0;
        o[11] = 0.0;
      }
      log_count[0] = idx + 1;
    }
  }
  if (start) s2 = half();
  st[0] = xi;
  st[1] = eta;
  st[2] = Q;
  st[3] = s1;
  st[4] = s2;
  st[5] = k2;
}

// Nose-Hoover half steps of every system (integrators.py:90-112), one block per system:
// k2 = sum m v^2 is twice the kinetic energy of the brick partials of the fused launch
// that just closed the step. With `end`, the half step that ends that step (and the
// record of the step, when `rec`: kinetic energy, momentum and angular momentum scaled
// as the velocities are, and the bath energy Q xi^2 / 2 + N_f kT eta); with `start`,
// the half step that opens the next one (k2 of the scaled velocities is k2 scale^2).
// nh[6 s ..]: xi, eta, Q (0 until the first half step), the velocity scales of the two
// half steps (1 for a half step not taken; the drift applies them one after the other,
// as the reference does), and k2 after the last half step (-1 before any; a run's first,
// start-only launch takes k2 from it, or from the partials of the set-up record).
__global__ void k_b_nh(const BSys* __restrict__ sys, const double* __restrict__ partials, double* __restrict__ nh,
                       BInteg ig, int end, int start, int rec, double* __restrict__ log, int log_cap,
                       int* __restrict__ log_count, const int* __restrict__ nreb) {
  __shared__ double red[kRecP * (kBT / 64)];
  const int si = blockIdx.x;
  const BSys& s = sys[si];
  double acc[kRecP] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double* __restrict__ ps = partials + (size_t)s.brick_off * kRecP;
  for (int b = threadIdx.x; b < s.nbricks; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < kRecP; ++k) acc[k] += ps[(size_t)b * kRecP + k];
  block_sum<kRecP>(acc, red);
  if (threadIdx.x != 0) return;
  nh_update(s, ig, acc, nh + 6 * si, end, start, rec, log + (size_t)si * log_cap * kRecW, log_cap, log_count + si,
            nreb[si]);
}


// Positions (which = 0), velocities (1) or forces (2) of every slot scattered to user order.
__global__ void k_b_scatter_user(int natoms, int which, const double4* __restrict__ posA,
                                 const double4* __restrict__ posB, const int* __restrict__ par, BState sta,
                                 const int* __restrict__ spar, const int* __restrict__ sysid, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= natoms) return;
  const int si = sysid[t], sp = spar[si];
  const int u = sta.perm[sp][t];
  const double4 v = which == 0 ? (par[si] ? posB : posA)[t] : (which == 1 ? ld_vel(sta.vel[sp], t) : sta.F[sp][t]);
  out[3 * (size_t)u] = v.x;
  out[3 * (size_t)u + 1] = v.y;
  out[3 * (size_t)u + 2] = v.z;
}

// ------------------------------------------------------------------ rigid bonds

constexpr int kMaxClusterBonds = 512;  // bonds of one cluster (its solve is one thread's)

// Constraint clusters of the Rattle integrator (integrators.py:119-234): connected
// components of the bond graph, one thread each. Bonds couple only through shared atoms,
// so the reference's dense Newton system is block diagonal over these clusters. Cluster c
// has bonds [coff[c], coff[c+1]) with local atom indices la / lb and lengths len, atoms
// [aoff[c], aoff[c+1]) with global user indices uatom; clusters are ordered by system
// (those of system s are [csoff[s], csoff[s+1])). islot[u] is the current slot of user
// index u, written by every list rebuild. Work memory per cluster: 8 doubles per bond
// (old bond vector, drifted bond vector, summed multiplier), 5 per atom (displacement,
// inverse mass, slot), K^2 + K at moff[c] (the Newton matrix and its right-hand side).
struct BCons {
  int ncl;
  const int *coff, *aoff, *la, *lb, *uatom, *csys, *csoff;
  const double* len;
  const long long* moff;
  int* islot;
  double *work, *disp, *mat, *cdelta;
  int* cerr;                     // [0] 1: no convergence in max_iter, 2: stalled; [1] cluster
  double* cres;                  // [0] residual at the failure
  unsigned long long* maxviol;   // bits of the worst relative bond-length error (non-negative double)
  double tol;
  int max_iter;
};

__device__ __forceinline__ double mi_np(double d, double L) { return d - L * rint(d / L); }

// A thread's view of one work array: element i at p[i * s]. Global work memory is
// contiguous per cluster (s = 1); in the LDS the threads of a block interleave (s = 64,
// element i of every lane in one row of 64 doubles), so lanes never share a bank.
struct Scr {
  double* p;
  int s;
  __device__ __forceinline__ double& operator()(int i) const { return p[i * s]; }
};

// LDS path: per-thread doubles for clusters of at most k bonds (8 per bond, 5 per atom,
// k^2 + k for the Newton system); clusters up to kLdsBonds bonds fit 64 threads in 64 KiB.
__host__ __device__ constexpr int cluster_doubles(int k) { return 8 * k + 5 * (k + 1) + k * k + k; }
constexpr int kLdsBonds = 6;

// A x = b for a K x K row-major A (overwritten) by Gaussian elimination with partial
// pivoting; x replaces b. False for a singular A.
__device__ bool gauss_solve(Scr A, Scr b, int K) {
  for (int j = 0; j < K; ++j) {
    int p = j;
    double best = fabs(A(j * K + j));
    for (int i = j + 1; i < K; ++i)
      if (fabs(A(i * K + j)) > best) {
        best = fabs(A(i * K + j));
        p = i;
      }
    if (best == 0.0) return false;
    if (p != j) {
      for (int l = j; l < K; ++l) {
        const double t = A(j * K + l);
        A(j * K + l) = A(p * K + l);
        A(p * K + l) = t;
      }
      const double t = b(j);
      b(j) = b(p);
      b(p) = t;
    }
    const double piv = A(j * K + j);
    for (int i = j + 1; i < K; ++i) {
      const double f = A(i * K + j) / piv;
      if (f == 0.0) continue;
      for (int l = j + 1; l < K; ++l) A(i * K + l) -= f * A(j * K + l);
      b(i) -= f * b(j);
    }
  }
  for (int j = K - 1; j >= 0; --j) {
    double s = b(j);
    for (int l = j + 1; l < K; ++l) s -= A(j * K + l) * b(l);
    b(j) = s / A(j * K + j);
  }
  return true;
}

// C[k][l] of a cluster: the change of bond vector k per unit multiplier on bond l, which
// moves atom la[l] by +inv and lb[l] by -inv along that bond (integrators.py:202-205).
__device__ __forceinline__ double coupling_kl(const int* la, const int* lb, Scr D, int k, int l) {
  const int ak = la[k], bk = lb[k], al = la[l], bl = lb[l];
  const double ia = D(5 * ak + 3), ib = D(5 * bk + 3);
  return ia * ((ak == al ? 1.0 : 0.0) - (ak == bl ? 1.0 : 0.0)) - ib * ((bk == al ? 1.0 : 0.0) - (bk == bl ? 1.0 : 0.0));
}

// D[j] = inv_j sum_l s(j, l) G_l ro_l: the displacement of every atom of the cluster for
// summed multipliers G (s = +1 for the bond's first atom, -1 for its second).
__device__ void cluster_disp(const int* la, const int* lb, Scr W, Scr D, int K, int na) {
  for (int j = 0; j < na; ++j) D(5 * j) = D(5 * j + 1) = D(5 * j + 2) = 0.0;
  for (int l = 0; l < K; ++l) {
    const int a = la[l], b = lb[l];
    const double g = W(8 * l + 6), ga = g * D(5 * a + 3), gb = g * D(5 * b + 3);
    for (int d = 0; d < 3; ++d) {
      const double w = W(8 * l + d);
      D(5 * a + d) += ga * w;
      D(5 * b + d) -= gb * w;
    }
  }
}

// The work arrays of cluster cl: in the LDS (LDS: this lane's interleaved rows, laid out
// for clusters of up to lds_k bonds) or in global memory. A compile-time choice, so each
// access is an LDS or a global instruction rather than a generic one.
template <bool LDS>
__device__ __forceinline__ void cluster_scratch(const BCons& c, int cl, int k0, int a0, int lds_k, double* lds,
                                                Scr& W, Scr& D, Scr& M) {
  if constexpr (LDS) {
    double* t = lds + (threadIdx.x & 63);
    W = {t, 64};
    D = {t + 64 * 8 * lds_k, 64};
    M = {t + 64 * (8 * lds_k + 5 * (lds_k + 1)), 64};
  } else {
    W = {c.work + 8 * (size_t)k0, 1};
    D = {c.disp + 5 * (size_t)a0, 1};
    M = {c.mat + c.moff[cl], 1};
  }
}

// SHAKE of every cluster after the drift (integrators.py:197-226): the drifted positions
// (the next buffer, not yet committed) move along the old bond vectors until every bond's
// relative residual |L^2 - r^2| / L^2 is at most tol, by Newton iterations on the summed
// multipliers G: A g = L^2 - r^2 with A[k][l] = 2 C[k][l] (r_k . ro_l). Positions and
// velocities take the total displacement (velocities / dt), positions are wrapped again,
// moved atoms are checked against the list's reference positions, and the worst relative
// bond-length error is kept. A failure leaves the cluster as drifted and sets cerr.
// Returns the cluster's relative bond-length error (0 after a failure).
template <bool LDS>
__device__ double shake_cluster(const BSys* __restrict__ sys, double4* __restrict__ posA, double4* __restrict__ posB,
                                const int* __restrict__ par, const BState& sta, const int* __restrict__ spar,
                                const BCons& c, int* __restrict__ flags, int S, int lds_k, double* cl_lds, int cl) {
#pragma clang fp contract(off)
  const int si = c.csys[cl];
  const BSys& s = sys[si];
  const double4* old = par[si] ? posB : posA;
  double4* nw = par[si] ? posA : posB;
  const int sp = spar[si];
  double* vel = sta.vel[sp];
  const double4* ref = sta.ref[sp];
  const int k0 = c.coff[cl], K = c.coff[cl + 1] - k0, a0 = c.aoff[cl], na = c.aoff[cl + 1] - a0;
  const int* la = c.la + k0;
  const int* lb = c.lb + k0;
  const double* len = c.len + k0;
  Scr W, D, M;
  cluster_scratch<LDS>(c, cl, k0, a0, lds_k, cl_lds, W, D, M);
  const Scr rhs = {&M(K * K), M.s};  // right-hand side, then the Newton step
  for (int j = 0; j < na; ++j) {
    const int t = c.islot[c.uatom[a0 + j]];
    D(5 * j + 3) = 1.0 / s.mass[(int)old[t].w];
    D(5 * j + 4) = (double)t;
  }
  for (int k = 0; k < K; ++k) {
    const int ta = (int)D(5 * la[k] + 4), tb = (int)D(5 * lb[k] + 4);
    const double4 pa = old[ta], pb = old[tb], qa = nw[ta], qb = nw[tb];
    W(8 * k) = mi_np(pa.x - pb.x, s.L[0]);
    W(8 * k + 1) = mi_np(pa.y - pb.y, s.L[1]);
    W(8 * k + 2) = mi_np(pa.z - pb.z, s.L[2]);
    W(8 * k + 3) = mi_np(qa.x - qb.x, s.L[0]);
    W(8 * k + 4) = mi_np(qa.y - qb.y, s.L[1]);
    W(8 * k + 5) = mi_np(qa.z - qb.z, s.L[2]);
    W(8 * k + 6) = 0.0;
  }
  cluster_disp(la, lb, W, D, K, na);
  double best = INFINITY;
  int stall = 0, it = 0;
  for (; it < c.max_iter; ++it) {
    // bond vectors of the displaced cluster: r_k = r0_k + D[a_k] - D[b_k]
    double res = 0.0;
    for (int k = 0; k < K; ++k) {
      const int a = la[k], b = lb[k];
      const double rx = W(8 * k + 3) + (D(5 * a) - D(5 * b)), ry = W(8 * k + 4) + (D(5 * a + 1) - D(5 * b + 1)),
                   rz = W(8 * k + 5) + (D(5 * a + 2) - D(5 * b + 2));
      const double L2 = len[k] * len[k];
      rhs(k) = L2 - (rx * rx + ry * ry + rz * rz);
      res = fmax(res, fabs(rhs(k)) / L2);
      for (int l = 0; l < K; ++l)
        M(k * K + l) = 2.0 * coupling_kl(la, lb, D, k, l) * (rx * W(8 * l) + ry * W(8 * l + 1) + rz * W(8 * l + 2));
    }
    if (res <= c.tol) break;
    if (res < 0.5 * best) {
      best = res;
      stall = 0;
    } else if (++stall >= 10) {
      if (atomicCAS(&c.cerr[0], 0, 2) == 0) {
        c.cerr[1] = cl;
        c.cres[0] = res;
      }
      return 0.0;
    }
    if (!gauss_solve(M, rhs, K)) {
      if (atomicCAS(&c.cerr[0], 0, 2) == 0) {
        c.cerr[1] = cl;
        c.cres[0] = res;
      }
      return 0.0;
    }
    for (int k = 0; k < K; ++k) W(8 * k + 6) += rhs(k);
    cluster_disp(la, lb, W, D, K, na);
  }
  if (it == c.max_iter) {
    if (atomicCAS(&c.cerr[0], 0, 1) == 0) c.cerr[1] = cl;
    return 0.0;
  }
  bool far = false;
  for (int j = 0; j < na; ++j) {
    const int t = (int)D(5 * j + 4);
    double4 x = nw[t], v = ld_vel(vel, t);
    x.x = np_mod2(x.x + D(5 * j), s.L[0]);
    x.y = np_mod2(x.y + D(5 * j + 1), s.L[1]);
    x.z = np_mod2(x.z + D(5 * j + 2), s.L[2]);
    v.x = v.x + D(5 * j) / s.dt;
    v.y = v.y + D(5 * j + 1) / s.dt;
    v.z = v.z + D(5 * j + 2) / s.dt;
    nw[t] = x;
    st_vel(vel, t, v);
    D(5 * j) = x.x;  // the displacement is applied: keep the final position for the check below
    D(5 * j + 1) = x.y;
    D(5 * j + 2) = x.z;
    if (sta.posf[0]) {
      const int4 f = fix_pos(s, x, sta.ccorner[sta.cell[sp][t]]);
      st_fix(sta.posf[par[si] ^ 1], t, sta.fxw, f);
      far |= moved_far_q(s, f, ld_fix(sta.refq[sp], t, sta.fxw));
    } else {
      const double4 r = ref[t];
      const double d0 = minimg_fast(x.x - r.x, s.L[0], s.inv[0]);
      const double d1 = minimg_fast(x.y - r.y, s.L[1], s.inv[1]);
      const double d2 = minimg_fast(x.z - r.z, s.L[2], s.inv[2]);
      far |= !(r2_np(d0, d1, d2) < s.half_skin2);
    }
  }
  if (far) {
    atomicOr(&flags[si], 1);
    atomicOr(&flags[S], 1);
  }
  double dev = 0.0;
  for (int k = 0; k < K; ++k) {
    const int a = la[k], b = lb[k];
    const double rx = mi_np(D(5 * a) - D(5 * b), s.L[0]), ry = mi_np(D(5 * a + 1) - D(5 * b + 1), s.L[1]),
                 rz = mi_np(D(5 * a + 2) - D(5 * b + 2), s.L[2]);
    dev = fmax(dev, fabs(sqrt(rx * rx + ry * ry + rz * rz) - len[k]) / len[k]);
  }
  return dev;
}

// Blocks of 64 threads (one wave); LDS: work arrays in the dynamic LDS (64 cluster_doubles(lds_k)).
// The worst bond-length error of a wave takes one atomic.
template <bool LDS>
__global__ void __launch_bounds__(64) k_b_shake(const BSys* __restrict__ sys, double4* __restrict__ posA,
                                                double4* __restrict__ posB, const int* __restrict__ par, BState sta,
                                                const int* __restrict__ spar, BCons c, int* __restrict__ flags, int S,
                                                int lds_k) {
  extern __shared__ double cl_lds[];
  const int cl = blockIdx.x * blockDim.x + threadIdx.x;
  double dev = 0.0;
  if (cl < c.ncl) dev = shake_cluster<LDS>(sys, posA, posB, par, sta, spar, c, flags, S, lds_k, cl_lds, cl);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) dev = fmax(dev, __shfl_xor(dev, o, 64));
  if (threadIdx.x == 0 && dev > 0.0) atomicMax(c.maxviol, (unsigned long long)__double_as_longlong(dev));
}

// RATTLE of every cluster after the closing kick (integrators.py:227-234): one solve of
// (C[k][l] r_k . r_l) mu = r_k . (v_a - v_b) removes the bond-stretching velocities,
// v_a -= inv_a mu r, v_b += inv_b mu r per bond. With rec, the change of the recorded sums
// (kinetic energy, momentum, angular momentum, as the fused launch forms them) per cluster
// goes to cdelta, for k_b_cdelta to add to the system's partials before the record.
template <bool LDS>
__global__ void __launch_bounds__(64) k_b_rattle(const BSys* __restrict__ sys, const double4* __restrict__ posA,
                                                 const double4* __restrict__ posB, const int* __restrict__ par,
                                                 BState sta, const int* __restrict__ spar, BCons c, int rec,
                                                 int lds_k) {
#pragma clang fp contract(off)
  extern __shared__ double cl_lds[];
  const int cl = blockIdx.x * blockDim.x + threadIdx.x;
  if (cl >= c.ncl) return;
  const int si = c.csys[cl];
  const BSys& s = sys[si];
  const double4* pos = par[si] ? posB : posA;
  double* vel = sta.vel[spar[si]];
  const int k0 = c.coff[cl], K = c.coff[cl + 1] - k0, a0 = c.aoff[cl], na = c.aoff[cl + 1] - a0;
  const int* la = c.la + k0;
  const int* lb = c.lb + k0;
  Scr W, D, M;
  cluster_scratch<LDS>(c, cl, k0, a0, lds_k, cl_lds, W, D, M);
  const Scr rhs = {&M(K * K), M.s};
  for (int j = 0; j < na; ++j) {
    const int t = c.islot[c.uatom[a0 + j]];
    D(5 * j + 3) = 1.0 / s.mass[(int)pos[t].w];
    D(5 * j + 4) = (double)t;
  }
  for (int k = 0; k < K; ++k) {
    const int ta = (int)D(5 * la[k] + 4), tb = (int)D(5 * lb[k] + 4);
    const double4 pa = pos[ta], pb = pos[tb], va = ld_vel(vel, ta), vb = ld_vel(vel, tb);
    const double rx = mi_np(pa.x - pb.x, s.L[0]), ry = mi_np(pa.y - pb.y, s.L[1]), rz = mi_np(pa.z - pb.z, s.L[2]);
    W(8 * k) = rx;
    W(8 * k + 1) = ry;
    W(8 * k + 2) = rz;
    rhs(k) = rx * (va.x - vb.x) + ry * (va.y - vb.y) + rz * (va.z - vb.z);
  }
  for (int k = 0; k < K; ++k)
    for (int l = 0; l < K; ++l)
      M(k * K + l) = coupling_kl(la, lb, D, k, l) *
                     (W(8 * k) * W(8 * l) + W(8 * k + 1) * W(8 * l + 1) + W(8 * k + 2) * W(8 * l + 2));
  if (!gauss_solve(M, rhs, K)) {
    if (atomicCAS(&c.cerr[0], 0, 3) == 0) c.cerr[1] = cl;
    return;
  }
  for (int k = 0; k < K; ++k) W(8 * k + 6) = -rhs(k);  // v_a -= inv_a mu r: the displacement rule with G = -mu
  cluster_disp(la, lb, W, D, K, na);
  double acc[kRecP] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < na; ++j) {
    const int t = (int)D(5 * j + 4);
    const double4 v = ld_vel(vel, t);
    double4 u = v;
    u.x = v.x + D(5 * j);
    u.y = v.y + D(5 * j + 1);
    u.z = v.z + D(5 * j + 2);
    st_vel(vel, t, u);
    if (rec) {
      const double4 x = pos[t];
      const double m = s.mass[(int)x.w];
      const double px = m * v.x, py = m * v.y, pz = m * v.z, qx = m * u.x, qy = m * u.y, qz = m * u.z;
      acc[1] += 0.5 * m * ((u.x * u.x + u.z * u.z) + u.y * u.y) - 0.5 * m * ((v.x * v.x + v.z * v.z) + v.y * v.y);
      acc[2] += qx - px;
      acc[3] += qy - py;
      acc[4] += qz - pz;
      acc[5] += (x.y * qz - x.z * qy) - (x.y * pz - x.z * py);
      acc[6] += (x.z * qx - x.x * qz) - (x.z * px - x.x * pz);
      acc[7] += (x.x * qy - x.y * qx) - (x.x * py - x.y * px);
    }
  }
  if (rec)
    for (int k = 0; k < kRecP; ++k) c.cdelta[(size_t)cl * kRecP + k] = acc[k];
}

// Per system (one block each): the clusters' record changes summed in cluster order and
// added to the partials of the system's first brick, ahead of the record launch.
__global__ void k_b_cdelta(const BSys* __restrict__ sys, BCons c, double* __restrict__ partials) {
  __shared__ double red[kRecP * 4];
  const int si = blockIdx.x;
  double acc[kRecP] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = c.csoff[si] + threadIdx.x; q < c.csoff[si + 1]; q += blockDim.x)
#pragma unroll
    for (int k = 0; k < kRecP; ++k) acc[k] += c.cdelta[(size_t)q * kRecP + k];
  block_sum<kRecP>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < kRecP; ++k) partials[(size_t)sys[si].brick_off * kRecP + k] += acc[k];
}

// ------------------------------------------------------------------ host engine

// The dynamic-LDS limit of every instantiation of the fused kernel.
template <typename P, typename F>
void set_fused_attrs(F&& set) {
  auto one = [&](auto on, auto en) {
    constexpr bool O = decltype(on)::value, E = decltype(en)::value;
    set((const void*)k_b_fused<P, true, true, O, E, true>);
    set((const void*)k_b_fused<P, true, false, O, E, true>);
    set((const void*)k_b_fused<P, false, true, O, E, true>);
    set((const void*)k_b_fused<P, true, true, O, E, false>);
    set((const void*)k_b_fused<P, true, false, O, E, false>);
    set((const void*)k_b_fused<P, false, true, O, E, false>);
  };
  one(std::true_type{}, std::true_type{});
  one(std::true_type{}, std::false_type{});
  one(std::false_type{}, std::true_type{});
  one(std::false_type{}, std::false_type{});
}

struct SysSpec {
  int natoms = 0;
  double box[3] = {0, 0, 0};
  int ns = 1;
  double mass[kMaxSpecies] = {0}, charge[kMaxSpecies] = {0}, diff[kMaxSpecies] = {0};
  bool has_lj = false, has_coul = false;
  double eps[64], sig2[64], rc2[64], shift = 0;
  double alpha = 0, rcut = 0, erc = 0, kc = 0;
  double rc_list = 0, skin = 0;
  int bdim[3] = {2, 2, 2};
  double dt = 0;
};

struct BMD {
  int S = 0, N = 0, precision = ACCU_FP64;
  std::vector<SysSpec> spec;
  hipStream_t st = nullptr;
  // tables
  BSys* d_sys = nullptr;
  BBrick* d_bricks = nullptr;
  BRegion* d_reg = nullptr;
  int* d_cellmap = nullptr;
  int* d_cellsys = nullptr;
  int* d_sysid = nullptr;
  int ncells = 0, nbricks = 0;
  std::vector<int> brick_off, atom_off;
  // state
  double4 *pos = nullptr, *pos_b = nullptr;  // position buffers A/B
  int* par = nullptr;                         // per system: 1 when its current positions are in pos_b
  BState sta = {};                            // the rest of the per-atom state, buffers [0] / [1]
  int* spar = nullptr;                        // per system: which state buffer is current
  int *count = nullptr, *start = nullptr, *tmp = nullptr, *slot = nullptr;
  int* bucket = nullptr;  // per-cell arrival buckets of the rebuild, bcap atoms each
  int bcap = 0;
  int* nn = nullptr;
  uint4* nbr = nullptr;  // 16-byte groups of slot indices (see w12), group g of atom a at [g * stride + a]
  int stride = 0, max_nbrs = 0;  // max_nbrs: list capacity per atom, a multiple of the entries per group
  // list entries: 12-bit slot indices, 10 per 16-byte group, when every region fits 4,095
  // slots (a fifth less list traffic per step); else 16-bit, 8 per group
  bool w12 = false;
  int epg() const { return w12 ? 10 : 8; }
  int* flags = nullptr;  // [S] per system + [S] any
  int* ints = nullptr;   // [0] max neighbours, [1] region overflow, [2] rebuild barrier timeout
  int* nreb = nullptr;
  long long* ndrift = nullptr;  // per system: drifts done (noise counter of the stochastic integrators)
  double* nh = nullptr;         // per system: Nose-Hoover xi, eta, Q, velocity scales s1, s2, k2
  BInteg ig = {kVV, 0.0, 0.0, 0.0, 0ull};
  // rigid bonds (Rattle): clusters and their device buffers (freed with cbuf)
  BCons cons = {};
  int ncl = 0, lds_k = 0;
  std::vector<int> hc_coff, hc_aoff, hc_la, hc_lb, hc_uatom, hc_csys;  // host copies of the clusters
  std::vector<int> hc_user;  // cluster index as the caller passed it, per device cluster
  std::vector<double> hc_len;
  std::vector<void*> cbuf;
  int* bflags = nullptr;   // rebuild flags handed to the list launch
  unsigned* bar = nullptr;  // rebuild launch: grid barrier count and generation, blocks done with the gather
  int* bsum = nullptr;      // per-block sums of the rebuild scan
  unsigned long long* look = nullptr;  // tagged per-block sums of the rebuild scan
  long long* tphase = nullptr;  // rebuild phase timestamps
  static constexpr int kMaxGrid = 1 << 16;
  int grid_rb = 0;  // resident grid of the rebuild launch
  int* log_count = nullptr;
  double* log = nullptr;
  int log_cap = 0;
  double* partials = nullptr;
  // snapshot ring: frames of N positions then N velocities (user order) of every
  // snap_every-th record; snap_count[0] = frames taken, snap_rec[f] = record index of frame f
  double* snap = nullptr;
  int snap_cap = 0, snap_every = 0;
  int* snap_count = nullptr;
  int* snap_rec = nullptr;
  long long rec_enq = 0;  // records enqueued so far (the set-up record is record 0)
  int cap_fused = 0, cap_build = 0;  // region capacity (atoms) for the LDS
  int threads = 256;                 // workgroup size of the per-brick kernels (64..512)
  double home_mean = 0.0;            // home atoms per brick at set-up (mean over bricks)
  // The list launch takes one thread per home atom: its workgroup is the multiple of 64 just
  // above 1.1x the mean brick (a larger brick takes a second pass), at most `threads`, so
  // that small bricks do not leave most lanes idle.
  int build_threads() const {
    const int t = ((int)(1.1 * home_mean) + 63) / 64 * 64;
    return std::max(64, std::min(threads, t));
  }
  bool ready = false;
  bool timing = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<long long> graph_key;
  long long graph_recs = 0;  // records in one replay of the graph

  void release() {
    void* ps[] = {d_sys, d_bricks, d_reg, d_cellmap, d_cellsys, d_sysid, pos, pos_b, par, spar, sta.vel[0],
                  sta.vel[1], sta.F[0], sta.F[1], sta.ref[0], sta.ref[1], sta.perm[0], sta.perm[1], sta.cell[0],
                  sta.cell[1], count, start, tmp, slot, bucket, nn, nbr, flags, ints, nreb, log_count, log, partials, bflags,
                  bar, bsum, look, tphase, snap, snap_count, snap_rec, ndrift, nh, sta.posf[0], sta.posf[1],
                  sta.refq[0], sta.refq[1], (void*)sta.ccorner};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    for (void* p : cbuf) (void)hipFree(p);
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (st) (void)hipStreamDestroy(st);
  }

  // staged slot bytes of the fused launch: double4, float4, or 3 floats (one species, LJ only; Pk3)
  size_t lds_fused() const {
    const size_t slot = precision == ACCU_FP64 ? sizeof(double4)
                        : (any_lj && !any_coul && one_sp) ? sizeof(Pk3) : sizeof(float4);
    return ((size_t)cap_fused * slot + 15) / 16 * 16;
  }
  // the list launch of this engine's list format
  void (*build_kernel() const)(RB) { return w12 ? k_b_build<true> : k_b_build<false>; }
  // the list launch stages 8-byte integer slots (stage_region_q), 8 past the capacity: it
  // reads up to 7 slots beyond a column's last one and discards them, so its candidate
  // addresses need no clamp
  size_t lds_build() const { return (size_t)(cap_build + 8) * sizeof(uint2); }

  RB rebuild_args(int bin_only, int flip_n, int rec = 0) const {
    RB p;
    p.sys = d_sys;
    p.bricks = d_bricks;
    p.reg = d_reg;
    p.cellmap = d_cellmap;
    p.cellsys = d_cellsys;
    p.sysid = d_sysid;
    p.posA = pos;
    p.posB = pos_b;
    p.par = par;
    p.st = sta;
    p.spar = spar;
    p.count = count;
    p.start = start;
    p.tmp = tmp;
    p.slot = slot;
    p.bucket = bucket;
    p.bcap = bcap;
    p.bsum = bsum;
    p.look = look;
    p.nn = nn;
    p.nbr = nbr;
    p.stride = stride;
    p.ngroups = max_nbrs / epg();
    p.cap = cap_build;
    p.ints = ints;
    p.tphase = tphase;
    p.bar = bar;
    p.flags = flags;
    p.bflags = bflags;
    p.nreb = nreb;
    p.ndrift = ndrift;
    p.islot = ncl ? cons.islot : nullptr;
    p.flip_n = flip_n;
    p.rec = rec;
    p.partials = partials;
    p.log = log;
    p.log_cap = log_cap;
    p.log_count = log_count;
    p.S = S;
    p.N = N;
    p.ncells = ncells;
    p.nbricks = nbricks;
    p.bin_only = bin_only;
    return p;
  }

  // Grid of the rebuild launch: one block of kRB threads per CU, all resident at once
  // (16 waves of a CU's 28 or more; the query must allow one block).
  int resident_grid() {
    int per = 0, dev = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_b_rebuild, kRB, 0) != hipSuccess)
      return 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    if (per < 1) return 0;
    return std::min(ncu, kMaxGrid);
  }

  // Rebuild of the flagged systems (binning, sort, lists); systems [0, flip_n) drifted
  // in the step before and flip their buffer parity; rec: record that step first.
  void launch_rebuild(int flip_n, int rec = 0) {
    const RB p = rebuild_args(0, flip_n, rec);
    hipLaunchKernelGGL(k_b_rebuild, dim3(grid_rb), dim3(kRB), 0, st, p);
    hipLaunchKernelGGL(build_kernel(), dim3(nbricks), dim3(build_threads()),
                       lds_build(), st, p);
  }

  template <typename P, bool ONE, bool EN>
  void fused_t(int nb, double wc, double wn, int drift_from, bool lj, bool co, bool take) {
    const size_t sh = lds_fused();
    double* sn = take ? snap : nullptr;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(nb), dim3(threads), sh, st, d_sys, d_bricks, d_reg, start, pos, pos_b, par, sta,
                         spar, nn, nbr, stride, cap_fused, wc, wn, drift_from, flags, S, partials, ints, sn,
                         snap_cap, snap_count, N);
    };
    if (w12) {
      if (lj && co) go(k_b_fused<P, true, true, ONE, EN, true>);
      else if (lj) go(k_b_fused<P, true, false, ONE, EN, true>);
      else go(k_b_fused<P, false, true, ONE, EN, true>);
    } else {
      if (lj && co) go(k_b_fused<P, true, true, ONE, EN, false>);
      else if (lj) go(k_b_fused<P, true, false, ONE, EN, false>);
      else go(k_b_fused<P, false, true, ONE, EN, false>);
    }
  }

  bool any_lj = false, any_coul = false, one_sp = true;

  // One fused (sub)step of the first `nb` bricks; bricks below drift_from also drift
  // (their systems' buffer parity flips with the rebuild launch that follows).
  void launch_fused(int nb, double wc, double wn, int drift_from, int rec) {
    // a recorded step of every system takes a snapshot on every snap_every-th record
    const bool take = rec && snap && nb == nbricks && rec_enq % snap_every == 0;
    if (rec) ++rec_enq;
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
    const bool l = any_lj, c = any_coul;
    if (precision == ACCU_FP64) {
      if (one_sp) rec ? fused_t<double, true, true>(nb, wc, wn, drift_from, l, c, take)
                      : fused_t<double, true, false>(nb, wc, wn, drift_from, l, c, false);
      else rec ? fused_t<double, false, true>(nb, wc, wn, drift_from, l, c, take)
               : fused_t<double, false, false>(nb, wc, wn, drift_from, l, c, false);
    } else {
      if (one_sp) rec ? fused_t<float, true, true>(nb, wc, wn, drift_from, l, c, take)
                      : fused_t<float, true, false>(nb, wc, wn, drift_from, l, c, false);
      else rec ? fused_t<float, false, true>(nb, wc, wn, drift_from, l, c, take)
               : fused_t<float, false, false>(nb, wc, wn, drift_from, l, c, false);
    }
    // the frame is complete: note its record index (the record is written after this) and advance the ring
    if (take) hipLaunchKernelGGL(k_b_snap_bump, dim3(1), dim3(64), 0, st, snap_count, snap_rec, snap_cap, log_count);
    if (timing) {
      (void)hipEventRecord(ev[ev_used + 1], st);
      ev_used += 2;
    }
  }

  void launch_record() {
    hipLaunchKernelGGL(k_b_record, dim3(S), dim3(512), 0, st, d_sys, partials, log, log_cap, log_count, nreb);
  }

  // Enqueue `coarse` coarse steps of `nticks` ticks; at tick t the first active[t]
  // systems advance one (sub)step with weights wc[t] (closing kick) and wn[t] (next
  // opening kick and drift). The last tick must advance every system. Each system's
  // buffer parity flips with every drift, on the device (inside the rebuild launch).
  void launch_drift(double wn, int drift) {
    const int T = 256;
    hipLaunchKernelGGL(k_b_drift, dim3((N + T - 1) / T), dim3(T), 0, st, d_sys, d_sysid, N, pos, pos_b, par, sta, spar,
                       wn, flags, S, ig, ndrift, nh, drift);
  }

  // the SHAKE / RATTLE launches keep each cluster's work arrays in the LDS when the
  // largest cluster has at most kLdsBonds bonds (lds_k = that size), else in global memory
  size_t lds_cons() const { return lds_k ? (size_t)64 * cluster_doubles(lds_k) * sizeof(double) : 0; }

  void launch_shake() {
    hipLaunchKernelGGL(lds_k ? k_b_shake<true> : k_b_shake<false>, dim3((ncl + 63) / 64), dim3(64), lds_cons(), st,
                       d_sys, pos, pos_b, par, sta, spar, cons, flags, S, lds_k);
  }

  void launch_rattle(int rec) {
    hipLaunchKernelGGL(lds_k ? k_b_rattle<true> : k_b_rattle<false>, dim3((ncl + 63) / 64), dim3(64), lds_cons(), st,
                       d_sys, pos, pos_b, par, sta, spar, cons, rec, lds_k);
    if (rec) hipLaunchKernelGGL(k_b_cdelta, dim3(S), dim3(256), 0, st, d_sys, cons, partials);
  }

  void launch_nh(int end, int start, int rec) {
    hipLaunchKernelGGL(k_b_nh, dim3(S), dim3(kBT), 0, st, d_sys, partials, nh, ig, end, start, rec, log, log_cap,
                       log_count, nreb);
  }

  // The thermostatted integrators and Rattle (one step per coarse step, every system
  // active): the fused launch stops after the forces and the closing kick (none for
  // Brownian, whose velocities are dx/dt), the drift launch does the rest of the step,
  // Nose-Hoover adds its half steps in between (which also write its records), and
  // Rattle its velocity solve after the closing kick and its SHAKE after the drift.
  void enqueue_thermo(int coarse, int record_every) {
    const bool nose = ig.mode == kNoseHoover, rat = ig.mode == kRattle;
    const double wc = ig.mode == kBrownian ? 0.0 : 1.0;
    if (nose) launch_nh(0, 1, 0);
    launch_drift(1.0, 1);
    if (rat) launch_shake();
    launch_rebuild(S);
    for (int c = 0; c < coarse; ++c) {
      const bool last = c == coarse - 1;
      const int rec = record_every > 0 && ((c + 1) % record_every == 0 || last);
      launch_fused(nbricks, wc, 0.0, 0, nose ? 1 : rec);
      if (rat) launch_rattle(rec);
      if (nose) launch_nh(1, !last, rec);
      else if (rec) launch_record();
      if (nose) launch_drift(1.0, !last);
      else if (!last) launch_drift(1.0, 1);
      if (rat && !last) launch_shake();
      if (!last) launch_rebuild(S);
    }
  }

  void enqueue(int coarse, int nticks, const int* active, const double* wc, const double* wn, int record_every) {
    if (ig.mode != kVV) {
      enqueue_thermo(coarse, record_every);
      return;
    }
    launch_drift(wc[0], 1);
    launch_rebuild(S);
    for (int c = 0; c < coarse; ++c) {
      for (int t = 0; t < nticks; ++t) {
        const int na = active[t];
        const int nb = brick_off[na];
        const bool last = (c == coarse - 1) && (t == nticks - 1);
        const int rec = (t == nticks - 1) && record_every > 0 && ((c + 1) % record_every == 0 || c == coarse - 1);
        launch_fused(nb, wc[t], wn[t], last ? 0 : nb, rec);
        const bool inline_rec = rec && !last && S <= kRecInline;
        if (rec && !inline_rec) launch_record();
        if (!last) launch_rebuild(na, inline_rec);
      }
    }
  }
};

static int build_tables(BMD& m) {
  const int S = m.S;
  std::vector<BSys> hs(S);
  std::vector<BBrick> hb;
  std::vector<BRegion> hr;
  std::vector<int> cellmap, cellsys;
  std::vector<std::array<int, 3>> corner;  // low corner (cell coords) of every brick
  std::vector<double4> hcc;                // corner of every cell (box frame), by global cell id
  m.brick_off.assign(S + 1, 0);
  m.atom_off.assign(S + 1, 0);
  int cell_off = 0;
  for (int si = 0; si < S; ++si) {
    const SysSpec& sp = m.spec[si];
    BSys& s = hs[si];
    memset(&s, 0, sizeof s);
    for (int d = 0; d < 3; ++d) {
      s.L[d] = sp.box[d];
      s.inv[d] = 1.0 / sp.box[d];
      s.nc[d] = (int)std::floor(sp.box[d] / sp.rc_list);
      ARG_CHECK(s.nc[d] >= 3, "system %d: box %.3g is under 3 list radii (%.3g) along axis %d", si, sp.box[d],
                sp.rc_list, d);
      ARG_CHECK(sp.bdim[d] >= 1 && sp.bdim[d] <= s.nc[d], "system %d: brick dim %d outside [1, %d]", si, sp.bdim[d],
                s.nc[d]);
      s.cinv[d] = s.nc[d] / sp.box[d];
    }
    s.cell_off = cell_off;
    s.ncells = s.nc[0] * s.nc[1] * s.nc[2];
    s.atom_off = m.atom_off[si];
    s.natoms = sp.natoms;
    m.atom_off[si + 1] = m.atom_off[si] + sp.natoms;
    s.brick_off = (int)hb.size();
    s.cellmap_off = (int)cellmap.size();
    s.ns = sp.ns;
    for (int k = 0; k < sp.ns; ++k) {
      s.mass[k] = sp.mass[k];
      s.charge[k] = sp.charge[k];
      s.diff[k] = sp.diff[k];
    }
    s.has_lj = sp.has_lj;
    if (sp.has_lj)
      for (int k = 0; k < sp.ns * sp.ns; ++k) {
        s.lj_eps[k] = sp.eps[k];
        s.lj_sig2[k] = sp.sig2[k];
        s.lj_rc2[k] = sp.rc2[k];
      }
    s.lj_shift = sp.shift;
    s.has_coul = sp.has_coul;
    if (sp.has_coul) {
      CoulParams& c = s.coul;
      c.alpha = sp.alpha;
      c.rc = sp.rcut;
      c.c = 2 * sp.alpha / std::sqrt(M_PI);
      c.erc_rc = sp.erc / sp.rcut;
      c.fshift = sp.erc / (sp.rcut * sp.rcut) + c.c * std::exp(-(sp.alpha * sp.rcut) * (sp.alpha * sp.rcut)) / sp.rcut;
      c.self = sp.erc / (2 * sp.rcut) + sp.alpha / std::sqrt(M_PI);
      c.k = sp.kc;
      c.shifted = 1;
    }
    s.rc_list2 = sp.rc_list * sp.rc_list;
    s.half_skin2 = sp.skin > 0 ? (0.5 * sp.skin) * (0.5 * sp.skin) : -1.0;
    s.dt = sp.dt;
    // fixed-point copy: the fused launch's staging unit (sigma for one-species LJ without
    // Coulomb, as the kernel decides it, else metres) and the largest power of two of
    // quanta per unit that keeps a staged coordinate (half a brick plus three cells)
    // below 2^30 quanta
    {
      const bool unit_sig = m.any_lj && m.one_sp && !m.any_coul && sp.has_lj && sp.eps[0] > 0.0 && sp.rc2[0] > 0.0;
      const double unit = unit_sig ? std::sqrt(sp.sig2[0]) : 1.0;
      double ext = 0.0;
      for (int d = 0; d < 3; ++d) {
        s.hc[d] = sp.box[d] / s.nc[d];
        ext = std::max(ext, (0.5 * sp.bdim[d] + 3.0) * s.hc[d] / unit);
      }
      const int k = (int)std::floor(std::log2((double)(1 << 30) / ext));
      ARG_CHECK(k > -100 && k < 100, "system %d: no fixed-point scale for cells of %.3g", si, s.hc[0]);
      s.fx_k = k;
      s.fx_q = std::ldexp(1.0, k) / unit;
      s.half_skin2_q = s.half_skin2 > 0.0 ? s.half_skin2 * s.fx_q * s.fx_q : -1.0;
      s.fx_s = std::ldexp(1.0f, -k);
      s.fx_sm = (float)(unit * std::ldexp(1.0, -k));
      // build quanta: the coarsest power-of-two multiple of the fixed-point quantum at or
      // below 1/8192 of the largest cell edge, coarsened until a region coordinate (half a
      // brick plus one cell) stays within int16
      const double fq = 1.0 / s.fx_q, hmax = std::max(s.hc[0], std::max(s.hc[1], s.hc[2]));
      const int bdmax = std::max(sp.bdim[0], std::max(sp.bdim[1], sp.bdim[2]));
      int sh = 0;
      while (fq * std::ldexp(1.0, sh + 1) <= hmax / 8192.0) ++sh;
      while ((0.5 * bdmax + 1.0) * hmax / (fq * std::ldexp(1.0, sh)) > 30000.0) ++sh;
      const double bq = fq * std::ldexp(1.0, sh);
      ARG_CHECK(sh < 31 && 2.0 * hmax / bq < 32000.0, "system %d: no integer list quanta for cells of %.3g", si, hmax);
      s.bq_sh = sh;
      s.bq_inv = 1.0 / bq;
      s.bq_m = (float)bq;
      // a listed pair: squared quantised distance below T, which covers rc_list and the
      // rounding of both atoms (each coordinate off by at most half a quantum)
      const double rcq = std::sqrt(s.rc_list2) / bq + std::sqrt(3.0);
      s.bq_T = (int)std::ceil(rcq * rcq) + 2;
      s.bq_W = (int)std::ceil(rcq) + 1;
      // the column prune tests the rounded atom position against exact cell boxes
      const double rcp = std::sqrt(s.rc_list2) + 3.0 * std::sqrt(3.0) * bq;
      s.bq_rcp2 = (float)(rcp * rcp * (1.0 + 1e-5));
    }
    // bricks and brick-major cell numbering
    const int nbx[3] = {(s.nc[0] + sp.bdim[0] - 1) / sp.bdim[0], (s.nc[1] + sp.bdim[1] - 1) / sp.bdim[1],
                        (s.nc[2] + sp.bdim[2] - 1) / sp.bdim[2]};
    cellmap.resize(cellmap.size() + s.ncells, -1);
    int gid = cell_off;
    for (int BX = 0; BX < nbx[0]; ++BX)
      for (int BY = 0; BY < nbx[1]; ++BY)
        for (int BZ = 0; BZ < nbx[2]; ++BZ) {
          BBrick b;
          b.sys = si;
          b.cell0 = gid;
          int lo[3] = {BX * sp.bdim[0], BY * sp.bdim[1], BZ * sp.bdim[2]};
          for (int d = 0; d < 3; ++d) {
            b.hd[d] = std::min(sp.bdim[d], s.nc[d] - lo[d]);
            b.rd[d] = b.hd[d] + 2;
            b.org[d] = (lo[d] + 0.5 * b.hd[d]) * (sp.box[d] / s.nc[d]);
          }
          b.nhome = b.hd[0] * b.hd[1] * b.hd[2];
          ARG_CHECK(b.rd[0] * b.rd[1] * b.rd[2] <= kMaxRegionCells && b.nhome <= kMaxHomeCells,
                    "brick region too large");
          for (int lx = 0; lx < b.hd[0]; ++lx)
            for (int ly = 0; ly < b.hd[1]; ++ly)
              for (int lz = 0; lz < b.hd[2]; ++lz) {
                const int cx = lo[0] + lx, cy = lo[1] + ly, cz = lo[2] + lz;
                cellmap[s.cellmap_off + (cx * s.nc[1] + cy) * s.nc[2] + cz] = gid++;
                cellsys.push_back(si);
              }
          b.reg_off = (int)hr.size();
          hb.push_back(b);
          corner.push_back({lo[0], lo[1], lo[2]});
        }
    hcc.resize(gid);
    for (int cx = 0; cx < s.nc[0]; ++cx)
      for (int cy = 0; cy < s.nc[1]; ++cy)
        for (int cz = 0; cz < s.nc[2]; ++cz)
          hcc[cellmap[s.cellmap_off + (cx * s.nc[1] + cy) * s.nc[2] + cz]] =
              make_double4(cx * s.hc[0], cy * s.hc[1], cz * s.hc[2], 0.0);
    // regions (need the complete cell map of the system)
    for (int k = s.brick_off; k < (int)hb.size(); ++k) {
      BBrick& b = hb[k];
      b.reg_off = (int)hr.size();
      const int* lo = corner[k].data();
      for (int rx = 0; rx < b.rd[0]; ++rx)
        for (int ry = 0; ry < b.rd[1]; ++ry)
          for (int rz = 0; rz < b.rd[2]; ++rz) {
            int raw[3] = {lo[0] + rx - 1, lo[1] + ry - 1, lo[2] + rz - 1}, w[3], sh[3];
            for (int d = 0; d < 3; ++d) {
              sh[d] = raw[d] < 0 ? -1 : (raw[d] >= s.nc[d] ? 1 : 0);
              w[d] = raw[d] - sh[d] * s.nc[d];
            }
            BRegion r;
            r.gid = cellmap[s.cellmap_off + (w[0] * s.nc[1] + w[1]) * s.nc[2] + w[2]];
            r.shift = (sh[0] + 1) + 3 * (sh[1] + 1) + 9 * (sh[2] + 1);
            hr.push_back(r);
          }
    }
    s.nbricks = (int)hb.size() - s.brick_off;
    m.brick_off[si + 1] = (int)hb.size();
    cell_off = gid;
  }
  m.ncells = cell_off;
  m.nbricks = (int)hb.size();
  HIP_TRY(hipMalloc(&m.d_sys, S * sizeof(BSys)));
  HIP_TRY(hipMemcpy(m.d_sys, hs.data(), S * sizeof(BSys), hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&m.d_bricks, hb.size() * sizeof(BBrick)));
  HIP_TRY(hipMemcpy(m.d_bricks, hb.data(), hb.size() * sizeof(BBrick), hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&m.d_reg, hr.size() * sizeof(BRegion)));
  HIP_TRY(hipMemcpy(m.d_reg, hr.data(), hr.size() * sizeof(BRegion), hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&m.d_cellmap, cellmap.size() * sizeof(int)));
  HIP_TRY(hipMemcpy(m.d_cellmap, cellmap.data(), cellmap.size() * sizeof(int), hipMemcpyHostToDevice));
  {
    double4* cc = nullptr;
    HIP_TRY(hipMalloc(&cc, hcc.size() * sizeof(double4)));
    HIP_TRY(hipMemcpy(cc, hcc.data(), hcc.size() * sizeof(double4), hipMemcpyHostToDevice));
    m.sta.ccorner = cc;
  }
  HIP_TRY(hipMalloc(&m.d_cellsys, cellsys.size() * sizeof(int)));
  HIP_TRY(hipMemcpy(m.d_cellsys, cellsys.data(), cellsys.size() * sizeof(int), hipMemcpyHostToDevice));
  std::vector<int> sysid(m.N);
  for (int si = 0; si < S; ++si)
    for (int a = m.atom_off[si]; a < m.atom_off[si + 1]; ++a) sysid[a] = si;
  HIP_TRY(hipMalloc(&m.d_sysid, m.N * sizeof(int)));
  HIP_TRY(hipMemcpy(m.d_sysid, sysid.data(), m.N * sizeof(int), hipMemcpyHostToDevice));
  return ACCU_OK;
}

}  // namespace accu

using namespace accu;

extern "C" {

int accu_bmd_create(accu_bmd** out, int nsys, const int* natoms, const double* pos_host, const double* vel_host,
                    const int* sp_host, const double* box_host, int precision) {
  ARG_CHECK(out && natoms && pos_host && vel_host && sp_host && box_host, "null argument");
  ARG_CHECK(nsys >= 1, "need at least one system");
  ARG_CHECK(precision == ACCU_FP64 || precision == ACCU_FP32_PAIR, "unknown precision %d", precision);
  BMD* m = new BMD();
  m->S = nsys;
  m->precision = precision;
  m->spec.resize(nsys);
  long long N = 0;
  for (int s = 0; s < nsys; ++s) {
    ARG_CHECK(natoms[s] >= 2, "system %d has %d atoms", s, natoms[s]);
    m->spec[s].natoms = natoms[s];
    for (int d = 0; d < 3; ++d) m->spec[s].box[d] = box_host[3 * s + d];
    N += natoms[s];
  }
  ARG_CHECK(N < (1LL << 30), "too many atoms");
  m->N = (int)N;
  const size_t d4 = (size_t)N * sizeof(double4), i4 = (size_t)N * sizeof(int);
  void** d4s[] = {(void**)&m->pos,        (void**)&m->pos_b,      (void**)&m->sta.vel[0], (void**)&m->sta.vel[1],
                  (void**)&m->sta.F[0],   (void**)&m->sta.F[1],   (void**)&m->sta.ref[0], (void**)&m->sta.ref[1]};
  for (void** p : d4s) HIP_TRY(hipMalloc(p, d4));
  void** i4s[] = {(void**)&m->sta.perm[0], (void**)&m->sta.perm[1], (void**)&m->sta.cell[0], (void**)&m->sta.cell[1],
                  (void**)&m->tmp,         (void**)&m->slot,        (void**)&m->nn};
  for (void** p : i4s) HIP_TRY(hipMalloc(p, i4 + 64));
  if (precision == ACCU_FP32_PAIR)
    for (int k = 0; k < 2; ++k) {
      HIP_TRY(hipMalloc(&m->sta.posf[k], (size_t)N * sizeof(int4) + 64));
      HIP_TRY(hipMalloc(&m->sta.refq[k], (size_t)N * sizeof(int4) + 64));
    }
  HIP_TRY(hipMalloc(&m->par, nsys * sizeof(int) + 64));
  HIP_TRY(hipMemset(m->par, 0, nsys * sizeof(int) + 64));
  HIP_TRY(hipMalloc(&m->spar, nsys * sizeof(int) + 64));
  HIP_TRY(hipMemset(m->spar, 0, nsys * sizeof(int) + 64));
  HIP_TRY(hipMalloc(&m->flags, (nsys + 1) * sizeof(int) + 64));
  HIP_TRY(hipMemset(m->flags, 0, (nsys + 1) * sizeof(int) + 64));
  HIP_TRY(hipMalloc(&m->ints, 64));
  HIP_TRY(hipMemset(m->ints, 0, 64));
  HIP_TRY(hipMalloc(&m->nreb, nsys * sizeof(int) + 64));
  HIP_TRY(hipMemset(m->nreb, 0, nsys * sizeof(int) + 64));
  HIP_TRY(hipMalloc(&m->ndrift, nsys * sizeof(long long)));
  HIP_TRY(hipMemset(m->ndrift, 0, nsys * sizeof(long long)));
  {
    std::vector<double> nh0((size_t)nsys * 6, 0.0);
    for (int k = 0; k < nsys; ++k) {
      nh0[6 * k + 3] = 1.0;
      nh0[6 * k + 4] = 1.0;
      nh0[6 * k + 5] = -1.0;
    }
    HIP_TRY(hipMalloc(&m->nh, nh0.size() * sizeof(double)));
    HIP_TRY(hipMemcpy(m->nh, nh0.data(), nh0.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMalloc(&m->log_count, nsys * sizeof(int) + 64));
  HIP_TRY(hipMemset(m->log_count, 0, nsys * sizeof(int) + 64));
  std::vector<double4> h4(N);
  std::vector<int> perm(N);
  for (long long i = 0; i < N; ++i) {
    h4[i] = make_double4(pos_host[3 * i], pos_host[3 * i + 1], pos_host[3 * i + 2], (double)sp_host[i]);
    perm[i] = (int)i;
  }
  HIP_TRY(hipMemcpy(m->pos, h4.data(), d4, hipMemcpyHostToDevice));
  // velocities: 3 doubles per atom, slots in user order until the first sort
  HIP_TRY(hipMemcpy(m->sta.vel[0], vel_host, (size_t)N * 3 * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(m->sta.F[0], 0, d4));
  HIP_TRY(hipMemcpy(m->sta.perm[0], perm.data(), i4, hipMemcpyHostToDevice));
  HIP_TRY(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
  // the fills above ran on the null stream, which a non-blocking stream does not wait for
  HIP_TRY(hipDeviceSynchronize());
  // species indices are checked when the species table arrives
  m->spec[0].ns = 0;
  *out = reinterpret_cast<accu_bmd*>(m);
  // keep a host copy of the species for validation
  return ACCU_OK;
}

int accu_bmd_destroy(accu_bmd* h) {
  if (h) {
    BMD* m = reinterpret_cast<BMD*>(h);
    m->release();
    delete m;
  }
  return ACCU_OK;
}

int accu_bmd_set_system(accu_bmd* h, int sys, int ns, const double* mass, const double* charge, double dt,
                        double rc_list, double skin, const int* brick_dims) {
  ARG_CHECK(h && mass && brick_dims, "null argument");
  BMD& m = *reinterpret_cast<BMD*>(h);
  ARG_CHECK(sys >= 0 && sys < m.S, "bad system %d", sys);
  ARG_CHECK(ns >= 1 && ns <= kMaxSpecies, "number of species %d outside [1, %d]", ns, kMaxSpecies);
  ARG_CHECK(rc_list > 0 && skin >= 0 && dt != 0, "bad list radius, skin or dt");
  SysSpec& s = m.spec[sys];
  s.ns = ns;
  for (int k = 0; k < ns; ++k) {
    s.mass[k] = mass[k];
    s.charge[k] = charge ? charge[k] : 0.0;
  }
  s.dt = dt;
  s.rc_list = rc_list;
  s.skin = skin;
  for (int d = 0; d < 3; ++d) s.bdim[d] = brick_dims[d];
  return ACCU_OK;
}

int accu_bmd_set_lj(accu_bmd* h, int sys, int ns, const double* eps, const double* sig2, const double* rc2,
                    double shift) {
  ARG_CHECK(h && eps && sig2 && rc2, "null argument");
  BMD& m = *reinterpret_cast<BMD*>(h);
  ARG_CHECK(sys >= 0 && sys < m.S && ns == m.spec[sys].ns, "bad system or species count");
  SysSpec& s = m.spec[sys];
  s.has_lj = true;
  for (int k = 0; k < ns * ns; ++k) {
    s.eps[k] = eps[k];
    s.sig2[k] = sig2[k];
    s.rc2[k] = rc2[k];
  }
  s.shift = shift;
  return ACCU_OK;
}

int accu_bmd_set_dsf(accu_bmd* h, int sys, double alpha, double rcut, double erc, double k) {
  ARG_CHECK(h, "null handle");
  BMD& m = *reinterpret_cast<BMD*>(h);
  ARG_CHECK(sys >= 0 && sys < m.S, "bad system %d", sys);
  SysSpec& s = m.spec[sys];
  s.has_coul = true;
  s.alpha = alpha;
  s.rcut = rcut;
  s.erc = erc;
  s.kc = k;
  return ACCU_OK;
}

// Build tables, sort every system into brick order, build the lists, compute the
// forces at t = 0 and record them. slack sizes the list and LDS capacities.
static int order_clusters(BMD& m);

int accu_bmd_init(accu_bmd* h, double slack, int log_cap) {
  ARG_CHECK(h, "null handle");
  BMD& m = *reinterpret_cast<BMD*>(h);
  bool lj0 = m.spec[0].has_lj, co0 = m.spec[0].has_coul;
  int maxns = 0;
  for (int s = 0; s < m.S; ++s) {
    ARG_CHECK(m.spec[s].ns >= 1 && m.spec[s].rc_list > 0, "system %d is not configured", s);
    ARG_CHECK(m.spec[s].has_lj == lj0 && m.spec[s].has_coul == co0, "all systems must use the same interactions");
    ARG_CHECK(lj0 || co0, "no interaction set");
    maxns = std::max(maxns, m.spec[s].ns);
  }
  m.any_lj = lj0;
  m.any_coul = co0;
  m.one_sp = maxns == 1;
  m.sta.fxw = m.one_sp ? 3 : 4;  // the fused launch's ONE_SP form reads 12-byte records
  int rc = build_tables(m);
  if (rc) return rc;
  HIP_TRY(hipMalloc(&m.count, (m.ncells + 1) * sizeof(int)));
  HIP_TRY(hipMalloc(&m.start, (m.ncells + 1) * sizeof(int)));
  HIP_TRY(hipMalloc(&m.partials, (size_t)m.nbricks * kRecP * sizeof(double)));
  m.log_cap = std::max(log_cap, 4);
  HIP_TRY(hipMalloc(&m.log, (size_t)m.S * m.log_cap * kRecW * sizeof(double)));
  m.stride = (m.N + 63) / 64 * 64;
  HIP_TRY(hipMemset(m.count, 0, (m.ncells + 1) * sizeof(int)));
  HIP_TRY(hipMalloc(&m.bflags, (m.S + 1) * sizeof(int) + 64));
  HIP_TRY(hipMemset(m.bflags, 0, (m.S + 1) * sizeof(int) + 64));
  HIP_TRY(hipMalloc(&m.bar, 64));
  HIP_TRY(hipMemset(m.bar, 0, 64));
  HIP_TRY(hipMalloc(&m.bsum, BMD::kMaxGrid * sizeof(int)));
  HIP_TRY(hipMalloc(&m.look, BMD::kMaxGrid * sizeof(unsigned long long)));
  HIP_TRY(hipMemset(m.look, 0, BMD::kMaxGrid * sizeof(unsigned long long)));
  HIP_TRY(hipMalloc(&m.tphase, 16 * sizeof(long long)));
  HIP_TRY(hipMemset(m.tphase, 0, 16 * sizeof(long long)));
  // null-stream fills complete before the first launch on the engine's non-blocking stream
  // (a barrier word still being zeroed would leave the rebuild's grid barrier waiting)
  HIP_TRY(hipDeviceSynchronize());
  // first binning of every system: measure the region sizes to size the LDS
  m.grid_rb = m.resident_grid();
  ARG_CHECK(m.grid_rb > 0, "the rebuild launch (%d threads per block) does not fit on a CU", kRB);
  hipLaunchKernelGGL(k_b_fill, dim3(1), dim3(256), 0, m.st, m.flags, m.S + 1, 1);
  hipLaunchKernelGGL(k_b_rebuild, dim3(m.grid_rb), dim3(kRB), 0, m.st, m.rebuild_args(1, 0));
  LAUNCH_CHECK();
  std::vector<int> hstart(m.ncells + 1);
  HIP_TRY(hipMemcpyAsync(hstart.data(), m.start, (m.ncells + 1) * sizeof(int), hipMemcpyDeviceToHost, m.st));
  {
    int e[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(e, m.ints, sizeof e, hipMemcpyDeviceToHost, m.st));
    HIP_TRY(hipStreamSynchronize(m.st));
    ARG_CHECK(e[2] == 0, "rebuild grid barrier timed out at the first binning: %d of %d blocks arrived", e[3],
              m.grid_rb);
  }
  {  // per-cell arrival buckets: twice the fullest cell of the first binning, plus 8; a
     // rebuild in which some cell outgrows its bucket groups its atoms through tmp instead
    int cmax = 0;
    for (long long c = 0; c < m.ncells; ++c) cmax = std::max(cmax, hstart[c + 1] - hstart[c]);
    const int bcap = (2 * cmax + 8 + 7) / 8 * 8;
    const size_t bytes = (size_t)m.ncells * bcap * sizeof(int);
    if (bytes <= std::max((size_t)64 << 20, (size_t)m.N * 16 * sizeof(int))) {
      HIP_TRY(hipMalloc(&m.bucket, bytes));
      m.bcap = bcap;
    }
  }
  std::vector<BBrick> hb(m.nbricks);
  HIP_TRY(hipMemcpy(hb.data(), m.d_bricks, m.nbricks * sizeof(BBrick), hipMemcpyDeviceToHost));
  {
    double home = 0.0;
    for (auto& b : hb) home += hstart[b.cell0 + b.nhome] - hstart[b.cell0];
    m.home_mean = m.nbricks ? home / m.nbricks : 0.0;
  }
  int regmax = 0;
  {
    std::vector<BRegion> hr;
    size_t nreg = 0;
    for (auto& b : hb) nreg = std::max(nreg, (size_t)(b.reg_off + b.rd[0] * b.rd[1] * b.rd[2]));
    hr.resize(nreg);
    HIP_TRY(hipMemcpy(hr.data(), m.d_reg, nreg * sizeof(BRegion), hipMemcpyDeviceToHost));
    for (auto& b : hb) {
      int tot = 0;
      for (int r = 0; r < b.rd[0] * b.rd[1] * b.rd[2]; ++r) {
        const int c = hr[b.reg_off + r].gid;
        tot += hstart[c + 1] - hstart[c];
      }
      regmax = std::max(regmax, tot);
    }
  }
  const int cap = (int)(regmax * (1.0 + slack)) + 64;  // + 1 sentinel slot included
  hipDeviceProp_t prop;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  const size_t lds_max = prop.sharedMemPerBlock;
  m.cap_build = cap;
  m.cap_fused = cap;
  ARG_CHECK(m.lds_build() + 6144 <= lds_max && m.lds_fused() + 6144 <= lds_max && cap < 65536,
            "brick region of %d atoms does not fit the LDS (%zu bytes): use smaller bricks", cap, lds_max);
  // 12-bit entries need every slot index (the sentinel's included) below 4,096; the float64
  // Coulomb step keeps 16-bit ones (with its sums merged it no longer spills with ten pairs
  // per group, but 12-bit lists did not make the NaCl fp64 step faster: 4,141 against
  // 4,162 steps/s)
  m.w12 = cap <= 4095 && !(m.precision == ACCU_FP64 && m.any_coul);
  (void)hipFuncSetAttribute((const void*)m.build_kernel(),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)m.lds_build());
  // neighbour capacity: measure with one group of 8, then size
  m.max_nbrs = m.epg();
  HIP_TRY(hipMalloc(&m.nbr, (size_t)m.stride * sizeof(uint4)));
  hipLaunchKernelGGL(k_b_fill, dim3(1), dim3(256), 0, m.st, m.flags, m.S + 1, 1);
  m.launch_rebuild(0);
  int ints[4] = {0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(ints, m.ints, sizeof ints, hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipStreamSynchronize(m.st));
  ARG_CHECK(ints[2] == 0, "rebuild grid barrier timed out at set-up: %d of %d blocks arrived", ints[3], m.grid_rb);
  ARG_CHECK(ints[1] == 0, "region overflow at the first build (%d atoms)", ints[1]);
  const int nbrs = ((int)(ints[0] * (1.0 + slack)) + 16 + m.epg() - 1) / m.epg() * m.epg();
  HIP_TRY(hipFree(m.nbr));
  HIP_TRY(hipMalloc(&m.nbr, (size_t)m.stride * (nbrs / m.epg()) * sizeof(uint4)));
  m.max_nbrs = nbrs;
  hipLaunchKernelGGL(k_b_fill, dim3(1), dim3(256), 0, m.st, m.flags, m.S + 1, 1);
  m.launch_rebuild(0);
  // the measuring builds count as one rebuild per system
  hipLaunchKernelGGL(k_b_fill, dim3(1), dim3(256), 0, m.st, m.nreb, m.S, 1);
  // forces at t = 0, first record
  auto set_attr = [&](const void* f) {
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)m.lds_fused());
  };
  set_fused_attrs<double>(set_attr);
  set_fused_attrs<float>(set_attr);
  m.launch_fused(m.nbricks, 0.0, 0.0, 0, 1);
  m.launch_record();
  LAUNCH_CHECK();
  HIP_TRY(hipStreamSynchronize(m.st));
  if (m.ncl) {
    rc = order_clusters(m);
    if (rc) return rc;
  }
  m.ready = true;
  return ACCU_OK;
}

// After the set-up builds: the clusters of each system ordered by the slot of their first
// atom, so that the 64 clusters of a wave in the SHAKE / RATTLE launches gather from
// nearby slots (atoms stay near their cells between re-sorts, so the order stays close).
static int order_clusters(BMD& m) {
  std::vector<int> islot(m.N);
  HIP_TRY(hipMemcpy(islot.data(), m.cons.islot, (size_t)m.N * sizeof(int), hipMemcpyDeviceToHost));
  const int ncl = m.ncl;
  std::vector<int> ord(ncl);
  for (int c = 0; c < ncl; ++c) ord[c] = c;
  auto key = [&](int c) { return std::make_pair(m.hc_csys[c], islot[m.hc_uatom[m.hc_aoff[c]]]); };
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key(a) < key(b); });
  std::vector<int> coff(1, 0), aoff(1, 0), la, lb, uatom, csys, user;
  std::vector<double> len;
  std::vector<long long> moff;
  long long mtot = 0;
  for (int c : ord) {
    user.push_back(m.hc_user[c]);
    for (int k = m.hc_coff[c]; k < m.hc_coff[c + 1]; ++k) {
      la.push_back(m.hc_la[k]);
      lb.push_back(m.hc_lb[k]);
      len.push_back(m.hc_len[k]);
    }
    for (int j = m.hc_aoff[c]; j < m.hc_aoff[c + 1]; ++j) uatom.push_back(m.hc_uatom[j]);
    coff.push_back((int)la.size());
    aoff.push_back((int)uatom.size());
    csys.push_back(m.hc_csys[c]);
    const long long K = m.hc_coff[c + 1] - m.hc_coff[c];
    moff.push_back(mtot);
    mtot += K * K + K;
  }
  auto put = [&](const void* dst, const void* src, size_t bytes) {
    return hipMemcpy(const_cast<void*>(dst), src, bytes, hipMemcpyHostToDevice);
  };
  HIP_TRY(put(m.cons.coff, coff.data(), coff.size() * sizeof(int)));
  HIP_TRY(put(m.cons.aoff, aoff.data(), aoff.size() * sizeof(int)));
  HIP_TRY(put(m.cons.la, la.data(), la.size() * sizeof(int)));
  HIP_TRY(put(m.cons.lb, lb.data(), lb.size() * sizeof(int)));
  HIP_TRY(put(m.cons.len, len.data(), len.size() * sizeof(double)));
  HIP_TRY(put(m.cons.uatom, uatom.data(), uatom.size() * sizeof(int)));
  HIP_TRY(put(m.cons.csys, csys.data(), csys.size() * sizeof(int)));
  HIP_TRY(put(m.cons.moff, moff.data(), moff.size() * sizeof(long long)));
  m.hc_coff.swap(coff);
  m.hc_aoff.swap(aoff);
  m.hc_la.swap(la);
  m.hc_lb.swap(lb);
  m.hc_len.swap(len);
  m.hc_uatom.swap(uatom);
  m.hc_csys.swap(csys);
  m.hc_user.swap(user);
  return ACCU_OK;
}

// coarse steps of nticks ticks each (see BMD::enqueue). use_graph captures one block.
int accu_bmd_set_integrator(accu_bmd* h, int mode, double temperature, double friction, double period,
                            unsigned long long seed) {
  ARG_CHECK(h, "null handle");
  BMD& m = *reinterpret_cast<BMD*>(h);
  ARG_CHECK(mode >= kVV && mode <= kRattle, "unknown integrator mode %d", mode);
  ARG_CHECK(mode != kRattle || m.ncl > 0, "Rattle needs constraints (accu_bmd_set_constraints)");
  ARG_CHECK(mode == kVV || temperature >= 0, "negative temperature");
  ARG_CHECK(mode != kBrownian || temperature > 0, "Brownian dynamics needs a positive temperature");
  ARG_CHECK(mode != kLangevin || friction >= 0, "negative friction");
  ARG_CHECK(mode != kNoseHoover || (period > 0 && temperature > 0), "Nose-Hoover needs a positive period and temperature");
  m.ig.mode = mode;
  m.ig.T = temperature;
  m.ig.gamma = friction;
  m.ig.tau = period;
  m.ig.seed = seed;
  if (m.gexec) {  // the captured schedule depends on the integrator
    (void)hipGraphExecDestroy(m.gexec);
    (void)hipGraphDestroy(m.graph);
    m.gexec = nullptr;
    m.graph = nullptr;
    m.graph_key.clear();
  }
  return ACCU_OK;
}

int accu_bmd_set_diffusion(accu_bmd* h, int sys, int ns, const double* D) {
  ARG_CHECK(h && D, "null argument");
  BMD& m = *reinterpret_cast<BMD*>(h);
  ARG_CHECK(!m.ready, "set diffusion coefficients before accu_bmd_init");
  ARG_CHECK(sys >= 0 && sys < m.S, "bad system %d", sys);
  ARG_CHECK(ns == m.spec[sys].ns, "system %d has %d species, not %d", sys, m.spec[sys].ns, ns);
  for (int k = 0; k < ns; ++k) {
    ARG_CHECK(D[k] >= 0, "negative diffusion coefficient");
    m.spec[sys].diff[k] = D[k];
  }
  return ACCU_OK;
}

int accu_bmd_get_thermostat(accu_bmd* h, double* out) {
  ARG_CHECK(h && out, "null argument");
  BMD& m = *reinterpret_cast<BMD*>(h);
  HIP_TRY(hipMemcpyAsync(out, m.nh, (size_t)m.S * 6 * sizeof(double), hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipStreamSynchronize(m.st));
  return ACCU_OK;
}

int accu_bmd_set_constraints(accu_bmd* h, int ncl, const int* coff, const int* aoff, const int* la, const int* lb,
                             const double* len, const int* uatom, const int* csys, double tol, int max_iter) {
  ARG_CHECK(h && coff && aoff && la && lb && len && uatom && csys, "null argument");
  BMD& m = *reinterpret_cast<BMD*>(h);
  ARG_CHECK(!m.ready, "set constraints before accu_bmd_init");
  ARG_CHECK(m.ncl == 0, "constraints are already set");
  ARG_CHECK(ncl >= 1, "no constraint clusters");
  ARG_CHECK(tol > 0 && max_iter >= 1, "bad tolerance or iteration limit");
  ARG_CHECK(coff[0] == 0 && aoff[0] == 0, "offsets must start at 0");
  std::vector<long long> sys_lo(m.S + 1, 0);
  for (int s = 0; s < m.S; ++s) sys_lo[s + 1] = sys_lo[s] + m.spec[s].natoms;
  std::vector<int> csoff(m.S + 1, 0);
  std::vector<long long> moff(ncl);
  std::vector<char> seen(m.N, 0);
  long long mtot = 0;
  int kmax = 0;
  for (int c = 0; c < ncl; ++c) {
    const int K = coff[c + 1] - coff[c], na = aoff[c + 1] - aoff[c], s = csys[c];
    ARG_CHECK(K >= 1 && K <= kMaxClusterBonds, "cluster %d has %d bonds (1 to %d are supported)", c, K,
              kMaxClusterBonds);
    ARG_CHECK(na >= 2 && na <= K + 1, "cluster %d: %d atoms for %d bonds", c, na, K);
    ARG_CHECK(s >= 0 && s < m.S && (c == 0 || s >= csys[c - 1]), "clusters must be ordered by system");
    for (int j = aoff[c]; j < aoff[c + 1]; ++j) {
      const int u = uatom[j];
      ARG_CHECK(u >= sys_lo[s] && u < sys_lo[s + 1], "cluster %d: atom %d is not in system %d", c, u, s);
      ARG_CHECK(!seen[u], "atom %d is in two clusters", u);
      seen[u] = 1;
    }
    for (int k = coff[c]; k < coff[c + 1]; ++k)
      ARG_CHECK(la[k] >= 0 && la[k] < na && lb[k] >= 0 && lb[k] < na && la[k] != lb[k] && len[k] > 0,
                "cluster %d: bad bond %d", c, k);
    moff[c] = mtot;
    mtot += (long long)K * K + K;
    kmax = std::max(kmax, K);
    for (int q = s + 1; q <= m.S; ++q) csoff[q] = c + 1;
  }
  const int nb = coff[ncl], na = aoff[ncl];
  auto up = [&](const void* src, size_t bytes) -> void* {
    void* d = nullptr;
    if (hipMalloc(&d, bytes + 64) != hipSuccess) return nullptr;
    m.cbuf.push_back(d);
    if (src && hipMemcpy(d, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    if (!src && hipMemset(d, 0, bytes + 64) != hipSuccess) return nullptr;
    return d;
  };
  std::vector<int> ident(m.N);
  for (int i = 0; i < m.N; ++i) ident[i] = i;
  BCons& cs = m.cons;
  cs.ncl = ncl;
  cs.coff = (const int*)up(coff, (ncl + 1) * sizeof(int));
  cs.aoff = (const int*)up(aoff, (ncl + 1) * sizeof(int));
  cs.la = (const int*)up(la, nb * sizeof(int));
  cs.lb = (const int*)up(lb, nb * sizeof(int));
  cs.len = (const double*)up(len, nb * sizeof(double));
  cs.uatom = (const int*)up(uatom, na * sizeof(int));
  cs.csys = (const int*)up(csys, ncl * sizeof(int));
  cs.csoff = (const int*)up(csoff.data(), (m.S + 1) * sizeof(int));
  cs.moff = (const long long*)up(moff.data(), ncl * sizeof(long long));
  cs.islot = (int*)up(ident.data(), m.N * sizeof(int));
  cs.work = (double*)up(nullptr, (size_t)nb * 8 * sizeof(double));
  cs.disp = (double*)up(nullptr, (size_t)na * 5 * sizeof(double));
  cs.mat = (double*)up(nullptr, (size_t)mtot * sizeof(double));
  cs.cdelta = (double*)up(nullptr, (size_t)ncl * kRecP * sizeof(double));
  cs.cerr = (int*)up(nullptr, 2 * sizeof(int));
  cs.cres = (double*)up(nullptr, sizeof(double));
  cs.maxviol = (unsigned long long*)up(nullptr, sizeof(unsigned long long));
  for (void* p : m.cbuf)
    if (!p) {
      set_error("device memory for %d constraint clusters", ncl);
      return ACCU_ERUNTIME;
    }
  cs.tol = tol;
  cs.max_iter = max_iter;
  m.ncl = ncl;
  m.hc_coff.assign(coff, coff + ncl + 1);
  m.hc_aoff.assign(aoff, aoff + ncl + 1);
  m.hc_la.assign(la, la + nb);
  m.hc_lb.assign(lb, lb + nb);
  m.hc_len.assign(len, len + nb);
  m.hc_uatom.assign(uatom, uatom + na);
  m.hc_csys.assign(csys, csys + ncl);
  m.hc_user.resize(ncl);
  for (int c = 0; c < ncl; ++c) m.hc_user[c] = c;
  m.lds_k = kmax <= kLdsBonds ? kmax : 0;
  if (m.lds_k) {
    (void)hipFuncSetAttribute((const void*)k_b_shake<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)m.lds_cons());
    (void)hipFuncSetAttribute((const void*)k_b_rattle<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)m.lds_cons());
  }
  HIP_TRY(hipDeviceSynchronize());  // null-stream uploads and fills before the engine stream uses them
  return ACCU_OK;
}

int accu_bmd_get_constraints(accu_bmd* h, double* out, int* err) {
  ARG_CHECK(h && out && err, "null argument");
  BMD& m = *reinterpret_cast<BMD*>(h);
  ARG_CHECK(m.ncl > 0, "no constraints are set");
  unsigned long long bits = 0;
  HIP_TRY(hipMemcpyAsync(&bits, m.cons.maxviol, sizeof bits, hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipMemcpyAsync(out + 1, m.cons.cres, sizeof(double), hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipMemcpyAsync(err, m.cons.cerr, 2 * sizeof(int), hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipStreamSynchronize(m.st));
  memcpy(out, &bits, sizeof(double));
  // the kernels see the clusters in slot order (order_clusters); report the caller's index
  if (err[0] != 0 && err[1] >= 0 && err[1] < (int)m.hc_user.size()) err[1] = m.hc_user[err[1]];
  return ACCU_OK;
}

int accu_bmd_run(accu_bmd* h, int coarse, int nticks, const int* active, const double* wc, const double* wn,
                 int record_every, int use_graph) {
  ARG_CHECK(h && active && wc && wn, "null argument");
  BMD& m = *reinterpret_cast<BMD*>(h);
  ARG_CHECK(m.ready, "call accu_bmd_init first");
  ARG_CHECK(coarse >= 1 && nticks >= 1 && nticks <= 16, "bad step counts");
  ARG_CHECK(m.ig.mode == kVV || (nticks == 1 && wc[0] == 1.0 && wn[0] == 1.0),
            "the thermostatted integrators advance every system one step of dt per coarse step");
  ARG_CHECK(!((m.ig.mode == kNoseHoover || m.ig.mode == kRattle) && m.snap),
            "snapshot frames are not taken with the Nose-Hoover and Rattle integrators");
  for (int t = 0; t < nticks; ++t) ARG_CHECK(active[t] >= 1 && active[t] <= m.S, "bad active count at tick %d", t);
  ARG_CHECK(active[nticks - 1] == m.S, "the last tick must advance every system");
  std::vector<int> cnt(m.S);
  HIP_TRY(hipMemcpyAsync(cnt.data(), m.log_count, m.S * sizeof(int), hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipStreamSynchronize(m.st));
  const int nrec = record_every > 0 ? coarse / record_every + 2 : 0;
  for (int s = 0; s < m.S; ++s) ARG_CHECK(cnt[s] + nrec <= m.log_cap, "energy log of system %d is full", s);
  if (!use_graph) {
    m.enqueue(coarse, nticks, active, wc, wn, record_every);
    LAUNCH_CHECK();
    return ACCU_OK;
  }
  std::vector<long long> key = {coarse, nticks, record_every};
  for (int t = 0; t < nticks; ++t) {
    key.push_back(active[t]);
    key.push_back((long long)(wc[t] * 1e15));
    key.push_back((long long)(wn[t] * 1e15));
  }
  // which records of the block take snapshots is fixed at capture: replay only in the same phase
  key.push_back(m.snap ? m.rec_enq % m.snap_every : -1);
  if (!(m.gexec && key == m.graph_key)) {
    if (m.gexec) HIP_TRY(hipGraphExecDestroy(m.gexec));
    if (m.graph) HIP_TRY(hipGraphDestroy(m.graph));
    m.gexec = nullptr;
    m.graph = nullptr;
    const bool tm = m.timing;
    m.timing = false;
    const long long rec0 = m.rec_enq;
    HIP_TRY(hipStreamBeginCapture(m.st, hipStreamCaptureModeThreadLocal));
    m.enqueue(coarse, nticks, active, wc, wn, record_every);
    hipError_t e = hipStreamEndCapture(m.st, &m.graph);
    m.timing = tm;
    m.graph_recs = m.rec_enq - rec0;
    m.rec_enq = rec0;
    if (e != hipSuccess) {
      set_error("graph capture failed: %s", hipGetErrorString(e));
      return ACCU_ERUNTIME;
    }
    HIP_TRY(hipGraphInstantiate(&m.gexec, m.graph, nullptr, nullptr, 0));
    m.graph_key = key;
  }
  HIP_TRY(hipGraphLaunch(m.gexec, m.st));
  m.rec_enq += m.graph_recs;
  return ACCU_OK;
}

int accu_bmd_sync(accu_bmd* h) {
  ARG_CHECK(h, "null handle");
  HIP_TRY(hipStreamSynchronize(reinterpret_cast<BMD*>(h)->st));
  return ACCU_OK;
}

// State in user order (host [N][3], any may be NULL), per-system logs
// log_host[S][log_max][8] and stats[S][4] = rebuilds, records, 0, 0; info[4] =
// max neighbours seen, capacity, region capacity, overflowing region (0 = none).
int accu_bmd_get(accu_bmd* h, double* pos_host, double* vel_host, double* F_host, double* log_host, int log_max,
                 long long* stats, long long* info) {
  ARG_CHECK(h, "null handle");
  BMD& m = *reinterpret_cast<BMD*>(h);
  const int N = m.N;
  double* tmp = nullptr;
  HIP_TRY(hipMallocAsync((void**)&tmp, (size_t)N * 3 * sizeof(double), m.st));
  double* dsts[3] = {pos_host, vel_host, F_host};
  for (int k = 0; k < 3; ++k) {
    if (!dsts[k]) continue;
    hipLaunchKernelGGL(k_b_scatter_user, dim3((N + 255) / 256), dim3(256), 0, m.st, N, k, m.pos, m.pos_b, m.par,
                       m.sta, m.spar, m.d_sysid, tmp);
    HIP_TRY(hipMemcpyAsync(dsts[k], tmp, (size_t)N * 3 * sizeof(double), hipMemcpyDeviceToHost, m.st));
  }
  std::vector<int> nreb(m.S), lc(m.S);
  int ints[6];
  HIP_TRY(hipMemcpyAsync(nreb.data(), m.nreb, m.S * sizeof(int), hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipMemcpyAsync(lc.data(), m.log_count, m.S * sizeof(int), hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipMemcpyAsync(ints, m.ints, sizeof ints, hipMemcpyDeviceToHost, m.st));
  HIP_TRY(hipStreamSynchronize(m.st));
  HIP_TRY(hipFreeAsync(tmp, m.st));
  if (log_host) {
    std::vector<double> lg((size_t)m.S * m.log_cap * kRecW);
    HIP_TRY(hipMemcpy(lg.data(), m.log, lg.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int s = 0; s < m.S; ++s)
      for (int r = 0; r < std::min(log_max, std::min(lc[s], m.log_cap)); ++r)
        memcpy(log_host + ((size_t)s * log_max + r) * kRecW, lg.data() + ((size_t)s * m.log_cap + r) * kRecW,
               kRecW * sizeof(double));
  }
  if (stats)
    for (int s = 0; s < m.S; ++s) {
      stats[4 * s] = nreb[s];
      stats[4 * s + 1] = std::min(lc[s], m.log_cap);
      stats[4 * s + 2] = 0;
      stats[4 * s + 3] = 0;
    }
  if (info) {
    info[0] = ints[0];
    info[1] = m.max_nbrs;
    info[2] = m.cap_fused;
    info[3] = ints[1];
    // [4] = current list entries (both directions) over all systems
    std::vector<int> nn(N);
    HIP_TRY(hipMemcpy(nn.data(), m.nn, (size_t)N * sizeof(int), hipMemcpyDeviceToHost));
    long long tot = 0;
    for (int v : nn) tot += v;
    info[4] = tot;
    // [5..9] phase ends of the last rebuild, [12] wall-clock rate (kHz)
    long long tp[6];
    HIP_TRY(hipMemcpy(tp, m.tphase, sizeof tp, hipMemcpyDeviceToHost));
    for (int k = 1; k < 6; ++k) info[4 + k] = tp[k] - tp[0];
    info[10] = ints[5];  // rebuilds in which some cell outgrew its bucket
    info[11] = m.bcap;
    int dev = 0, khz = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    info[12] = khz;
  }
  if (ints[2] != 0) {
    set_error("rebuild grid barrier timed out: %d of %d blocks arrived", ints[3], m.grid_rb);
    return ACCU_ERUNTIME;
  }
  if (ints[1] > 0) {
    set_error("brick region overflow: %d atoms > LDS capacity %d (raise the slack or use smaller bricks)", ints[1],
              m.cap_fused);
    return ACCU_EOVERFLOW;
  }
  if (ints[0] > m.max_nbrs) {
    set_error("neighbour list overflow: %d > capacity %d (raise the slack)", ints[0], m.max_nbrs);
    return ACCU_EOVERFLOW;
  }
  return ACCU_OK;
}

// Workgroup size of the per-brick kernels (a multiple of 64 in [64, 512]).
int accu_bmd_set_threads(accu_bmd* h, int threads) {
  ARG_CHECK(h, "null handle");
  ARG_CHECK(threads >= 64 && threads <= kBT && threads % 64 == 0, "threads %d not a multiple of 64 in [64, %d]",
            threads, kBT);
  BMD& m = *reinterpret_cast<BMD*>(h);
  m.threads = threads;
  if (m.gexec) {
    (void)hipGraphExecDestroy(m.gexec);
    (void)hipGraphDestroy(m.graph);
    m.gexec = nullptr;
    m.graph = nullptr;
    m.graph_key.clear();
  }
  return ACCU_OK;
}

int accu_bmd_set_snapshots(accu_bmd* h, int cap_frames, int every) {
  ARG_CHECK(h, "null handle");
  ARG_CHECK(cap_frames >= 0 && every >= 1, "bad snapshot ring (capacity %d, every %d)", cap_frames, every);
  BMD& m = *reinterpret_cast<BMD*>(h);
  HIP_TRY(hipStreamSynchronize(m.st));
  for (void* p : {(void*)m.snap, (void*)m.snap_count, (void*)m.snap_rec})
    if (p) HIP_TRY(hipFree(p));
  m.snap = nullptr;
  m.snap_count = m.snap_rec = nullptr;
  m.snap_cap = 0;
  m.snap_every = every;
  if (cap_frames > 0) {
    HIP_TRY(hipMalloc(&m.snap, (size_t)cap_frames * m.N * 6 * sizeof(double)));
    HIP_TRY(hipMalloc(&m.snap_count, 64));
    HIP_TRY(hipMalloc(&m.snap_rec, (size_t)cap_frames * sizeof(int)));
    HIP_TRY(hipMemset(m.snap_count, 0, 64));
    HIP_TRY(hipMemset(m.snap_rec, 0xff, (size_t)cap_frames * sizeof(int)));
    HIP_TRY(hipDeviceSynchronize());  // null-stream fills before the engine stream's next step
    m.snap_cap = cap_frames;
  }
  // the record phase restarts at the next record; a captured graph holds the old ring
  m.rec_enq = 0;
  if (m.gexec) {
    (void)hipGraphExecDestroy(m.gexec);
    (void)hipGraphDestroy(m.graph);
    m.gexec = nullptr;
    m.graph = nullptr;
    m.graph_key.clear();
  }
  return ACCU_OK;
}

int accu_bmd_get_snapshots(accu_bmd* h, int first, int count, double* pos_host, double* vel_host, int* rec_host,
                           int* taken_host) {
  ARG_CHECK(h, "null handle");
  BMD& m = *reinterpret_cast<BMD*>(h);
  int taken = 0;
  if (m.snap_count) {
    HIP_TRY(hipMemcpyAsync(&taken, m.snap_count, sizeof(int), hipMemcpyDeviceToHost, m.st));
    HIP_TRY(hipStreamSynchronize(m.st));
  }
  if (taken_host) *taken_host = taken;
  if (count == 0) return ACCU_OK;
  const int stored = std::min(taken, m.snap_cap);
  ARG_CHECK(first >= 0 && count > 0 && first + count <= stored, "frames [%d, %d) not stored (%d stored)", first,
            first + count, stored);
  const size_t n3 = (size_t)m.N * 3;
  for (int f = 0; f < count; ++f) {
    const double* src = m.snap + (size_t)(first + f) * n3 * 2;
    if (pos_host) HIP_TRY(hipMemcpy(pos_host + f * n3, src, n3 * sizeof(double), hipMemcpyDeviceToHost));
    if (vel_host) HIP_TRY(hipMemcpy(vel_host + f * n3, src + n3, n3 * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (rec_host) HIP_TRY(hipMemcpy(rec_host, m.snap_rec + first, count * sizeof(int), hipMemcpyDeviceToHost));
  return ACCU_OK;
}

int accu_bmd_snapshot_frames(accu_bmd* h, void** frames_out, int* stored_out) {
  ARG_CHECK(h && frames_out && stored_out, "null argument");
  BMD& m = *reinterpret_cast<BMD*>(h);
  int taken = 0;
  if (m.snap_count) {
    HIP_TRY(hipMemcpyAsync(&taken, m.snap_count, sizeof(int), hipMemcpyDeviceToHost, m.st));
    HIP_TRY(hipStreamSynchronize(m.st));
  }
  *frames_out = m.snap;
  *stored_out = std::min(taken, m.snap_cap);
  return ACCU_OK;
}

int accu_bmd_set_timing(accu_bmd* h, int enable) {
  ARG_CHECK(h, "null handle");
  BMD& m = *reinterpret_cast<BMD*>(h);
  m.timing = enable != 0;
  m.ev_used = 0;
  return ACCU_OK;
}

int accu_bmd_timing(accu_bmd* h, double* total_ms, long long* launches) {
  ARG_CHECK(h && total_ms && launches, "null argument");
  BMD& m = *reinterpret_cast<BMD*>(h);
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
