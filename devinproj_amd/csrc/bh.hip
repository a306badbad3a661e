// Barnes-Hut gravity (tree.py:19-78): octree built on the host exactly as the
// reference recursion builds it (root cube from the bounding box with
// half * (1 + 1e-9), octant bits (pos > center) . [1, 2, 4], split while more than
// leaf_size particles and depth < 48), then walked on the GPU.
//
// Nodes are stored in depth-first pre-order, so a node's first child is node + 1 and
// the node after its subtree is next[node]: the walk needs no stack. The set of nodes
// a target accepts is the reference's (accept = target outside the node box and
// (2 half)^2 < theta^2 r^2 with softened r^2); only the summation order differs.
// Optional quadrupole terms (traceless Q about the centre of mass, softened radius)
// extend the reference's monopole expansion.
#include "common.h"
#include "../../include/accusim_hip.h"

#include <cmath>
#include <vector>

namespace accu {

struct BHTree {
  std::vector<double4> com;     // com x, y, z, mass
  std::vector<double4> cube;    // center x, y, z, half
  std::vector<int> next;        // node after this subtree (-1 at the end)
  std::vector<int> leaf;        // 1 if leaf
  std::vector<int> lstart, lcount;
  std::vector<double> quad;     // 6 per node: xx, yy, zz, xy, xz, yz
  std::vector<int> perm;        // particle order (leaves contiguous)
};

struct Builder {
  const double* pos;
  const double* m;
  int leaf_size;
  int want_quad;
  BHTree& t;
  std::vector<int> tmp;

  int build(int start, int count, double cx, double cy, double cz, double half, int depth) {
    const int node = (int)t.com.size();
    double mass = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = start; k < start + count; ++k) {
      const int p = t.perm[k];
      mass += m[p];
    }
    for (int k = start; k < start + count; ++k) {
      const int p = t.perm[k];
      sx += m[p] * pos[3 * (size_t)p];
      sy += m[p] * pos[3 * (size_t)p + 1];
      sz += m[p] * pos[3 * (size_t)p + 2];
    }
    double4 c;
    if (mass > 0) c = make_double4(sx / mass, sy / mass, sz / mass, mass);
    else c = make_double4(cx, cy, cz, mass);
    t.com.push_back(c);
    t.cube.push_back(make_double4(cx, cy, cz, half));
    t.next.push_back(-1);
    t.leaf.push_back(1);
    t.lstart.push_back(start);
    t.lcount.push_back(count);
    if (want_quad) {
      double q[6] = {0, 0, 0, 0, 0, 0};
      for (int k = start; k < start + count; ++k) {
        const int p = t.perm[k];
        const double x = pos[3 * (size_t)p] - c.x, y = pos[3 * (size_t)p + 1] - c.y, z = pos[3 * (size_t)p + 2] - c.z;
        const double r2 = x * x + y * y + z * z, mp = m[p];
        q[0] += mp * (3 * x * x - r2);
        q[1] += mp * (3 * y * y - r2);
        q[2] += mp * (3 * z * z - r2);
        q[3] += mp * 3 * x * y;
        q[4] += mp * 3 * x * z;
        q[5] += mp * 3 * y * z;
      }
      for (int k = 0; k < 6; ++k) t.quad.push_back(q[k]);
    }
    if (count > leaf_size && depth < 48) {
      t.leaf[node] = 0;
      // stable partition of perm[start, start+count) by octant 0..7
      int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      std::vector<unsigned char> oct(count);
      for (int k = 0; k < count; ++k) {
        const int p = t.perm[start + k];
        const int o = (pos[3 * (size_t)p] > cx ? 1 : 0) | (pos[3 * (size_t)p + 1] > cy ? 2 : 0) |
                      (pos[3 * (size_t)p + 2] > cz ? 4 : 0);
        oct[k] = (unsigned char)o;
        ++cnt[o];
      }
      int off[8];
      int acc = 0;
      for (int o = 0; o < 8; ++o) {
        off[o] = acc;
        acc += cnt[o];
      }
      for (int k = 0; k < count; ++k) tmp[start + off[oct[k]]++] = t.perm[start + k];
      for (int k = 0; k < count; ++k) t.perm[start + k] = tmp[start + k];
      int s = start;
      for (int o = 0; o < 8; ++o) {
        if (cnt[o] == 0) continue;
        // tree.py:41-42: child center = center + sign * half / 2, child half = half / 2
        const double sxn = (o & 1) ? 1.0 : -1.0, syn = (o & 2) ? 1.0 : -1.0, szn = (o & 4) ? 1.0 : -1.0;
        build(s, cnt[o], cx + sxn * half / 2, cy + syn * half / 2, cz + szn * half / 2, half / 2, depth + 1);
        s += cnt[o];
      }
    }
    t.next[node] = (int)t.com.size();  // first node after the subtree
    return node;
  }
};

struct BHDev {
  double4* com = nullptr;
  double4* cube = nullptr;
  int* next = nullptr;
  int* leaf = nullptr;
  int* lstart = nullptr;
  int* lcount = nullptr;
  double* quad = nullptr;
  int* perm = nullptr;
  double4* ppos = nullptr;  // particles in tree order: x, y, z, m
  size_t cap[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

static BHDev g_bh;

template <bool QUAD>
__global__ void __launch_bounds__(256) k_bh_walk(int n, int nnodes, const double4* __restrict__ com,
                                                 const double4* __restrict__ cube, const int* __restrict__ next,
                                                 const int* __restrict__ leaf, const int* __restrict__ lstart,
                                                 const int* __restrict__ lcount, const double* __restrict__ quad,
                                                 const int* __restrict__ perm, const double4* __restrict__ ppos,
                                                 double G, double eps2, double th2, double* __restrict__ F,
                                                 double* partials, unsigned* counter, double* Uout) {
  __shared__ double red[16];
  const int k = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  double U[1] = {0.0};
  if (k < n) {
    const double4 p = ppos[k];
    double ax = 0, ay = 0, az = 0, phi = 0;
    int node = 0;
    while (node >= 0 && node < nnodes) {
      if (leaf[node]) {
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
        node = next[node];
        continue;
      }
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
        node = next[node];
      } else {
        node = node + 1;  // first child
      }
    }
    const int i = perm[k];
    F[3 * (size_t)i] = G * p.w * ax;
    F[3 * (size_t)i + 1] = G * p.w * ay;
    F[3 * (size_t)i + 2] = G * p.w * az;
    U[0] = 0.5 * G * p.w * phi;
  }
  if (Uout) grid_sum<1>(U, partials, counter, Uout, red);
}

template <typename T>
static int upload(T** d, size_t* cap, const std::vector<T>& v, hipStream_t st) {
  if (v.size() > *cap || !*d) {
    if (*d) HIP_TRY(hipFree(*d));
    size_t c = v.size() + v.size() / 4 + 16;
    HIP_TRY(hipMalloc(d, c * sizeof(T)));
    *cap = c;
  }
  HIP_TRY(hipMemcpyAsync(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return ACCU_OK;
}

}  // namespace accu

using namespace accu;

extern "C" {

// pos_host/m_host: host arrays (the tree is built on the host); F and U: device.
int accu_bh_forces(const double* pos_host, const double* m_host, int n, double G, double eps2, double theta,
                   int leaf_size, int quadrupole, double* F, double* U, long long* nnodes_out, void* stream) {
  ARG_CHECK(pos_host && m_host && F, "null argument");
  ARG_CHECK(leaf_size >= 1, "leaf_size must be >= 1");
  hipStream_t st = (hipStream_t)stream;
  if (n < 2) return ACCU_OK;
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) lo[d] = hi[d] = pos_host[d];
  for (int i = 1; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      const double v = pos_host[3 * (size_t)i + d];
      lo[d] = v < lo[d] ? v : lo[d];
      hi[d] = v > hi[d] ? v : hi[d];
    }
  double ext = 0.0;
  for (int d = 0; d < 3; ++d) ext = (hi[d] - lo[d]) > ext ? (hi[d] - lo[d]) : ext;
  double half = ext / 2 * (1 + 1e-9);
  half = half > 1e-300 ? half : 1e-300;
  BHTree t;
  t.perm.resize(n);
  for (int i = 0; i < n; ++i) t.perm[i] = i;
  t.com.reserve(n / 2 + 16);
  Builder b{pos_host, m_host, leaf_size, quadrupole, t, std::vector<int>(n)};
  b.build(0, n, (lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2, half, 0);
  const int nn = (int)t.com.size();
  for (int k = 0; k < nn; ++k)
    if (t.next[k] >= nn) t.next[k] = -1;
  std::vector<double4> pp(n);
  for (int k = 0; k < n; ++k) {
    const int p = t.perm[k];
    pp[k] = make_double4(pos_host[3 * (size_t)p], pos_host[3 * (size_t)p + 1], pos_host[3 * (size_t)p + 2], m_host[p]);
  }
  BHDev& d = g_bh;
  int rc;
  if ((rc = upload(&d.com, &d.cap[0], t.com, st))) return rc;
  if ((rc = upload(&d.cube, &d.cap[1], t.cube, st))) return rc;
  if ((rc = upload(&d.next, &d.cap[2], t.next, st))) return rc;
  if ((rc = upload(&d.leaf, &d.cap[3], t.leaf, st))) return rc;
  if ((rc = upload(&d.lstart, &d.cap[4], t.lstart, st))) return rc;
  if ((rc = upload(&d.lcount, &d.cap[5], t.lcount, st))) return rc;
  if (quadrupole && (rc = upload(&d.quad, &d.cap[6], t.quad, st))) return rc;
  if ((rc = upload(&d.perm, &d.cap[7], t.perm, st))) return rc;
  if ((rc = upload(&d.ppos, &d.cap[8], pp, st))) return rc;
  Workspace& ws = default_workspace();
  const int nb = (n + 255) / 256;
  if ((rc = ws.ensure((size_t)nb * 2))) return rc;
  if (quadrupole)
    hipLaunchKernelGGL(k_bh_walk<true>, dim3(nb), dim3(256), 0, st, n, nn, d.com, d.cube, d.next, d.leaf, d.lstart,
                       d.lcount, d.quad, d.perm, d.ppos, G, eps2, theta * theta, F, ws.partials, ws.counter, U);
  else
    hipLaunchKernelGGL(k_bh_walk<false>, dim3(nb), dim3(256), 0, st, n, nn, d.com, d.cube, d.next, d.leaf, d.lstart,
                       d.lcount, d.quad, d.perm, d.ppos, G, eps2, theta * theta, F, ws.partials, ws.counter, U);
  LAUNCH_CHECK();
  if (nnodes_out) *nnodes_out = nn;
  // the host vectors above are released on return: finish the copies first
  HIP_TRY(hipStreamSynchronize(st));
  return ACCU_OK;
}

}  // extern "C"
