// Structural and transport observables of accusim/analysis.py on the device:
//   rdf pair counts (analysis.py:7-24): every pair i < j, minimum image, |d| binned
//       exactly as numpy.histogram bins it (the edges come from the host's linspace);
//   msd (analysis.py:27-43): displacements accumulated frame to frame (minimum-image
//       steps in a periodic box), mean |disp|^2 per frame;
//   vacf (analysis.py:77-87): <v(t0) . v(t0 + lag)> over atoms and time origins.
// Dot products and squared norms of 3-vectors keep numpy's orders (einsum: (x x' + z z')
// + y y'; sum over an axis of 3: (x^2 + y^2) + z^2).
// The pair counts are integers (exact); the float64 means are summed per block in a
// fixed order and then over blocks in block order, so results do not depend on timing.
#include "common.h"
#include "../../include/accusim_hip.h"

namespace accu {

constexpr int kRdfT = 256;        // threads per block (one i-atom each), j-tile size
constexpr int kRdfMaxBins = 2048;  // histogram bins held in LDS

// Bin of r among `bins` uniform bins of [0, rmax] with edges e[0..bins], as
// numpy.histogram (numpy/lib/_histograms_impl.py) computes it: index from
// (r - 0) / rmax * bins, the last edge folded into the last bin, then corrected by
// comparisons with the edges. r must lie in [0, rmax].
__device__ __forceinline__ int np_bin(double r, double rmax, int bins, const double* e) {
#pragma clang fp contract(off)
  int i = (int)((r - 0.0) / rmax * (double)bins);
  if (i == bins) i -= 1;
  if (r < e[i]) i -= 1;
  if (r >= e[i + 1] && i != bins - 1) i += 1;
  return i;
}

// Pair counts of the tile pair (I, J) = (blockIdx.y, blockIdx.x), atoms [I*T, I*T+T) x
// [J*T, J*T+T); blocks below the diagonal (J < I) leave at once.
__global__ void __launch_bounds__(kRdfT) k_rdf(const double* __restrict__ pos, int n, double L0,
                                               double L1, double L2, int periodic, double rmax, int bins,
                                               const double* __restrict__ edges,
                                               unsigned long long* __restrict__ counts) {
  __shared__ double xj[kRdfT][3];
  __shared__ unsigned hist[kRdfMaxBins];
  __shared__ double e[kRdfMaxBins + 1];
  const int I = blockIdx.y, J = blockIdx.x;
  if (J < I) return;
  for (int k = threadIdx.x; k < bins; k += blockDim.x) hist[k] = 0u;
  for (int k = threadIdx.x; k <= bins; k += blockDim.x) e[k] = edges[k];
  const int j0 = J * kRdfT;
  {
    const int j = j0 + threadIdx.x;
    if (j < n) {
      xj[threadIdx.x][0] = pos[3 * (size_t)j];
      xj[threadIdx.x][1] = pos[3 * (size_t)j + 1];
      xj[threadIdx.x][2] = pos[3 * (size_t)j + 2];
    }
  }
  __syncthreads();
  const int i = I * kRdfT + threadIdx.x;
  if (i < n) {
    const double xi0 = pos[3 * (size_t)i], xi1 = pos[3 * (size_t)i + 1], xi2 = pos[3 * (size_t)i + 2];
    const int jend = min(kRdfT, n - j0);
    // i < j: in the diagonal tile only the j past i
    for (int t = (I == J ? threadIdx.x + 1 : 0); t < jend; ++t) {
#pragma clang fp contract(off)
      double d0 = xj[t][0] - xi0, d1 = xj[t][1] - xi1, d2 = xj[t][2] - xi2;
      if (periodic) {
        d0 = minimg_exact(d0, L0);
        d1 = minimg_exact(d1, L1);
        d2 = minimg_exact(d2, L2);
      }
      const double r = sqrt(r2_np(d0, d1, d2));
      if (r <= rmax) atomicAdd(&hist[np_bin(r, rmax, bins, e)], 1u);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < bins; k += blockDim.x)
    if (hist[k]) atomicAdd(&counts[k], (unsigned long long)hist[k]);
}

// Unwrapped mean square displacement, one thread per atom walking the frames: frame f
// of atom a is at frames[f * stride + 3 a]. Block partial sums of |disp|^2 per frame go
// to part[block][f] (fixed-order block reduction).
__global__ void k_msd(const double* __restrict__ frames, int nf, long long stride, int n, double L0, double L1,
                      double L2, int periodic, double* __restrict__ part) {
  __shared__ double red[kRdfT / 64];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double dx = 0, dy = 0, dz = 0;
  double px = 0, py = 0, pz = 0;
  if (a < n) {
    px = frames[3 * (size_t)a];
    py = frames[3 * (size_t)a + 1];
    pz = frames[3 * (size_t)a + 2];
  }
  for (int f = 1; f < nf; ++f) {
    double v = 0.0;
    if (a < n) {
#pragma clang fp contract(off)
      const double* q = frames + (size_t)f * stride + 3 * (size_t)a;
      const double bx = q[0], by = q[1], bz = q[2];
      double sx = bx - px, sy = by - py, sz = bz - pz;
      if (periodic) {
        sx = minimg_exact(sx, L0);
        sy = minimg_exact(sy, L1);
        sz = minimg_exact(sz, L2);
      }
      dx = dx + sx;
      dy = dy + sy;
      dz = dz + sz;
      px = bx;
      py = by;
      pz = bz;
      v = (dx * dx + dy * dy) + dz * dz;
    }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
      part[(size_t)blockIdx.x * nf + f] = t;
    }
    __syncthreads();
  }
}

// Velocity autocorrelation partial sums: part[block][k] = sum over the block's atoms and
// the time origins t < nf - k of v(t) . v(t + k).
__global__ void k_vacf(const double* __restrict__ frames, int nf, long long stride, int n, int max_lag,
                       double* __restrict__ part) {
  __shared__ double red[kRdfT / 64];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = 0; k < max_lag; ++k) {
    double c = 0.0;
    if (a < n)
      for (int t = 0; t + k < nf; ++t) {
        const double* u = frames + (size_t)t * stride + 3 * (size_t)a;
        const double* w = frames + (size_t)(t + k) * stride + 3 * (size_t)a;
        c += (u[0] * w[0] + u[2] * w[2]) + u[1] * w[1];  // einsum's order for a 3-vector
      }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
      part[(size_t)blockIdx.x * max_lag + k] = t;
    }
    __syncthreads();
  }
}

// out[j] = sum over blocks (in block order) of part[b][j], j < m.
__global__ void k_sum_blocks(const double* __restrict__ part, int nb, int m, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  double t = 0.0;
  for (int b = 0; b < nb; ++b) t += part[(size_t)b * m + j];
  out[j] = t;
}

}  // namespace accu

using namespace accu;

extern "C" {

int accu_rdf_counts(const double* pos, int n, const double* box_host, int periodic, double rmax,
                    const double* edges_host, int bins, unsigned long long* counts_host, void* stream) {
  ARG_CHECK(pos && edges_host && counts_host, "null argument");
  ARG_CHECK(n >= 0 && bins >= 1 && bins <= kRdfMaxBins, "bins %d outside [1, %d]", bins, kRdfMaxBins);
  ARG_CHECK(rmax > 0, "rmax must be positive");
  ARG_CHECK(!periodic || box_host, "a periodic rdf needs the box");
  hipStream_t st = (hipStream_t)stream;
  double* e = nullptr;
  unsigned long long* c = nullptr;
  HIP_TRY(hipMallocAsync((void**)&e, (bins + 1) * sizeof(double), st));
  HIP_TRY(hipMallocAsync((void**)&c, bins * sizeof(unsigned long long), st));
  HIP_TRY(hipMemcpyAsync(e, edges_host, (bins + 1) * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemsetAsync(c, 0, bins * sizeof(unsigned long long), st));
  const int nt = (n + kRdfT - 1) / kRdfT;
  ARG_CHECK(nt <= 65535, "too many atoms for the all-pairs rdf (%d)", n);
  const double L[3] = {periodic ? box_host[0] : 1.0, periodic ? box_host[1] : 1.0, periodic ? box_host[2] : 1.0};
  if (nt > 0)
    hipLaunchKernelGGL(k_rdf, dim3(nt, nt), dim3(kRdfT), 0, st, pos, n, L[0], L[1], L[2], periodic, rmax, bins, e,
                       c);
  LAUNCH_CHECK();
  HIP_TRY(hipMemcpyAsync(counts_host, c, bins * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipFreeAsync(e, st));
  HIP_TRY(hipFreeAsync(c, st));
  HIP_TRY(hipStreamSynchronize(st));
  return ACCU_OK;
}

int accu_msd(const double* frames, int nframes, long long stride, int n, const double* box_host, int periodic,
             double* out_host, void* stream) {
  ARG_CHECK(frames && out_host, "null argument");
  ARG_CHECK(nframes >= 1 && n >= 1 && stride >= 3ll * n, "bad frame layout");
  ARG_CHECK(!periodic || box_host, "periodic unwrapping needs the box");
  hipStream_t st = (hipStream_t)stream;
  const int nb = (n + kRdfT - 1) / kRdfT;
  double *part = nullptr, *sum = nullptr;
  HIP_TRY(hipMallocAsync((void**)&part, (size_t)nb * nframes * sizeof(double), st));
  HIP_TRY(hipMallocAsync((void**)&sum, (size_t)nframes * sizeof(double), st));
  HIP_TRY(hipMemsetAsync(part, 0, (size_t)nb * nframes * sizeof(double), st));
  const double L[3] = {periodic ? box_host[0] : 1.0, periodic ? box_host[1] : 1.0, periodic ? box_host[2] : 1.0};
  hipLaunchKernelGGL(k_msd, dim3(nb), dim3(kRdfT), 0, st, frames, nframes, stride, n, L[0], L[1], L[2], periodic,
                     part);
  hipLaunchKernelGGL(k_sum_blocks, dim3((nframes + 255) / 256), dim3(256), 0, st, part, nb, nframes, sum);
  LAUNCH_CHECK();
  HIP_TRY(hipMemcpyAsync(out_host, sum, (size_t)nframes * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipFreeAsync(part, st));
  HIP_TRY(hipFreeAsync(sum, st));
  HIP_TRY(hipStreamSynchronize(st));
  for (int f = 0; f < nframes; ++f) out_host[f] /= (double)n;
  out_host[0] = 0.0;
  return ACCU_OK;
}

int accu_vacf(const double* frames, int nframes, long long stride, int n, int max_lag, double* out_host,
              void* stream) {
  ARG_CHECK(frames && out_host, "null argument");
  ARG_CHECK(nframes >= 1 && n >= 1 && stride >= 3ll * n, "bad frame layout");
  ARG_CHECK(max_lag >= 1 && max_lag <= nframes, "max_lag %d outside [1, %d]", max_lag, nframes);
  hipStream_t st = (hipStream_t)stream;
  const int nb = (n + kRdfT - 1) / kRdfT;
  double *part = nullptr, *sum = nullptr;
  HIP_TRY(hipMallocAsync((void**)&part, (size_t)nb * max_lag * sizeof(double), st));
  HIP_TRY(hipMallocAsync((void**)&sum, (size_t)max_lag * sizeof(double), st));
  hipLaunchKernelGGL(k_vacf, dim3(nb), dim3(kRdfT), 0, st, frames, nframes, stride, n, max_lag, part);
  hipLaunchKernelGGL(k_sum_blocks, dim3((max_lag + 255) / 256), dim3(256), 0, st, part, nb, max_lag, sum);
  LAUNCH_CHECK();
  HIP_TRY(hipMemcpyAsync(out_host, sum, (size_t)max_lag * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipFreeAsync(part, st));
  HIP_TRY(hipFreeAsync(sum, st));
  HIP_TRY(hipStreamSynchronize(st));
  for (int k = 0; k < max_lag; ++k) out_host[k] /= (double)(nframes - k) * (double)n;
  return ACCU_OK;
}

}  // extern "C"
