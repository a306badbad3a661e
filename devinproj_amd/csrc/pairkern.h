// Pair interactions (LJ, DSF, Ewald real space) in float64 or float32 pair math, and
// the generic list / dense loops that apply them. Accumulation is always float64.
#pragma once
#include "common.h"

namespace accu {

constexpr int kMaxSpecies = 8;

// ---------------------------------------------------------------- Lennard-Jones
// forces.py:119-142 (list) and 157-175 (dense):
//   mask  r2 < (cutoff*sig)^2 and eps > 0
//   sr6 = (sig^2 / r2)^3, u = 4 eps (sr6^2 - sr6 - (src6^2 - src6))
//   coef = -24 eps (2 sr6^2 - sr6) / r2, f_i += coef * d (d = r_j - r_i)
struct LJParams {
  int ns;
  double shift;  // src6^2 - src6 with src6 = (1/cutoff^2)^3
  double eps[kMaxSpecies * kMaxSpecies];
  double sig2[kMaxSpecies * kMaxSpecies];
  double rc2[kMaxSpecies * kMaxSpecies];
};

template <typename P>
struct LJ {
  const int* __restrict__ sp;
  // tables staged in LDS
  P* eps;
  P* sig2;
  P* rc2;
  P shift;

  __device__ void stage(const LJParams& p, double* lds) {
    P* base = reinterpret_cast<P*>(lds);
    eps = base;
    sig2 = base + kMaxSpecies * kMaxSpecies;
    rc2 = base + 2 * kMaxSpecies * kMaxSpecies;
    const int m = p.ns * p.ns;
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
      eps[k] = (P)p.eps[k];
      sig2[k] = (P)p.sig2[k];
      rc2[k] = (P)p.rc2[k];
    }
    shift = (P)p.shift;
  }
  __device__ __forceinline__ int row(int i, int ns) const { return sp[i] * ns; }
};

// One LJ pair in float64; returns true when inside.
__device__ __forceinline__ void lj_pair64(double dx, double dy, double dz, double r2, double eps, double sig2,
                                          double rc2, double shift, double& fx, double& fy, double& fz,
                                          double& u) {
  if (r2 < rc2 && eps > 0.0) {
    const double ir2 = rcp64(r2);
    const double x = sig2 * ir2;
    const double sr6 = x * x * x;
    u += 2.0 * eps * (sr6 * sr6 - sr6 - shift);  // half of 4 eps (...): full list counts each pair twice
    const double coef = -24.0 * eps * (2.0 * sr6 * sr6 - sr6) * ir2;
    fx = fma(coef, dx, fx);
    fy = fma(coef, dy, fy);
    fz = fma(coef, dz, fz);
  }
}

// Same pair with float32 arithmetic; the caller accumulates in float64.
__device__ __forceinline__ void lj_pair32(float dx, float dy, float dz, float r2, float eps, float sig2, float rc2,
                                          float shift, float& fx, float& fy, float& fz, float& u) {
  if (r2 < rc2 && eps > 0.0f) {
    const float ir2 = __builtin_amdgcn_rcpf(r2);
    const float x = sig2 * ir2;
    const float sr6 = x * x * x;
    u += 2.0f * eps * (sr6 * sr6 - sr6 - shift);
    const float coef = -24.0f * eps * (2.0f * sr6 * sr6 - sr6) * ir2;
    fx = fmaf(coef, dx, fx);
    fy = fmaf(coef, dy, fy);
    fz = fmaf(coef, dz, fz);
  }
}

// ---------------------------------------------------------------- DSF / Ewald real
// forces.py:408-438: pair kept when sqrt(r2) < rc;
//   u = qq (erfc(a r)/r - erc/rc + fshift (r - rc))
//   f_i -= qq ((erfc(a r)/r^2 + c exp(-(a r)^2)/r) - fshift) / r * d
// forces.py:371-379 (Ewald real space, shift-free): u = qq erfc(a r)/r,
//   f_i -= qq (erfc(a r)/r + c exp(-(a r)^2)) / r^2 * d
struct CoulParams {
  double alpha, rc, c;  // c = 2 alpha / sqrt(pi)
  double erc_rc;        // erfc(alpha rc) / rc
  double fshift;        // DSF force shift (0 for Ewald)
  double self;          // per-ion self energy factor: erc/(2rc) + alpha/sqrt(pi) (DSF) or alpha/sqrt(pi) (Ewald)
  double k;             // Coulomb constant applied to F and U
  int shifted;          // 1 = DSF, 0 = Ewald real space
  int nex;              // number of excluded id pairs (sorted keys lo*n+hi)
  const long long* ex;
};

__device__ __forceinline__ bool excluded(const CoulParams& p, int i, int j, int n) {
  if (p.nex == 0) return false;
  const long long lo = i < j ? i : j, hi = i < j ? j : i;
  const long long key = lo * (long long)n + hi;
  int a = 0, b = p.nex - 1;
  while (a <= b) {
    int m = (a + b) >> 1;
    long long v = p.ex[m];
    if (v == key) return true;
    if (v < key) a = m + 1;
    else b = m - 1;
  }
  return false;
}

__device__ __forceinline__ void coul_pair64(const CoulParams& p, double dx, double dy, double dz, double r2,
                                            double qq, double& fx, double& fy, double& fz, double& u) {
  const double r = sqrt(r2);
  if (r < p.rc) {
    const double ar = p.alpha * r;
    const double er = erfc(ar);
    const double ir = rcp64(r);
    const double ex = exp(-(ar * ar));
    double fm;
    if (p.shifted) {
      u += 0.5 * qq * (er * ir - p.erc_rc + p.fshift * (r - p.rc));
      fm = qq * ((er * ir * ir + p.c * ex * ir) - p.fshift) * ir;
    } else {
      u += 0.5 * qq * er * ir;
      fm = qq * (er * ir + p.c * ex) * ir * ir;
    }
    fx = fma(-fm, dx, fx);
    fy = fma(-fm, dy, fy);
    fz = fma(-fm, dz, fz);
  }
}

// erfc(x) for 0 <= x < 4, given ex = exp(-x^2): ex R(t) with t = 1 / (1 + p x), R a degree-8
// polynomial in t without constant term. The form is that of Abramowitz and Stegun (1964),
// eq. 7.1.26; p and the coefficients are fitted here to erfcx(x) = exp(x^2) erfc(x) on
// [0, 4] for the least largest relative error (7.5e-9 in exact arithmetic; 8.3e-7 against
// erfc in float32 with the exponential's rounding included).
__device__ __forceinline__ float erfc_from_exp(float x, float ex) {
  const float t = __builtin_amdgcn_rcpf(fmaf(0.406f, x, 1.0f));
  float r = 0.0799551249f;
  r = fmaf(r, t, -0.403218155f);
  r = fmaf(r, t, 0.666917063f);
  r = fmaf(r, t, -0.356064685f);
  r = fmaf(r, t, 0.400140398f);
  r = fmaf(r, t, 0.144174441f);
  r = fmaf(r, t, 0.239786361f);
  r = fmaf(r, t, 0.228309452f);
  return ex * (r * t);
}

// The float32 DSF pair (mixed precision): r and 1/r from one reciprocal square root, and
// erfc from the exponential the force needs anyway (the library erfcf beyond x = 4).
__device__ __forceinline__ void coul_pair32(float alpha, float rc, float c, float erc_rc, float fshift, int shifted,
                                            float dx, float dy, float dz, float r2, float qq, float& fx, float& fy,
                                            float& fz, float& u) {
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float r = r2 * ir;
  if (r < rc) {
    const float ar = alpha * r;
    const float ex = __expf(-(ar * ar));
    const float er = ar < 4.0f ? erfc_from_exp(ar, ex) : erfcf(ar);
    float fm;
    if (shifted) {
      u += 0.5f * qq * (er * ir - erc_rc + fshift * (r - rc));
      fm = qq * ((er * ir * ir + c * ex * ir) - fshift) * ir;
    } else {
      u += 0.5f * qq * er * ir;
      fm = qq * (er * ir + c * ex) * ir * ir;
    }
    fx = fmaf(-fm, dx, fx);
    fy = fmaf(-fm, dy, fy);
    fz = fmaf(-fm, dz, fz);
  }
}

}  // namespace accu
