// wf_device.h -- device helpers shared by the step kernels (wf_kernels.hip)
// and their adjoint (wf_adjoint.hip): launch shapes, block reductions,
// the cubic-spline gradient factor, yield laws, amplitudes.
#pragma once
#include <math.h>

#include "wf_internal.h"

namespace wf {

#define WF_KERNEL __global__ __launch_bounds__(kBlock)

static inline int nblk(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// grid of the neighbour-build kernels: they run every step but exit at once
// unless a search is due, so they use a modest grid with grid-stride loops
static inline int nblk_small(long long n) {
  const int b = nblk(n);
  return b < 1024 ? b : 1024;
}

// XCD-aware tile index.  The dispatcher hands consecutive workgroups to the
// 8 XCDs in turn and each XCD has its own L2; remapping makes XCD x walk one
// contiguous run of tiles, so a particle and its (Morton-adjacent)
// neighbours are fetched into the same L2.  Bijective for any grid size.
constexpr int kXcd = 8;
__device__ inline int xcd_tile(int b, int nb) {
  const int q = nb / kXcd, r = nb % kXcd;
  const int x = b % kXcd, k = b / kXcd;
  return x * q + (x < r ? x : r) + k;
}

// block partial of reduction slot `slot` (slot-major: coalesced in finalize)
__device__ inline double &part(const Dev &d, int slot, int block) {
  return d.partials[(long long)slot * d.nblocks + block];
}

// ---------------------------------------------------------------------------
// reductions

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

// block reduction in a fixed order; result valid in thread 0
template <int OP>
__device__ inline double block_reduce(double v, double *sh) {
  v = OP == 0 ? wave_sum(v) : (OP == 1 ? wave_max(v) : wave_min(v));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    r = sh[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
      r = OP == 0 ? r + sh[k] : (OP == 1 ? fmax(r, sh[k]) : fmin(r, sh[k]));
  }
  __syncthreads();
  return r;
}

// lanes > 1: lane `sub` > 0 of each particle hands its NV partial sums to
// lane 0 through LDS (red: (LANES - 1) * NV * kBlock values); lane 0 adds
// them in lane order.  Every thread of the block must call this.
template <int LANES, int NV, typename R>
__device__ inline void combine_lanes(R (&v)[NV], R *red) {
  if constexpr (LANES > 1) {
    const int lane = threadIdx.x % kBlock, sub = threadIdx.x / kBlock;
    if (sub > 0)
      for (int c = 0; c < NV; ++c) red[((sub - 1) * NV + c) * kBlock + lane] = v[c];
    __syncthreads();
    if (sub == 0)
      for (int q = 1; q < LANES; ++q)
        for (int c = 0; c < NV; ++c) v[c] += red[((q - 1) * NV + c) * kBlock + lane];
  }
}

// ---------------------------------------------------------------------------
// kernel functions (Source/Functions.cpp:74-87) and the neighbour predicate

// cubic spline gradient factor GK with GK * x_ij = grad_i W_ij (3D),
// Source/Functions.cpp:74-87:  GK = C/(q h) * P(q), C = 1/(pi h^4), so for
// q < 1: GK = c5 (-3 + 9/4 q) and for 1 <= q < 2: GK = c5 (-3/4 (2-q)^2) / q
// with c5 = 1/(pi h^5).  Reciprocals use the hardware estimate refined by two
// Newton steps (a few ulp from the correctly rounded quotient).
__device__ inline double rcp_d(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
__device__ inline double rsqrt_d(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// kernel constants of smoothing length h in dimension dim: the gradient
// factor's c5 = 1/(pi h^5) in 3D, 10/(7 pi h^4) in 2D (plane strain; the
// cubic spline's C/h with C from Source/Functions.cpp:81)
struct KernH {
  double h, ih, c5;  // pair smoothing length, 1/h, c5
  int dim;
};
__device__ inline KernH kern_h(double h, int dim = 3) {
  KernH k;
  k.h = h;
  k.ih = rcp_d(h);
  k.dim = dim;
  const double ih2 = k.ih * k.ih;
  k.c5 = dim == 2 ? ih2 * ih2 * (10.0 / (7.0 * M_PI)) : ih2 * ih2 * k.ih * (1.0 / M_PI);
  return k;
}
// h of the pair (h_i + h_j)/2; reuses particle i's constants when h_j == h_i
__device__ inline KernH pair_h(const KernH &ki, double hi, double hj) {
  if (hj == hi) return ki;
  return kern_h((hi + hj) / 2, ki.dim);
}
__device__ inline double grad_kernel_r2(double r2, const KernH &k) {
  if (r2 == 0.0) return -3.0 * k.c5;
  const double ri = rsqrt_d(r2);
  const double q = r2 * ri * k.ih;
  if (q < 1.0) return k.c5 * (-3.0 + 2.25 * q);
  if (q < 2.0) {
    const double t = 2.0 - q;
    return k.c5 * (-0.75 * t * t) * (k.h * ri);
  }
  return 0.0;
}

// the same kernel math for either pair precision R (double or float):
// float uses the 1-ulp hardware reciprocal / reciprocal square root
__device__ inline double rcp_t(double x) { return rcp_d(x); }
__device__ inline float rcp_t(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double rsqrt_t(double x) { return rsqrt_d(x); }
__device__ inline float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }

template <typename R>
struct KH {
  R h, ih, c5;
  int dim;
};
template <typename R>
__device__ inline KH<R> kern_h_t(R h, int dim = 3) {
  KH<R> k;
  k.h = h;
  k.ih = rcp_t(h);
  k.dim = dim;
  const R ih2 = k.ih * k.ih;
  k.c5 = dim == 2 ? ih2 * ih2 * (R)(10.0 / (7.0 * M_PI)) : ih2 * ih2 * k.ih * (R)(1.0 / M_PI);
  return k;
}
template <typename R>
__device__ inline KH<R> pair_h_t(const KH<R> &ki, R hi, R hj) {
  if (hj == hi) return ki;
  return kern_h_t<R>((hi + hj) / (R)2, ki.dim);
}
// both branches evaluated and selected (no divergent branch in the pair
// loops, so loads of later neighbours can be scheduled early); each branch
// rounds exactly as a branchy version would.  r2 = 0 makes ri infinite and q
// NaN, which selects nothing until the last select.
template <typename R>
__device__ inline R gk_t(R r2, const KH<R> &k) {
  const R ri = rsqrt_t(r2);
  const R q = r2 * ri * k.ih;
  const R inner = k.c5 * ((R)-3 + (R)2.25 * q);
  const R t = (R)2 - q;
  const R outer = k.c5 * ((R)-0.75 * t * t) * (k.h * ri);
  R g = q < (R)1 ? inner : (q < (R)2 ? outer : (R)0);
  return r2 == (R)0 ? (R)-3 * k.c5 : g;
}


// amplitude::getValAtTime (Source/Domain.h:75-87)
__device__ inline double amp_value(const DevAmp &a, double t) {
  int i = a.n - 1;
  bool end = false;
  while (!end) {
    i--;
    if (i == 0 || t > a.t[i]) end = true;
  }
  return a.v[i] + (a.v[i + 1] - a.v[i]) / (a.t[i + 1] - a.t[i]) * (t - a.t[i]);
}

__device__ inline double jc_yield(const DevMat &M, double strain, double sr, double T) {
  const double T_h = (T - M.jc_Tt) / (M.jc_Tm - M.jc_Tt);
  double f = 1.0;
  if (sr > M.jc_eps0) f = 1.0 + M.jc_C * log(sr / M.jc_eps0);
  return (M.jc_A + M.jc_B * pow(strain, M.jc_n)) * f * (1.0 - pow(T_h, M.jc_m));
}

// JohnsonCookDamage::CalcFractureStrain (Source/Material.cpp:76-92): the
// strain at fracture at plastic strain rate sr, triaxiality sig_as and
// temperature T, with the material's T_t, T_m and eps_0; capped at 10
__device__ inline double jc_fracture_strain(const DevMat &M, double sr, double sig_as, double T) {
  const double T_h = (T - M.jc_Tt) / (M.jc_Tm - M.jc_Tt);
  double f_sr = 1.0;
  double f_triax = M.dam[0];
  if (sr > M.jc_eps0) f_sr = 1.0 + log(sr / M.jc_eps0) * M.dam[3];
  if (sig_as < 1.5) f_triax = M.dam[0] + M.dam[1] * exp(M.dam[2] * sig_as);
  double ef = f_triax * f_sr * (1.0 + M.dam[4] * T_h);
  if (ef > 10.0) ef = 10.0;
  return ef;
}

// the damage weight of a pair: 1 - D_pair with D_pair = (D_i + D_j) / 2,
// clamped to 1 (the Johnson-Cook criterion's pair value, Source/Damage.cpp:
// 141-150; Source/InteractionAlt.cpp:54-60); not positive: the pair is
// skipped
__device__ inline double dam_weight(double Di, double Dj) {
  const double Dp = (Di + Dj) / 2.0;
  double f = 1.0;
  if (Dp > 0.0) f = 1.0 - Dp;
  if (f > 1.0) f = 1.0;
  return f;
}

__device__ inline double hol_yield(const DevMat &M, double strain) {
  if (strain + M.hol_eps0 > M.hol_eps1) return M.hol_K * pow(strain + M.hol_eps0, M.hol_m);
  return M.hol_sy0;
}

__device__ inline void cross_d(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace wf
