// Closed-form symmetric eigenvalues and the Frangi vesselness formula, as
// device functions shared by the fused Hessian epilogue and the standalone
// kernels.
//
// Reference: feature/corner.py:473-696 (eigen kernels), filters/ridges.py:
// 228-319 (Frangi inner kernels), paths relative to
// python/cucim/src/cucim/skimage/.  Transcendentals are HIP's device math
// library (OCML) functions.
#pragma once

#include "hc_common.h"

namespace hc {

template <typename F>
__device__ __forceinline__ void swapv(F& x, F& y) {
  F t = y;
  y = x;
  x = t;
}

// 2x2: the temporaries tmp1/tmp2 are of the element type F (corner.py:475-490),
// so for float32 input the arithmetic after each double product is float32.
template <typename F>
__device__ __forceinline__ void eig22(F M00, F M01, F M11, int descending, int abs_sort, F& l1,
                                      F& l2) {
  const double m00 = (double)M00, m01 = (double)M01, m11 = (double)M11;
  F tmp1 = (F)(m01 * m01);
  tmp1 *= (F)4;
  F tmp2 = (F)(m00 - m11);
  tmp2 *= tmp2;
  tmp2 += tmp1;
  tmp2 = sqrt(tmp2);
  tmp2 /= (F)2;
  tmp1 = (F)(m00 + m11);
  tmp1 /= (F)2;
  if (!descending) {
    l1 = tmp1 - tmp2;
    l2 = tmp1 + tmp2;
    if (abs_sort && fabs(l1) > fabs(l2)) swapv(l1, l2);
  } else {
    l1 = tmp1 + tmp2;
    l2 = tmp1 - tmp2;
    if (abs_sort && fabs(l1) < fabs(l2)) swapv(l1, l2);
  }
}

// cos(t) and sin(t) for t in [0, pi/3], in double: the Taylor series to the
// t^18 and t^19 terms (truncation below 2e-18 on that range), by Horner's
// rule in t^2.  The closed form below needs the cosines of (phi + k pi) / 3
// with phi in [0, pi]; with t = phi / 3 they are cos t and
// cos t / 2 -+ (sqrt(3) / 2) sin t, so one pair on [0, pi/3] serves all
// three and the general argument reduction of the library functions is not
// needed.
__device__ __forceinline__ void sincos_small(double t, double& c, double& s) {
  const double z = t * t;
  double pc = 1.0 / 6402373705728000.0;  // 1/18!
  pc = fma(pc, z, -1.0 / 20922789888000.0);  // -1/16!
  pc = fma(pc, z, 1.0 / 87178291200.0);      // 1/14!
  pc = fma(pc, z, -1.0 / 479001600.0);       // -1/12!
  pc = fma(pc, z, 1.0 / 3628800.0);          // 1/10!
  pc = fma(pc, z, -1.0 / 40320.0);           // -1/8!
  pc = fma(pc, z, 1.0 / 720.0);              // 1/6!
  pc = fma(pc, z, -1.0 / 24.0);              // -1/4!
  pc = fma(pc, z, 0.5);                      // 1/2!
  c = fma(-pc, z, 1.0);
  double ps = -1.0 / 121645100408832000.0;   // -1/19!
  ps = fma(ps, z, 1.0 / 355687428096000.0);  // 1/17!
  ps = fma(ps, z, -1.0 / 1307674368000.0);   // -1/15!
  ps = fma(ps, z, 1.0 / 6227020800.0);       // 1/13!
  ps = fma(ps, z, -1.0 / 39916800.0);        // -1/11!
  ps = fma(ps, z, 1.0 / 362880.0);           // 1/9!
  ps = fma(ps, z, -1.0 / 5040.0);            // -1/7!
  ps = fma(ps, z, 1.0 / 120.0);              // 1/5!
  ps = fma(ps, z, -1.0 / 6.0);               // -1/3!
  s = fma(ps * z, t, t);
}

// 3x3, Deledalle et al. (2017) closed form in double (corner.py:566-612) for
// the matrix [[a, d, f], [d, b, e], [f, e, c]], then the reference's three
// compare-swap steps (corner.py:613-647).  With abs_sort the comparison keys
// are the absolute values taken before any swap and are not updated by the
// swaps; this ordering is reproduced as is.
template <typename F>
__device__ __forceinline__ void eig33(F aa, F bb, F cc, F dd, F ee, F ff, int descending,
                                      int abs_sort, F& l1, F& l2, F& l3) {
  const double a = aa, b = bb, c = cc, d = dd, e = ee, f = ff;
  const double d_sq = d * d, e_sq = e * e, f_sq = f * f;
  const double tmpa = (2 * a - b - c);
  const double tmpb = (2 * b - a - c);
  const double tmpc = (2 * c - a - b);
  double x2 = -tmpa * tmpb * tmpc;
  x2 += 9 * (tmpc * d_sq + tmpb * f_sq + tmpa * e_sq);
  x2 -= 54 * (d * e * f);
  double x1 = a * a + b * b + c * c - a * b - a * c - b * c + 3 * (d_sq + e_sq + f_sq);
  x1 = fmax(x1, 0.0);
  double phi;
  if (x2 == 0.0) {
    phi = M_PI / 2.0;
  } else {
    const double arg = fmax(4 * x1 * x1 * x1 - x2 * x2, 0.0);
    phi = atan(sqrt(arg) / x2);
    if (x2 < 0) phi += M_PI;
  }
  // the reference divides by 3; a product with the double nearest 1/3
  // differs by at most an ulp of double, far below the float rounding that
  // follows, and avoids four IEEE divisions
  constexpr double third = 1.0 / 3.0;
  const double x1_term = (2.0 / 3.0) * sqrt(x1);
  const double abc = (a + b + c) * third;
  // cos(phi / 3), cos((phi - pi) / 3), cos((phi + pi) / 3)
  double ct, st;
  sincos_small(phi * third, ct, st);
  const double half_c = 0.5 * ct, r3_s = 0.86602540378443864676 * st;  // sqrt(3) / 2
  F lam1 = (F)(abc - x1_term * ct);
  F lam2 = (F)(abc + x1_term * (half_c + r3_s));
  F lam3 = (F)(abc + x1_term * (half_c - r3_s));
  // the network runs on (var1, lam2, var3) = (lam1, lam2, lam3) ascending,
  // (lam3, lam2, lam1) descending
  F& v1 = descending ? lam3 : lam1;
  F& v3 = descending ? lam1 : lam3;
  if (abs_sort) {
    const F k1 = fabs(v1), k2 = fabs(lam2), k3 = fabs(v3);
    if (k1 > k2) swapv(v1, lam2);
    if (k1 > k3) swapv(v1, v3);
    if (k2 > k3) swapv(lam2, v3);
  } else {
    if (v1 > lam2) swapv(v1, lam2);
    if (v1 > v3) swapv(v1, v3);
    if (lam2 > v3) swapv(lam2, v3);
  }
  l1 = lam1;
  l2 = lam2;
  l3 = lam3;
}

// r_g = sum lambda^2 in F (ridges.py:228-240, 270-283)
template <typename F>
__device__ __forceinline__ F frangi_rg2(F l1, F l2) {
  F r = l1 * l1;
  return fma_(l2, l2, r);
}
template <typename F>
__device__ __forceinline__ F frangi_rg3(F l1, F l2, F l3) {
  F r = l1 * l1;
  r = fma_(l2, l2, r);
  return fma_(l3, l3, r);
}

// ridges.py:243-267: eigenvalues ascending by |value|
template <typename F>
__device__ __forceinline__ F frangi_v2(F l1, F l2, F r_g, double beta_sq, double gamma_sq) {
  F r_b = fabs(l1) / fmax(l2, (F)1.0e-10);
  r_b *= r_b;
  F result = (F)exp(-(double)r_b / beta_sq);
  result = (F)((double)result * (1.0 - exp(-(double)r_g / gamma_sq)));
  return result;
}

// ridges.py:286-319
template <typename F>
__device__ __forceinline__ F frangi_v3(F l1, F l2, F l3, F r_g, double alpha_sq, double beta_sq,
                                       double gamma_sq) {
  const F lam2 = fmax(l2, (F)1.0e-10);
  const F lam3 = fmax(l3, (F)1.0e-10);
  F r_a = lam2 / lam3;
  r_a *= r_a;
  const F r_b = (l1 * l1) / (lam2 * lam3);
  F result = (F)(1.0 - exp(-(double)r_a / alpha_sq));
  result = (F)((double)result * exp(-(double)r_b / beta_sq));
  result = (F)((double)result * (1.0 - exp(-(double)r_g / gamma_sq)));
  return result;
}

// The fused passes' forms of frangi_v2 / frangi_v3: the same double
// exponentials, with each argument -r / s_sq computed as r * (-1 / s_sq) and
// the reciprocals taken once per kernel.  A product with the rounded
// reciprocal differs from the quotient by at most an ulp of double, far below
// the rounding to F that follows, and saves three double divisions per voxel.
struct FrangiNegRecip {
  double a, b, g;  // -1 / alpha_sq, -1 / beta_sq, -1 / gamma_sq
};
// a wave-uniform double kept in scalar registers (the divisions run on the
// vector ALU, which would otherwise hold the results in vector registers)
__device__ __forceinline__ double uniform_d(double x) {
  const unsigned long long u = __double_as_longlong(x);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// alpha_sq, beta_sq, gamma_sq must be the same in every lane of a wave
__device__ __forceinline__ FrangiNegRecip frangi_neg_recip(double alpha_sq, double beta_sq,
                                                           double gamma_sq) {
  return {uniform_d(-1.0 / alpha_sq), uniform_d(-1.0 / beta_sq), uniform_d(-1.0 / gamma_sq)};
}

template <typename F>
__device__ __forceinline__ F frangi_v2r(F l1, F l2, F r_g, const FrangiNegRecip& q) {
  F r_b = fabs(l1) / fmax(l2, (F)1.0e-10);
  r_b *= r_b;
  F result = (F)exp((double)r_b * q.b);
  result = (F)((double)result * (1.0 - exp((double)r_g * q.g)));
  return result;
}

template <typename F>
__device__ __forceinline__ F frangi_v3r(F l1, F l2, F l3, F r_g, const FrangiNegRecip& q) {
  const F lam2 = fmax(l2, (F)1.0e-10);
  const F lam3 = fmax(l3, (F)1.0e-10);
  F r_a = lam2 / lam3;
  r_a *= r_a;
  const F r_b = (l1 * l1) / (lam2 * lam3);
  F result = (F)(1.0 - exp((double)r_a * q.a));
  result = (F)((double)result * exp((double)r_b * q.b));
  result = (F)((double)result * (1.0 - exp((double)r_g * q.g)));
  return result;
}

// float32 forms of the two above, used by the fused float32 passes: the
// exponentials in float (1 - e^x as -expm1(x), accurate where x is small)
// instead of double.  The result then differs from the reference's
// double-then-round value by a few float ulps (each exponential within about
// an ulp of its argument's relative error, one float rounding per product),
// far below the rounding of the float32 Hessian the vesselness is built
// from; it saves the three double exponentials per voxel and sigma.
template <>
__device__ __forceinline__ float frangi_v2r<float>(float l1, float l2, float r_g,
                                                   const FrangiNegRecip& q) {
  float r_b = fabsf(l1) / fmaxf(l2, 1.0e-10f);
  r_b *= r_b;
  const float result = expf(r_b * (float)q.b);
  return result * -expm1f(r_g * (float)q.g);
}

template <>
__device__ __forceinline__ float frangi_v3r<float>(float l1, float l2, float l3, float r_g,
                                                   const FrangiNegRecip& q) {
  const float lam2 = fmaxf(l2, 1.0e-10f);
  const float lam3 = fmaxf(l3, 1.0e-10f);
  float r_a = lam2 / lam3;
  r_a *= r_a;
  const float r_b = (l1 * l1) / (lam2 * lam3);
  float result = -expm1f(r_a * (float)q.a);
  result *= expf(r_b * (float)q.b);
  return result * -expm1f(r_g * (float)q.g);
}

// cp.maximum: NaN-propagating
template <typename F>
__device__ __forceinline__ F nanmax(F a, F b) {
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}

// gamma from the first sigma (ridges.py:447-453), in double as the host
// Python computes it: gamma = sqrt(s)/2; gamma_sq = 2.0 if s == 0 else
// max(2*gamma*gamma, 1e-10) (Python's max keeps a NaN first argument).
__device__ __forceinline__ double gamma_sq_from_max(double s_max) {
  const double g = sqrt(s_max) / 2.0;
  if (s_max == 0.0) return 2.0;
  const double g2 = 2 * g * g;
  return (1e-10 > g2) ? 1e-10 : g2;
}

}  // namespace hc
