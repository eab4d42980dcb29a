// FP64 elementary functions for the cascade kernel, without tables or
// branches:
//   exp, log, pow   Taylor series with Cody-Waite argument reduction; the
//                   device libm's routines cost about 40, 95 and 220 VALU
//                   instructions, these about 25, 35 and 60
//   rcp, rsqrt,     the hardware reciprocal / reciprocal-square-root
//   sqrt            estimates (about 2^-23 relative) refined by Newton steps
//                   (5, 9 and 11 instructions against 8, 29 and 18)
//   sincos2pi       sin and cos of 2 pi u for u in [0, 1] by quadrant and
//                   Taylor series (about 35 against sincospi's 67)
// Errors against exact values stay within 1 ulp (exp, rcp, rsqrt, sqrt),
// 2 ulp (log, sincos2pi) and 1.7 ulp per unit of |y log x| (pow);
// tests/test_fmath.py measures the maxima against long double.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace mt {

// ln 2 split so that k * kLn2Hi is exact for |k| < 2^20 (the high part keeps
// 32 significant bits)
constexpr double kLn2 = 0.69314718055994530942;
constexpr double kLn2Hi = 0.69314718036912381649017333984375;    // ln 2 rounded to 32 bits
constexpr double kLn2Lo = 1.9082149292705878161442656807550e-10;  // ln 2 - kLn2Hi, to 53 bits
constexpr double kInvLn2 = 1.4426950408889634074;
constexpr double kSqrtHalf = 0.70710678118654752440;

__host__ __device__ __forceinline__ double fx_frexp(double x, int& e) {
#if defined(__HIP_DEVICE_COMPILE__)
  e = __builtin_amdgcn_frexp_exp(x);
  return __builtin_amdgcn_frexp_mant(x);
#else
  return std::frexp(x, &e);
#endif
}

__host__ __device__ __forceinline__ double fx_ldexp(double x, int e) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_ldexp(x, e);
#else
  return std::ldexp(x, e);
#endif
}

// Hardware estimates of 1/x and 1/sqrt(x).  The host build (used to test the
// refinement) starts from single precision, an error of the same order.
__host__ __device__ __forceinline__ double fx_rcp_seed(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcp(x);
#else
  return (double)(1.0f / (float)x);
#endif
}

__host__ __device__ __forceinline__ double fx_rsq_seed(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsq(x);
#else
  return (double)(1.0f / std::sqrt((float)x));
#endif
}

// 1/x for finite nonzero x of magnitude in [1e-300, 1e300]: two Newton steps
// y += y (1 - x y), each squaring the relative error
__host__ __device__ __forceinline__ double fx_rcp(double x) {
  double y = fx_rcp_seed(x);
  y = fma(y, fma(-x, y, 1.0), y);
  return fma(y, fma(-x, y, 1.0), y);
}

// 1/sqrt(x) for finite x > 0: two Newton steps y += (y / 2)(1 - x y^2)
__host__ __device__ __forceinline__ double fx_rsqrt(double x) {
  double y = fx_rsq_seed(x);
  y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
  return fma(0.5 * y, fma(-x * y, y, 1.0), y);
}

// sqrt(x) for finite x >= 0: s = x y from one Newton step of 1/sqrt(x),
// then the correction s += (y / 2)(x - s^2); sqrt(0) = 0
__host__ __device__ __forceinline__ double fx_sqrt(double x) {
  double y = fx_rsq_seed(x);
  y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
  const double s = x * y;
  const double r = fma(0.5 * y, fma(-s, s, x), s);
  return x == 0.0 ? x : r;
}

// sin(2 pi u) and cos(2 pi u) for u in [0, 1].  n = rint(4u) picks the
// quadrant, f = u - n/4 in [-1/8, 1/8] is exact, x = 2 pi f (two-part 2 pi)
// lies in [-pi/4, pi/4], where the Taylor series to x^17 (sin) and x^16 (cos)
// leave less than 1e-17 relative
__host__ __device__ __forceinline__ void fx_sincos2pi(double u, double& s, double& c) {
  constexpr double k2PiHi = 6.283185307179586232;      // 2 pi rounded to double
  constexpr double k2PiLo = 2.4492935982947063545e-16;  // 2 pi - k2PiHi
  const double n = rint(4.0 * u);
  const double f = fma(n, -0.25, u);
  const double x = fma(f, k2PiHi, f * k2PiLo);
  const double z = x * x;
  double ps = -1.0 / 355687428096000.0;  // -1/17!
  ps = fma(ps, z, 1.0 / 1307674368000.0);
  ps = fma(ps, z, -1.0 / 6227020800.0);
  ps = fma(ps, z, 1.0 / 39916800.0);
  ps = fma(ps, z, -1.0 / 362880.0);
  ps = fma(ps, z, 1.0 / 5040.0);
  ps = fma(ps, z, -1.0 / 120.0);
  ps = fma(ps, z, 1.0 / 6.0);
  const double sx = fma(-x * z, ps, x);
  double pc = 1.0 / 20922789888000.0;  // 1/16!
  pc = fma(pc, z, -1.0 / 87178291200.0);
  pc = fma(pc, z, 1.0 / 479001600.0);
  pc = fma(pc, z, -1.0 / 3628800.0);
  pc = fma(pc, z, 1.0 / 40320.0);
  pc = fma(pc, z, -1.0 / 720.0);
  pc = fma(pc, z, 1.0 / 24.0);
  pc = fma(pc, z, -0.5);
  const double cx = fma(z, pc, 1.0);
  const int q = (int)n;
  const double sv = (q & 1) ? cx : sx;
  const double cv = (q & 1) ? sx : cx;
  s = (q & 2) ? -sv : sv;
  c = ((q + 1) & 2) ? -cv : cv;
}

// exp(r) for |r| <= ln2 / 2: degree-11 polynomial fitted to exp on
// [-ln2/2, ln2/2] by Chebyshev interpolation (maximum error 3.2e-18 before
// the coefficients are rounded to double; the Taylor series needs degree 13
// for the same accuracy)
__host__ __device__ __forceinline__ double fx_exp_core(double r) {
  double p = 2.5110037605963777e-08;
  p = fma(p, r, 2.763263963904103e-07);
  p = fma(p, r, 2.755724091857897e-06);
  p = fma(p, r, 2.4801485482328494e-05);
  p = fma(p, r, 0.00019841269890047113);
  p = fma(p, r, 0.0013888888952314775);
  p = fma(p, r, 0.008333333333319601);
  p = fma(p, r, 0.0416666666664881);
  p = fma(p, r, 0.1666666666666668);
  p = fma(p, r, 0.5000000000000019);
  p = fma(p, r, 1.0);
  return fma(p, r, 1.0);
}

// exp(x) = 2^k exp(r), k = rint(x / ln 2), r = x - k ln 2 (two-part ln 2).
// x is clamped to [-1100, 1100]; the final scaling by 2^k then gives the
// subnormal, zero and infinite results of the full range; NaN propagates.
__host__ __device__ __forceinline__ double fx_exp(double x) {
  const double xc = fmin(fmax(x, -1100.0), 1100.0);
  const double k = rint(xc * kInvLn2);
  const double r = fma(-k, kLn2Lo, fma(-k, kLn2Hi, xc));
  const double v = fx_ldexp(fx_exp_core(r), (int)k);
  return x == x ? v : x;
}

// log(x), x = 2^e m with m in [sqrt(1/2), sqrt(2)): log m = 2 atanh(s),
// s = (m - 1) / (m + 1), |s| <= 0.1716, series in s^2 to s^19 (the next
// term is below 2.3e-17 relative).  Subnormal x is exact through frexp; zero, negative,
// infinite and NaN arguments are selected at the end (-inf, NaN, inf, NaN).
__host__ __device__ __forceinline__ double fx_log(double x) {
  int e;
  double m = fx_frexp(x, e);  // m in [0.5, 1) for finite nonzero x
  if (m < kSqrtHalf) {
    m *= 2.0;
    --e;
  }
  const double f = m - 1.0;
  // s = f / (m + 1) from the refined reciprocal and one residual correction
  const double d = m + 1.0, y = fx_rcp(d);
  const double s0 = f * y;
  const double s = fma(fma(-d, s0, f), y, s0);
  const double z = s * s;
  double p = 1.0 / 19.0;
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  // log m = 2s + 2s z p
  const double two_s = 2.0 * s;
  const double lm = fma(two_s * z, p, two_s);
  const double de = (double)e;
  double v = fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
  v = (x == 0.0) ? -HUGE_VAL : v;
  v = (x == HUGE_VAL) ? HUGE_VAL : v;
  return (x >= 0.0) ? v : __builtin_nan("");  // negative or NaN: NaN
}

// The physics only takes exp of finite arguments and logarithms of finite
// positive ones; these variants leave out the clamping and the special-value
// selects (5 to 8 instructions each).  exp_fin needs |x| < 2^30; NaN still
// gives NaN.
__host__ __device__ __forceinline__ double fx_exp_fin(double x) {
  const double k = rint(x * kInvLn2);
  const double r = fma(-k, kLn2Lo, fma(-k, kLn2Hi, x));
  return fx_ldexp(fx_exp_core(r), (int)k);
}

__host__ __device__ __forceinline__ double fx_log_pos(double x) {
  int e;
  double m = fx_frexp(x, e);
  if (m < kSqrtHalf) {
    m *= 2.0;
    --e;
  }
  const double f = m - 1.0;
  const double d = m + 1.0, y = fx_rcp(d);
  const double s0 = f * y;
  const double s = fma(fma(-d, s0, f), y, s0);
  const double z = s * s;
  double p = 1.0 / 19.0;
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  const double two_s = 2.0 * s;
  const double de = (double)e;
  return fma(de, kLn2Hi, fma(de, kLn2Lo, fma(two_s * z, p, two_s)));
}

// pow for x > 0 finite and y != 0 (the physics' constant exponents)
__host__ __device__ __forceinline__ double fx_pow_pos(double x, double y) { return fx_exp_fin(y * fx_log_pos(x)); }

// pow(x, y) = exp(y log x) for x >= 0 (error grows with |y log x|: about
// 1.7 ulp per unit of |y log x|, 7e-15 relative at |y log x| = 64); x = 0
// and x = inf follow from log's -inf / inf; negative x gives NaN
__host__ __device__ __forceinline__ double fx_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  return fx_exp(y * fx_log(x));
}

}  // namespace mt
