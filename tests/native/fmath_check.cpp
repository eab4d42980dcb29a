// Accuracy of mt_fmath.h (host instantiation of the same code the kernel
// runs) against long-double references: prints the worst error in ulp for
// random arguments over given ranges.
//   fmath_check N seed
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "mt_fmath.h"

static double ulp_err(double got, long double want) {
  const double w = (double)want;
  if (w == 0.0) return got == 0.0 ? 0.0 : 1e300;
  const double u = std::nextafter(std::fabs(w), INFINITY) - std::fabs(w);
  return (double)(std::fabs((long double)got - want) / u);
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
  std::mt19937_64 g(argc > 2 ? std::atoll(argv[2]) : 1);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  double e_exp = 0, e_log = 0, e_pow = 0, e_pow_rel = 0, e_powp = 0;
  double e_rcp = 0, e_rsqrt = 0, e_sqrt = 0, e_sin = 0, e_cos = 0;
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (long i = 0; i < n; ++i) {
    const double xe = -700.0 + 1400.0 * u(g);
    e_exp = std::fmax(e_exp, ulp_err(mt::fx_exp(xe), std::exp((long double)xe)));
    e_exp = std::fmax(e_exp, ulp_err(mt::fx_exp_fin(xe), std::exp((long double)xe)));
    const double xl = std::pow(10.0, -300.0 + 600.0 * u(g));
    e_log = std::fmax(e_log, ulp_err(mt::fx_log(xl), std::log((long double)xl)));
    e_log = std::fmax(e_log, ulp_err(mt::fx_log_pos(xl), std::log((long double)xl)));
    const double xl2 = 0.5 + 1.5 * u(g);  // around 1, where log is small
    e_log = std::fmax(e_log, ulp_err(mt::fx_log(xl2), std::log((long double)xl2)));
    e_log = std::fmax(e_log, ulp_err(mt::fx_log_pos(xl2), std::log((long double)xl2)));
    // pow over the physics' ranges: x in [1e-8, 1e6], y in [-2, 2]
    const double xp = std::pow(10.0, -8.0 + 14.0 * u(g)), yp = -2.0 + 4.0 * u(g);
    const long double wp = std::pow((long double)xp, (long double)yp);
    const double ep = std::fmax(ulp_err(mt::fx_pow(xp, yp), wp), ulp_err(mt::fx_pow_pos(xp, yp), wp));
    e_pow = std::fmax(e_pow, ep);
    const double t = std::fabs(yp * std::log(xp));
    e_pow_rel = std::fmax(e_pow_rel, ep / (t + 1.0));
    e_powp = std::fmax(e_powp, (double)(std::fabs((long double)mt::fx_pow(xp, yp) - wp) / std::fabs(wp)));
    // rcp / rsqrt / sqrt over [1e-30, 1e30] (the host seed is single precision)
    const double xr = std::pow(10.0, -30.0 + 60.0 * u(g)) * (u(g) < 0.5 ? -1.0 : 1.0);
    e_rcp = std::fmax(e_rcp, ulp_err(mt::fx_rcp(xr), 1.0L / (long double)xr));
    const double xs = std::fabs(xr);
    e_rsqrt = std::fmax(e_rsqrt, ulp_err(mt::fx_rsqrt(xs), 1.0L / std::sqrt((long double)xs)));
    e_sqrt = std::fmax(e_sqrt, ulp_err(mt::fx_sqrt(xs), std::sqrt((long double)xs)));
    // sin / cos of 2 pi u on the (0, 1] grid of 32-bit draws; error in ulp of 1
    // (absolute), since both pass through zero inside the interval
    const double uu = ((double)(uint32_t)g() + 1.0) * (1.0 / 4294967296.0);
    double sn, cs;
    mt::fx_sincos2pi(uu, sn, cs);
    const long double ang = two_pi * (long double)uu;
    e_sin = std::fmax(e_sin, (double)(std::fabs((long double)sn - std::sin(ang)) / 0x1p-53L));
    e_cos = std::fmax(e_cos, (double)(std::fabs((long double)cs - std::cos(ang)) / 0x1p-53L));
  }
  const bool zero_ok = mt::fx_sqrt(0.0) == 0.0;
  // special arguments: log(0) = -inf, log(inf) = inf, log(<0) and log(NaN)
  // NaN, exp(-inf) = 0, exp(inf) = inf, exp(NaN) NaN, pow(x, 0) = 1,
  // exp(-740) subnormal, exp(-800) = 0
  const double inf = HUGE_VAL, nan = std::nan("");
  const bool special_ok = mt::fx_log(0.0) == -inf && mt::fx_log(inf) == inf && std::isnan(mt::fx_log(-1.0)) &&
                          std::isnan(mt::fx_log(nan)) && mt::fx_exp(-inf) == 0.0 && mt::fx_exp(inf) == inf &&
                          std::isnan(mt::fx_exp(nan)) && mt::fx_pow(0.0, 0.0) == 1.0 && mt::fx_pow(2.0, 0.0) == 1.0 &&
                          mt::fx_exp(-740.0) > 0.0 && mt::fx_exp(-740.0) < 1e-300 && mt::fx_exp(-800.0) == 0.0;
  std::printf("special_ok %d\n", (int)special_ok);
  std::printf("rcp_ulp %.3f\nrsqrt_ulp %.3f\nsqrt_ulp %.3f\nsin_abs_ulp1 %.3f\ncos_abs_ulp1 %.3f\nsqrt_zero_ok %d\n",
              e_rcp, e_rsqrt, e_sqrt, e_sin, e_cos, (int)zero_ok);
  std::printf("exp_ulp %.3f\nlog_ulp %.3f\npow_ulp %.3f\npow_ulp_per_unit_ylogx %.3f\npow_rel %.3e\n", e_exp, e_log,
              e_pow, e_pow_rel, e_powp);
  return 0;
}
