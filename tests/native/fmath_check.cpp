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
  for (long i = 0; i < n; ++i) {
    const double xe = -700.0 + 1400.0 * u(g);
    e_exp = std::fmax(e_exp, ulp_err(mt::fx_exp(xe), std::exp((long double)xe)));
    const double xl = std::pow(10.0, -300.0 + 600.0 * u(g));
    e_log = std::fmax(e_log, ulp_err(mt::fx_log(xl), std::log((long double)xl)));
    const double xl2 = 0.5 + 1.5 * u(g);  // around 1, where log is small
    e_log = std::fmax(e_log, ulp_err(mt::fx_log(xl2), std::log((long double)xl2)));
    // pow over the physics' ranges: x in [1e-8, 1e6], y in [-2, 2]
    const double xp = std::pow(10.0, -8.0 + 14.0 * u(g)), yp = -2.0 + 4.0 * u(g);
    const long double wp = std::pow((long double)xp, (long double)yp);
    const double ep = ulp_err(mt::fx_pow(xp, yp), wp);
    e_pow = std::fmax(e_pow, ep);
    const double t = std::fabs(yp * std::log(xp));
    e_pow_rel = std::fmax(e_pow_rel, ep / (t + 1.0));
    e_powp = std::fmax(e_powp, (double)(std::fabs((long double)mt::fx_pow(xp, yp) - wp) / std::fabs(wp)));
  }
  std::printf("exp_ulp %.3f\nlog_ulp %.3f\npow_ulp %.3f\npow_ulp_per_unit_ylogx %.3f\npow_rel %.3e\n", e_exp, e_log,
              e_pow, e_pow_rel, e_powp);
  return 0;
}
