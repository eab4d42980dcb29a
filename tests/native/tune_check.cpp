// Launch-rate projection of the autotuner (mt_tune.h, the host code
// mt_autotune runs).  Builds the phases of synthetic launches whose ramp,
// steady rate and drain are known, and prints the projected rates the test
// compares with the closed form.
//   tune_check
#include <cstdio>

#include "mt_tune.h"

using mt::LaunchPhases;
using mt::projected_rate;

// a launch of h histories with c collisions each, f flights each: it ramps
// for r seconds (doing wr collisions), runs at s collisions per second, and
// drains for d seconds (doing wd collisions)
static LaunchPhases synth(long long h, double c, double f, double r, double wr, double s, double d, double wd) {
  LaunchPhases p;
  p.n_hist = h;
  p.n_coll = (long long)(h * c);
  p.n_flight = (long long)(h * f);
  const double cpf = c / f;
  const double steady_coll = double(p.n_coll) - wr - wd;
  p.half_s = r;
  p.admit_s = r + steady_coll / s;
  p.kernel_s = p.admit_s + d;
  p.flights_half = (long long)(wr / cpf);
  p.flights_admit = p.n_flight - (long long)(wd / cpf);
  return p;
}

int main() {
  // C2-like sample: 2e5 histories of 59 222 collisions in 6 569 flights
  const LaunchPhases a = synth(200000, 59222.0, 6569.0, 0.30, 5.0e9, 2.0e10, 0.09, 6.0e8);
  std::printf("same %.6e %.6e\n", projected_rate(a, a.n_hist), double(a.n_coll) / a.kernel_s);
  std::printf("smaller %.6e\n", projected_rate(a, a.n_hist / 2));
  const double n = 59222.0 * 1e6;
  std::printf("batch %.6e %.6e\n", projected_rate(a, 1000000), n / (0.30 + (n - 5.0e9 - 6.0e8) / 2.0e10 + 0.09));
  std::printf("huge %.6e\n", projected_rate(a, 1000000000000LL));
  // missing stamps, stamps out of order, a steady phase under a quarter of the flights
  LaunchPhases b = a;
  b.half_s = 0.0;
  std::printf("nostamp %.6e\n", projected_rate(b, 1000000));
  b = a;
  b.admit_s = b.half_s;
  std::printf("order %.6e\n", projected_rate(b, 1000000));
  b = a;
  b.flights_half = b.flights_admit - b.n_flight / 5;
  std::printf("short %.6e\n", projected_rate(b, 1000000));
  return 0;
}
