// Launch phases of a cascade launch and the rate they project for a launch of
// more histories (host code of the autotuner, mt_autotune).
//
// A launch ramps up until the admission of its middle primary, runs at a
// steady rate until the admission of its last one, then drains: the last
// cascades' longest flight chains run on an emptying device.  The kernel
// stamps both admissions (wall clock, flights ended by then).  A launch of N
// collisions is projected to take the sample's ramp and drain (their time
// and the collisions done in them) plus the rest of its collisions at the
// steady rate.
#pragma once
#include <algorithm>

namespace mt {

struct LaunchPhases {
  double kernel_s = 0.0;             // kernel time
  double half_s = 0.0;               // start to the middle admission (0: not stamped)
  double admit_s = 0.0;              // start to the last admission (0: not stamped)
  long long n_coll = 0, n_flight = 0;
  long long flights_half = 0;        // flights ended by the middle admission
  long long flights_admit = 0;       // flights ended by the last admission
  long long n_hist = 0;              // histories of the launch
};

// Collisions per second projected for a launch of n_batch histories at the
// sample's collisions per history.  The sample's own rate (collisions over
// kernel time) where n_batch is not larger than the sample, where the stamps
// are missing or out of order, or where the steady phase holds under a
// quarter of the flights (all admissions in the first round: fewer histories
// than the cascades the device admits at once).
inline double projected_rate(const LaunchPhases& p, long long n_batch) {
  const double whole = p.kernel_s > 0.0 ? double(p.n_coll) / p.kernel_s : 0.0;
  if (n_batch <= p.n_hist || p.n_hist <= 0 || p.n_flight <= 0 || p.kernel_s <= 0.0) return whole;
  if (!(p.half_s > 0.0 && p.admit_s > p.half_s && p.flights_admit > p.flights_half)) return whole;
  if (4 * (p.flights_admit - p.flights_half) < p.n_flight) return whole;
  const double cpf = double(p.n_coll) / double(p.n_flight);
  const double steady = double(p.flights_admit - p.flights_half) * cpf / (p.admit_s - p.half_s);
  const double w_ramp = double(p.flights_half) * cpf;
  const double w_drain = double(p.n_flight - p.flights_admit) * cpf;
  const double drain = std::max(0.0, p.kernel_s - p.admit_s);
  const double n = double(p.n_coll) / double(p.n_hist) * double(n_batch);
  return n / (p.half_s + std::max(0.0, n - w_ramp - w_drain) / steady + drain);
}

}  // namespace mt
