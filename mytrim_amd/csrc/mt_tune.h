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
#include <vector>

#include "mytrim_gpu.h"

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

// A launch shape timed by the autotuner: its tuning, its (projected) rate and
// the waves per CU its launch ran at.
struct TimedShape {
  mt_tuning t;
  double rate;
  int occ;
};

inline bool same_shape(const mt_tuning& a, const mt_tuning& b) {
  return a.block == b.block && a.blocks_per_cu == b.blocks_per_cu && a.stack_cap == b.stack_cap;
}

// Index of the pick among v (-1 if none has a rate): the fastest, except that
// among shapes within 2 % of it (the spread of one timing) the deepest LDS
// stack wins, the fastest of those at that depth.  On full launches shallower
// stacks than a sample's pick measured slower (C3: 96 entries picked over
// 128, -8 %).
inline int pick_shape(const std::vector<TimedShape>& v) {
  double top = 0.0;
  for (const auto& e : v) top = std::max(top, e.rate);
  int best = -1;
  for (int i = 0; i < (int)v.size(); ++i) {
    const TimedShape& e = v[size_t(i)];
    if (!(e.rate > 0.0) || e.rate < 0.98 * top) continue;
    if (best < 0 || e.t.stack_cap > v[size_t(best)].t.stack_cap ||
        (e.t.stack_cap == v[size_t(best)].t.stack_cap && e.rate > v[size_t(best)].rate))
      best = i;
  }
  return best;
}

// Finalists of one group of timed shapes: for each occupancy (ascending), its
// pick if that is within `within` of the group's fastest; with runner_up,
// also the fastest shape of another block size, blocks per CU or stack depth
// at that occupancy within the same margin.
inline std::vector<TimedShape> occupancy_finalists(const std::vector<TimedShape>& v, bool runner_up,
                                                   double within = 0.9) {
  std::vector<TimedShape> out;
  double top = 0.0;
  for (const auto& e : v) top = std::max(top, e.rate);
  if (!(top > 0.0)) return out;
  std::vector<int> occs;
  for (const auto& e : v) occs.push_back(e.occ);
  std::sort(occs.begin(), occs.end());
  occs.erase(std::unique(occs.begin(), occs.end()), occs.end());
  for (int oc : occs) {
    std::vector<TimedShape> g;
    for (const auto& e : v)
      if (e.occ == oc) g.push_back(e);
    const int p = pick_shape(g);
    if (p < 0 || g[size_t(p)].rate < within * top) continue;
    out.push_back(g[size_t(p)]);
    if (!runner_up) continue;
    int second = -1;
    for (int i = 0; i < (int)g.size(); ++i)
      if (!same_shape(g[size_t(i)].t, g[size_t(p)].t) && g[size_t(i)].rate >= within * top &&
          (second < 0 || g[size_t(i)].rate > g[size_t(second)].rate))
        second = i;
    if (second >= 0) out.push_back(g[size_t(second)]);
  }
  return out;
}

}  // namespace mt
