// Host code of the autotuner (mt_tune.h, what mt_autotune runs).  Builds the
// phases of synthetic launches whose ramp, steady rate and drain are known and
// prints the projected rates the test compares with the closed form; then
// prints the picks and finalists of synthetic sets of timed shapes.
//   tune_check
#include <cstdio>

#include "mt_tune.h"

using mt::LaunchPhases;
using mt::occupancy_finalists;
using mt::pick_shape;
using mt::projected_rate;
using mt::TimedShape;

static TimedShape shape(int block, int bpc, int cap, double rate, int occ, int lds = 0) {
  TimedShape s{};
  s.t.block = block;
  s.t.blocks_per_cu = bpc;
  s.t.stack_cap = cap;
  s.t.lds_hist = lds;
  s.rate = rate;
  s.occ = occ;
  return s;
}

static void print_shapes(const char* tag, const std::vector<TimedShape>& v) {
  std::printf("%s %zu", tag, v.size());
  for (const auto& s : v) std::printf(" %d:%d:%d:%d", s.t.block, s.t.blocks_per_cu, s.t.stack_cap, s.t.lds_hist);
  std::printf("\n");
}

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

  // picks: the fastest; within 2 % of it the deepest stack; none without a rate
  std::vector<TimedShape> v = {shape(256, 2, 128, 10.0, 8), shape(256, 2, 96, 10.1, 8), shape(256, 2, 192, 9.7, 8)};
  // 96 entries are the fastest, 128 (index 0) is within 2 % and deeper, 192 is 4 % behind
  std::printf("pick_fastest %d\n", pick_shape(v));
  v = {shape(256, 2, 96, 10.0, 8), shape(256, 2, 192, 9.85, 8), shape(256, 2, 128, 9.9, 8)};
  std::printf("pick_deeper %d\n", pick_shape(v));
  v = {shape(256, 2, 96, 0.0, 8)};
  std::printf("pick_none %d\n", pick_shape(v));
  // finalists per occupancy, ascending: 8 waves per CU (pick and runner-up),
  // 12 (pick only: its other shape is more than 10 % behind the fastest), 4
  // (none: more than 10 % behind)
  v = {shape(256, 2, 192, 18.4, 8), shape(256, 2, 128, 18.3, 8), shape(128, 4, 128, 17.9, 8, -1),
       shape(256, 3, 96, 19.8, 12), shape(128, 6, 96, 16.0, 12), shape(256, 1, 384, 15.0, 4)};
  print_shapes("final_runner", occupancy_finalists(v, true));
  print_shapes("final_pick", occupancy_finalists(v, false));
  return 0;
}
