// Torn-entry detection of the queue checksum (mt_qsum.h, host instantiation of
// the code the kernel runs).  A reader that loads a ten-field entry while its
// writer's stores are still arriving sees some fields of the slot's previous
// lap; the checksum over what it loaded must differ from the one the writer
// published.  Trials draw the previous and the new entry three ways (random
// fields, fields one bit apart, fields a small integer apart: nearby doubles
// and counters), take any non-empty strict subset of fields from the previous
// entry, and also check the whole previous entry under the new lap.  Prints
// trials and undetected mixtures; a 32-bit check should miss about one in 4e9.
//   qsum_check N seed
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "mt_qsum.h"

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
  std::mt19937_64 g(argc > 2 ? std::atoll(argv[2]) : 1);
  long long trials = 0, missed = 0, same_ok = 0;
  for (long t = 0; t < n; ++t) {
    uint64_t prev[10], next[10], seen[10];
    const uint32_t lap = 2u + (uint32_t)(g() % 100000u);
    const int mode = (int)(t % 3);
    for (int i = 0; i < 10; ++i) {
      prev[i] = g();
      next[i] = mode == 0 ? g() : mode == 1 ? prev[i] ^ (1ull << (g() % 64)) : prev[i] + 1 + g() % 15;
    }
    const uint32_t stale = 1u + (uint32_t)(g() % 1023u);  // fields still of the previous lap
    for (int i = 0; i < 10; ++i) seen[i] = (stale >> i & 1u) ? prev[i] : next[i];
    mt::QSum hn = mt::sum_init(lap), hs = mt::sum_init(lap), hp = mt::sum_init(lap), hr = mt::sum_init(lap);
    for (int i = 0; i < 10; ++i) {
      hn = mt::sum_add(hn, next[i]);
      hs = mt::sum_add(hs, seen[i]);
      hp = mt::sum_add(hp, prev[i]);
      hr = mt::sum_add(hr, next[i]);
    }
    same_ok += mt::sum_end(hr) == mt::sum_end(hn);
    trials += 2;
    missed += mt::sum_end(hs) == mt::sum_end(hn);
    missed += mt::sum_end(hp) == mt::sum_end(hn);
  }
  std::printf("trials %lld\nmissed %lld\nsame_ok %lld\n", trials, missed, same_ok);
  return 0;
}
