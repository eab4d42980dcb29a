// Vacancy and replacement counts per 1 A of depth, written to <base>_vac.dat
// (apps/include/TrimVacCount.h, apps/src/TrimVacCount.C:31-79).
#ifndef MYTRIM_AMD_TRIMVACCOUNT_H
#define MYTRIM_AMD_TRIMVACCOUNT_H

#include <vector>

#include "ThreadedTrimBase.h"

namespace MyTRIM_NS {

class TrimVacCount : public ThreadedTrimBase {
 public:
  TrimVacCount(SimconfType* simconf, SampleBase* sample) : ThreadedTrimBase(simconf, sample) {}
  virtual void threadJoin(const ThreadedTrimBase& other);
  virtual void writeOutput();

  // depth bin of x the way the reference truncates it (int(x), x < 0 dropped)
  static void count(std::vector<unsigned int>& bins, Real x);
  std::vector<unsigned int> _vac_bin, _repl_bin;

 protected:
  virtual void vacancyCreation();
  virtual void replacementCollision();
};

}  // namespace MyTRIM_NS

#endif
