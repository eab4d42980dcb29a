// Vacancies per (int ln E_recoil, 1 A depth bin), written to <base>_evac.dat
// (apps/include/TrimVacEnergyCount.h, apps/src/TrimVacEnergyCount.C:31-81).
#ifndef MYTRIM_AMD_TRIMVACENERGYCOUNT_H
#define MYTRIM_AMD_TRIMVACENERGYCOUNT_H

#include <vector>

#include "ThreadedTrimBase.h"

namespace MyTRIM_NS {

class TrimVacEnergyCount : public ThreadedTrimBase {
 public:
  TrimVacEnergyCount(SimconfType* simconf, SampleBase* sample) : ThreadedTrimBase(simconf, sample) {}
  virtual void threadJoin(const ThreadedTrimBase& other);
  virtual void writeOutput();

  std::vector<std::vector<unsigned int>> _evac_bin;

 protected:
  virtual void vacancyCreation();
};

}  // namespace MyTRIM_NS

#endif
