// Primaries-only run that records where sub-threshold recoils come to rest,
// per recoil Z, and counts damage with the NRT estimate; histograms written to
// <base>_ranges.dat (apps/include/TrimRange.h, apps/src/TrimRange.C:31-121).
#ifndef MYTRIM_AMD_TRIMRANGE_H
#define MYTRIM_AMD_TRIMRANGE_H

#include <vector>

#include "ThreadedTrimBase.h"

namespace MyTRIM_NS {

class TrimRange : public ThreadedTrimBase {
 public:
  TrimRange(SimconfType* simconf, SampleBase* sample) : ThreadedTrimBase(simconf, sample), _range(112) {}
  virtual void threadJoin(const ThreadedTrimBase& other);
  virtual void writeOutput();

  std::vector<std::vector<Real>> _range;  // x of resting recoils, per Z

 protected:
  virtual void vacancyCreation();
  virtual void dissipateRecoilEnergy();
  virtual bool followRecoil() { return _recoil->_gen < 1; }
};

}  // namespace MyTRIM_NS

#endif
