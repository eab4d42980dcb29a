// TrimBase with a per-thread join step, the base of the runmytrim output
// modules (apps/include/ThreadedTrimBase.h of the reference).
//
// The reference leaves _primaries_only uninitialised unless the input sets
// output.primaries_only; here it defaults to false (every recoil followed).
#ifndef MYTRIM_AMD_THREADEDTRIMBASE_H
#define MYTRIM_AMD_THREADEDTRIMBASE_H

#include "trim.h"

namespace MyTRIM_NS {

class ThreadedTrimBase : public TrimBase {
 public:
  ThreadedTrimBase(SimconfType* simconf, SampleBase* sample) : TrimBase(simconf, sample) {}
  virtual void threadJoin(const ThreadedTrimBase&) = 0;
  bool _primaries_only = false;

 protected:
  virtual bool followRecoil() { return !_primaries_only; }
};

}  // namespace MyTRIM_NS

#endif
