#ifndef MYTRIM_AMD_SAMPLE_DYNAMIC_H
#define MYTRIM_AMD_SAMPLE_DYNAMIC_H

#include "sample_layers.h"
#include "simconf.h"

namespace MyTRIM_NS {

// layers whose composition and thickness change as atoms are added to or
// removed from them (sample_dynamic.h, sample_dynamic.C).  On the device it is
// a SampleLayers: trim() and TrimGPU::runCascades upload the current layers
// and compositions on their next call, so composition updates run as host
// steps between flights or batches.
class SampleDynamic : public SampleLayers {
 public:
  SampleDynamic(SimconfType* simconf, Real x, Real y, Real z);

  virtual void averages(const IonBase* pka);
  virtual MaterialBase* lookupMaterial(Point& pos);

  // add n atoms (n < 0: remove) of element Z to `layer`, keeping its density:
  // the thickness grows by their mass over rho w1 w2, the stoichiometry by
  // n over the layer's atom count (sample_dynamic.C:67-119)
  virtual void addAtomsToLayer(int layer, int n, int Z);

 protected:
  const IonBase* pka;
  SimconfType* _simconf;
};

}  // namespace MyTRIM_NS

#endif
