#ifndef MYTRIM_AMD_SAMPLE_WIRE_H
#define MYTRIM_AMD_SAMPLE_WIRE_H

#include "sample.h"

namespace MyTRIM_NS {

// nanowire along z: the cylinder inscribed in the w[0] x w[1] cross-section
// holds material[0], the rest is vacuum; CUT boundaries in x and y
// (sample_wire.h, sample_wire.C:30-46)
class SampleWire : public SampleBase {
 public:
  SampleWire(Real x = 10000.0, Real y = 10000.0, Real z = 10000.0);
  virtual MaterialBase* lookupMaterial(Point& pos);
  virtual int geometry() const { return 3; }
};

}  // namespace MyTRIM_NS

#endif
