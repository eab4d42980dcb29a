#ifndef MYTRIM_AMD_SAMPLE_BURRIED_WIRE_H
#define MYTRIM_AMD_SAMPLE_BURRIED_WIRE_H

#include "sample_wire.h"

namespace MyTRIM_NS {

// a SampleWire under a 250 A cover of material[1] (-250 <= z < 0), with
// material[1] around the wire up to z = w[2] and vacuum beyond; INF
// boundaries (sample_burried_wire.h, sample_burried_wire.C:29-55)
class SampleBurriedWire : public SampleWire {
 public:
  SampleBurriedWire(Real x = 10000.0, Real y = 10000.0, Real z = 10000.0);
  virtual MaterialBase* lookupMaterial(Point& pos);
  virtual int geometry() const { return 4; }
};

}  // namespace MyTRIM_NS

#endif
