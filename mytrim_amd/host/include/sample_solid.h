#ifndef MYTRIM_AMD_SAMPLE_SOLID_H
#define MYTRIM_AMD_SAMPLE_SOLID_H

#include "sample.h"

namespace MyTRIM_NS {

// homogeneous sample: material[0] everywhere (sample_solid.h)
struct SampleSolid : public SampleBase {
 public:
  SampleSolid(Real x, Real y, Real z) : SampleBase(x, y, z) {}
  virtual MaterialBase* lookupMaterial(Point&) { return material[0]; }
  virtual int geometry() const { return 0; }
};

}  // namespace MyTRIM_NS

#endif
