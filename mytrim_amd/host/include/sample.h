// SampleBase of the MyTRIM API (sample.h).  The device runs the sample kinds
// SampleSolid, SampleLayers, sampleClusters, SampleWire and
// SampleBurriedWire; geometry() reports which one a sample is.
#ifndef MYTRIM_AMD_SAMPLE_H
#define MYTRIM_AMD_SAMPLE_H

#include <vector>

#include "ion.h"
#include "material.h"

namespace MyTRIM_NS {

class SampleBase {
 public:
  SampleBase(Real x = 10000.0, Real y = 10000.0, Real z = 10000.0);
  virtual ~SampleBase() {}

  virtual void averages(const IonBase* pka);
  virtual MaterialBase* lookupMaterial(Point& pos) = 0;
  virtual Real rangeMaterial(Point& pos, Point& dir);

  std::vector<MaterialBase*> material;
  Real w[3];
  enum sampleBoundary { PBC, INF, CUT };
  sampleBoundary bc[3];

  // device geometry kind: 0 solid, 1 layers, 2 clusters, 3 wire, 4 buried
  // wire, -1 not supported
  virtual int geometry() const { return -1; }
};

}  // namespace MyTRIM_NS

#endif
