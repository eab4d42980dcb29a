#ifndef MYTRIM_AMD_SAMPLE_LAYERS_H
#define MYTRIM_AMD_SAMPLE_LAYERS_H

#include <vector>

#include "sample.h"

namespace MyTRIM_NS {

// slabs stacked along x (sample_layers.h)
class SampleLayers : public SampleBase {
 public:
  SampleLayers(Real x, Real y, Real z) : SampleBase(x, y, z) {}
  virtual MaterialBase* lookupMaterial(Point& pos);
  virtual Real rangeMaterial(Point& pos, Point& dir);
  virtual int lookupLayer(Point& pos);
  virtual int geometry() const { return 1; }
  std::vector<Real> layerThickness;
};

}  // namespace MyTRIM_NS

#endif
