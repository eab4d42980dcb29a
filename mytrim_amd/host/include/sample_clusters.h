#ifndef MYTRIM_AMD_SAMPLE_CLUSTERS_H
#define MYTRIM_AMD_SAMPLE_CLUSTERS_H

#include <vector>

#include "sample.h"
#include "simconf.h"

namespace MyTRIM_NS {

// spheres of material[1] in material[0], found through a cell hash
// (sample_clusters.h)
struct sampleClusters : SampleBase {
  Real sd, kd[3];
  int* sh;
  int kn[3];
  int *cl, cn, cnm;
  Real* c[4];  // x, y, z, r
  Real cmr;

  sampleClusters(Real x = 10000.0, Real y = 10000.0, Real z = 10000.0);
  virtual ~sampleClusters();
  virtual MaterialBase* lookupMaterial(Point& pos);
  virtual int geometry() const { return 2; }

  int lookupCluster(Point& pos, Real dr = 0.0);
  void initSpatialhash(int x, int y, int z);
  void addCluster(Real x, Real y, Real z, Real r);
  void addRandomClusters(unsigned int n, Real r, Real dr, SimconfType* simconf);

 protected:
  void clearClusters();
  void clearSpatialHash();
  void reallocClusters(int n);

 private:
  std::vector<int> _sh, _cl;
  std::vector<Real> _c[4];
  void sync();
};

}  // namespace MyTRIM_NS

#endif
