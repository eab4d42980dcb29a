// SampleBase, SampleLayers, sampleClusters, SampleWire and SampleBurriedWire
// geometry on the host (sample.C, sample_layers.C, sample_clusters.C,
// sample_wire.C, sample_burried_wire.C).  The device has its own copies of
// the lookups; these serve sample construction and host queries.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "functions.h"
#include "sample.h"
#include "sample_clusters.h"
#include "sample_burried_wire.h"
#include "sample_dynamic.h"
#include "sample_layers.h"
#include "sample_wire.h"

namespace MyTRIM_NS {

SampleBase::SampleBase(Real x, Real y, Real z) {
  w[0] = x;
  w[1] = y;
  w[2] = z;
  for (int i = 0; i < 3; ++i) bc[i] = PBC;
}

void SampleBase::averages(const IonBase* pka) {
  for (auto* m : material) m->average(pka);
}

Real SampleBase::rangeMaterial(Point&, Point&) { return 100000.0; }

SampleWire::SampleWire(Real x, Real y, Real z) : SampleBase(x, y, z) {  // sample_wire.C:30-34
  bc[0] = CUT;
  bc[1] = CUT;
}

MaterialBase* SampleWire::lookupMaterial(Point& pos) {  // sample_wire.C:37-46
  const Real x = (pos(0) / w[0]) * 2.0 - 1.0;
  const Real y = (pos(1) / w[1]) * 2.0 - 1.0;
  return (x * x + y * y) > 1.0 ? nullptr : material[0];
}

SampleBurriedWire::SampleBurriedWire(Real x, Real y, Real z) : SampleWire(x, y, z) {  // sample_burried_wire.C:29-34
  bc[0] = INF;
  bc[1] = INF;
  bc[2] = INF;
}

MaterialBase* SampleBurriedWire::lookupMaterial(Point& pos) {  // sample_burried_wire.C:37-55
  if (pos(2) < 0.0 && pos(2) >= -250.0) return material[1];  // cover layer
  if (pos(2) > w[2] || pos(2) < -250.0) return nullptr;      // above the sample or below the cover
  MaterialBase* m = SampleWire::lookupMaterial(pos);
  return m ? m : material[1];
}

SampleDynamic::SampleDynamic(SimconfType* simconf, Real x, Real y, Real z)  // sample_dynamic.C:28-34
    : SampleLayers(x, y, z), pka(nullptr), _simconf(simconf) {
  bc[0] = CUT;
  bc[1] = PBC;
  bc[2] = PBC;
}

void SampleDynamic::averages(const IonBase* p) {  // sample_dynamic.C:36-46: averages on demand
  pka = p;
  for (MaterialBase* m : material) m->_dirty = true;
}

MaterialBase* SampleDynamic::lookupMaterial(Point& pos) {  // sample_dynamic.C:48-62
  MaterialBase* m = material[lookupLayer(pos)];
  if (m->_dirty && pka) m->average(pka);
  return m;
}

void SampleDynamic::addAtomsToLayer(int layer, int n, int Z) {  // sample_dynamic.C:64-119
  MaterialBase* mat = material[layer];
  std::vector<Element>& el = mat->_element;
  // the element's index in the layer, and the volume of one atom of the
  // layer's composition at natural density (weighted 1/atrho)
  unsigned ne = unsigned(el.size());
  Real rnorm = 0.0;
  for (unsigned i = 0; i < el.size(); ++i) {
    if (el[i]._Z == Z) ne = i;
    rnorm += el[i]._t / _simconf->scoef[el[i]._Z - 1].atrho;
  }
  (void)rnorm;  // computed, not used, as in the reference
  if (ne == el.size()) {  // a new element of the layer
    Element e;
    e._Z = Z;
    e._m = _simconf->scoef[Z - 1].m1;
    e._t = 0.0;
    el.push_back(e);
  }
  const int nl = int(mat->_arho * w[1] * w[2] * layerThickness[layer]);  // atoms in the layer
  const Real ma = 0.6022 * el[ne]._m * Real(n);                         // mass change
  layerThickness[layer] += ma / (mat->_rho * w[1] * w[2]);               // density kept
  if (el[ne]._t * nl + Real(n) < 0.0) {
    std::cout << "Crap, t*nl=" << el[ne]._t * nl << ", but n=" << n << std::endl;
    el[ne]._t = 0.0;
  } else {
    el[ne]._t += Real(n) / nl;  // prepare() renormalises the stoichiometry
  }
  mat->_dirty = true;
  mat->prepare();
}

int SampleLayers::lookupLayer(Point& pos) {  // sample_layers.C:26-43
  unsigned int i;
  Real d = 0.0;
  for (i = 0; i < layerThickness.size(); ++i) {
    d += layerThickness[i];
    if (pos(0) < d) break;
  }
  if (i >= material.size()) i = unsigned(material.size()) - 1;
  return int(i);
}

MaterialBase* SampleLayers::lookupMaterial(Point& pos) { return material[lookupLayer(pos)]; }

Real SampleLayers::rangeMaterial(Point& pos, Point& dir) {  // sample_layers.C:51-92
  Real d = 0.0;
  const Real unrestricted = 1.0e6;
  if (dir(0) == 0.0) return unrestricted;
  const Real epsilon = std::abs(1.0e-10 / dir(0));
  if (pos(0) < 0.0) return dir(0) < 0.0 ? unrestricted : -pos(0) / dir(0) + epsilon;
  for (unsigned i = 0; i < layerThickness.size(); ++i) {
    if (pos(0) >= d && pos(0) < d + layerThickness[i])
      return dir(0) < 0 ? (d - pos(0)) / dir(0) + epsilon : (d + layerThickness[i] - pos(0)) / dir(0) + epsilon;
    d += layerThickness[i];
  }
  return dir(0) > 0.0 ? unrestricted : (d - pos(0)) / dir(0) + epsilon;
}

sampleClusters::sampleClusters(Real x, Real y, Real z) : SampleBase(x, y, z), sd(0), sh(nullptr), cl(nullptr), cn(0), cnm(0), cmr(0) {
  for (int i = 0; i < 3; ++i) {
    kd[i] = 0;
    kn[i] = 0;
  }
  for (int i = 0; i < 4; ++i) c[i] = nullptr;
}

sampleClusters::~sampleClusters() {}

void sampleClusters::sync() {
  sh = _sh.empty() ? nullptr : _sh.data();
  cl = _cl.empty() ? nullptr : _cl.data();
  for (int i = 0; i < 4; ++i) c[i] = _c[i].empty() ? nullptr : _c[i].data();
}

MaterialBase* sampleClusters::lookupMaterial(Point& pos) {  // sample_clusters.C:43-55
  const int l = lookupCluster(pos, 0.0);
  if (l == -2) return nullptr;
  if (l == -1) return material[0];
  material[1]->_tag = l;
  return material[1];
}

int sampleClusters::lookupCluster(Point& pos, Real dr) {  // sample_clusters.C:59-133
  int k[3], k1[3], k2[3], j[3];
  for (int i = 0; i < 3; ++i) {
    k[i] = int(std::floor((pos(i) * kn[i]) / w[i]));
    if (pos(i) < 0.0 || pos(i) >= w[i]) {
      if (bc[i] == CUT) return -2;
      if (bc[i] == INF) return -1;
      k[i] = k[i] % kn[i];
      if (k[i] < 0) k[i] += kn[i];
    }
    const int ks = int((cmr + dr) / kd[i]) + 1;
    k1[i] = k[i] - ks;
    k2[i] = k[i] + ks;
    if (k1[i] < 0 && bc[i] != PBC) k1[i] = 0;
    if (k2[i] >= kn[i] && bc[i] != PBC) k2[i] = kn[i] - 1;
  }
  for (j[0] = k1[0]; j[0] <= k2[0]; ++j[0])
    for (j[1] = k1[1]; j[1] <= k2[1]; ++j[1])
      for (j[2] = k1[2]; j[2] <= k2[2]; ++j[2]) {
        int m[3];
        for (int i = 0; i < 3; ++i) {
          m[i] = j[i] % kn[i];
          if (m[i] < 0) m[i] += kn[i];
        }
        int l = _sh[m[0] + kn[0] * (m[1] + kn[1] * m[2])];
        while (l >= 0) {
          Real dif[3];
          for (int i = 0; i < 3; ++i) {
            dif[i] = pos(i) - _c[i][l];
            if (bc[i] == PBC) dif[i] -= ::round(dif[i] / w[i]) * w[i];
          }
          if (v_dot(dif, dif) < sqr(_c[3][l] + dr)) return l;
          l = _cl[l];
        }
      }
  return -1;
}

void sampleClusters::initSpatialhash(int x, int y, int z) {  // sample_clusters.C:135-153
  kn[0] = x;
  kn[1] = y;
  kn[2] = z;
  _sh.assign(size_t(x) * y * z, -1);
  sd = 0.0;
  for (int i = 0; i < 3; ++i) {
    kd[i] = w[i] / Real(kn[i]);
    sd += kd[i];
  }
  sd = 0.5 * std::sqrt(sd);
  cmr = 0.0;
  sync();
}

void sampleClusters::clearSpatialHash() {
  std::fill(_sh.begin(), _sh.end(), -1);
}

void sampleClusters::reallocClusters(int n) {
  if (n > cnm) {
    _cl.resize(n, -1);
    for (int i = 0; i < 4; ++i) _c[i].resize(n, 0.0);
    cnm = n;
    sync();
  }
}

void sampleClusters::clearClusters() {
  cn = 0;
  clearSpatialHash();
  std::fill(_cl.begin(), _cl.end(), -1);
}

void sampleClusters::addCluster(Real x, Real y, Real z, Real r) {  // sample_clusters.C:187-224
  if (cn >= cnm) reallocClusters(cnm + cnm / 10 + 10);
  _c[0][cn] = x;
  _c[1][cn] = y;
  _c[2][cn] = z;
  _c[3][cn] = r;
  int k[3];
  for (int i = 0; i < 3; ++i) {
    k[i] = int(std::floor((_c[i][cn] * kn[i]) / w[i])) % kn[i];
    k[i] = (k[i] < 0) ? k[i] + kn[i] : k[i];
  }
  int l = k[0] + kn[0] * (k[1] + kn[1] * k[2]);
  if (_sh[l] < 0) {
    _sh[l] = cn;
  } else {
    l = _sh[l];
    while (_cl[l] >= 0) l = _cl[l];
    _cl[l] = cn;
  }
  _cl[cn] = -1;
  if (r > cmr) cmr = r;
  cn++;
}

void sampleClusters::addRandomClusters(unsigned int n, Real r, Real dr, SimconfType* simconf) {
  reallocClusters(int(n + n / 10));
  Point npos;
  for (unsigned int i = 0; i < n; ++i) {
    while (true) {
      for (int j = 0; j < 3; ++j) npos(j) = simconf->drand() * w[j];
      if (lookupCluster(npos, dr + r) == -1) {
        addCluster(npos(0), npos(1), npos(2), r);
        break;
      }
    }
  }
}

}  // namespace MyTRIM_NS
