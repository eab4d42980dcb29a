// SampleDynamic::addAtomsToLayer through the C++ drop-in API, on the sequence
// the reference fixture harness runs (oracle/ref_harness.cpp, `dynamic`);
// prints the layer thicknesses and compositions after each step in the same
// format.  Host code only: no device call is made.
#include <cstdio>

#include "material.h"
#include "sample_dynamic.h"
#include "simconf.h"

using namespace MyTRIM_NS;

int main() {
  SimconfType sc;
  SampleDynamic smp(&sc, 1000.0, 100.0, 100.0);
  MaterialBase* si = new MaterialBase(&sc, 2.33);
  Element e;
  e._Z = 14;
  e._m = 28.086;
  e._t = 1.0;
  si->_element.push_back(e);
  si->prepare();
  MaterialBase* ox = new MaterialBase(&sc, 2.2);
  ox->_element.push_back(e);
  e._Z = 8;
  e._m = 15.999;
  e._t = 2.0;
  ox->_element.push_back(e);
  ox->prepare();
  smp.material.push_back(si);
  smp.material.push_back(ox);
  smp.layerThickness.push_back(500.0);
  smp.layerThickness.push_back(10000.0);
  const int steps[][3] = {{0, 1000, 5}, {0, -10, 14}, {1, 500, 5}, {0, 2000, 32}, {1, -100000000, 8}, {1, 7, 8}};
  for (const auto& st : steps) {
    smp.addAtomsToLayer(st[0], st[1], st[2]);
    std::printf("T %.17g %.17g\n", smp.layerThickness[0], smp.layerThickness[1]);
    for (MaterialBase* m : smp.material) {
      std::printf("M %.17g %.17g %.17g %.17g", m->_rho, m->_arho, m->_am, m->_az);
      for (const Element& x : m->_element) std::printf(" %d %.17g %.17g", x._Z, x._m, x._t);
      std::printf("\n");
    }
  }
  return 0;
}
