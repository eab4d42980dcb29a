// SampleDynamic between batches through the C++ API: TrimGPU::runCascades
// with the layer-transport tally on a 100 A Si / 400 A SiO2 stack (30 keV Xe,
// histories 0..n-1 of stream 2344), TrimGPU::applyTransport, and a second
// batch (histories n..2n-1) on the updated layers.  Prints, per batch, the
// non-zero transport entries ("T layer Z n") and then the layers
// ("L thickness rho n_el" + "Z m t Edisp Elbind" per element);
// tests/test_dynamic.py reruns both batches through the Python engine from
// the printed layers.
#include <cstdio>
#include <cstdlib>

#include "element.h"
#include "ion.h"
#include "material.h"
#include "sample_dynamic.h"
#include "simconf.h"
#include "mytrim_gpu.h"
#include "trim_gpu.h"

using namespace MyTRIM_NS;

static void print_layers(SampleDynamic* s) {
  for (unsigned int i = 0; i < s->material.size(); ++i) {
    const MaterialBase* m = s->material[i];
    std::printf("L %.17g %.17g %d", s->layerThickness[i], m->_rho, int(m->_element.size()));
    for (const Element& e : m->_element) std::printf(" %d %.17g %.17g %.17g %.17g", e._Z, e._m, e._t, e._Edisp, e._Elbind);
    std::printf("\n");
  }
}

int main(int argc, char** argv) {
  const long long n = argc > 1 ? std::atoll(argv[1]) : 200;
  SimconfType* simconf = new SimconfType(1);
  simconf->tmin = 0.2;
  SampleDynamic* sample = new SampleDynamic(simconf, 500.0, 100.0, 100.0);
  Element si;
  si._Z = 14;
  si._m = 28.086;
  si._t = 1.0;
  si._Edisp = 25.0;
  MaterialBase* m0 = new MaterialBase(simconf, 2.33);
  m0->_element.push_back(si);
  m0->prepare();
  MaterialBase* m1 = new MaterialBase(simconf, 2.2);
  m1->_element.push_back(si);
  Element ox;
  ox._Z = 8;
  ox._m = 15.999;
  ox._t = 2.0;
  ox._Edisp = 28.0;
  m1->_element.push_back(ox);
  m1->prepare();
  sample->material.push_back(m0);
  sample->layerThickness.push_back(100.0);
  sample->material.push_back(m1);
  sample->layerThickness.push_back(400.0);

  TrimGPU trim(simconf, sample);
  IonBase xe;
  xe._Z = 54;
  xe._m = 131.0;
  xe._E = 3.0e4;
  xe._pos = Point(0.0, 50.0, 50.0);
  xe._dir = Point(1.0, 0.0, 0.0);
  xe.setEf();
  TrimGPU::Options opt;
  opt.log_mask = MT_LOG_TRANSPORT;
  opt.max_records = 0;
  for (int batch = 0; batch < 2; ++batch) {
    print_layers(sample);
    const TrimGPU::Tallies t = trim.runCascades(xe, n, 2344u, opt, batch * n);
    for (size_t k = 0; k < sample->material.size(); ++k)
      for (int Z = 0; Z < 93; ++Z)
        if (t.transport[k * 93 + Z]) std::printf("T %d %d %lld\n", int(k), Z, t.transport[k * 93 + Z]);
    std::printf("B %lld %lld\n", t.collisions, t.vacancies);
    TrimGPU::applyTransport(*sample, t);
  }
  print_layers(sample);
  return 0;
}
