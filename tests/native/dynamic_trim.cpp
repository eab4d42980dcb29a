// The drain loop of apps/mytrim_dynamic.C (every flight's atom leaves the
// layer it starts in and joins the layer it stops in, through
// SampleDynamic::addAtomsToLayer) with a fixed seed, a short Xe series and a
// small two-layer sample.  The same source builds against the reference
// library (tests/golden/make_fixtures.py, the fixture) and against the C++
// drop-in API, whose TrimBase::trim runs every flight on the GPU
// (tests/test_gpu_dynamic.py).  Prints one line per flight (generation, Z,
// final position, state) and the layers at the end.
#include <cstdio>
#include <cstdlib>
#include <queue>

#include "element.h"
#include "ion.h"
#include "material.h"
#include "sample_dynamic.h"
#include "simconf.h"
#include "trim.h"

using namespace MyTRIM_NS;

int main(int argc, char** argv) {
  const int nions = argc > 1 ? std::atoi(argv[1]) : 12;
  SimconfType* simconf = new SimconfType(20231);
  simconf->fullTraj = false;
  simconf->tmin = 0.2;
  SampleDynamic* sample = new SampleDynamic(simconf, 500.0, 100.0, 100.0);
  TrimBase* trim = new TrimBase(simconf, sample);

  Element element;
  MaterialBase* si = new MaterialBase(simconf, 2.33);
  element._Z = 14;
  element._m = 28.086;
  element._t = 1.0;
  si->_element.push_back(element);
  si->prepare();
  sample->material.push_back(si);
  sample->layerThickness.push_back(100.0);
  MaterialBase* ox = new MaterialBase(simconf, 2.2);
  ox->_element.push_back(element);
  element._Z = 8;
  element._m = 15.999;
  element._t = 2.0;
  ox->_element.push_back(element);
  ox->prepare();
  sample->material.push_back(ox);
  sample->layerThickness.push_back(400.0);

  std::queue<IonBase*> recoils;
  for (int n = 0; n < nions; ++n) {
    IonBase* ff1 = new IonBase;
    ff1->_gen = 0;
    ff1->_tag = -1;
    ff1->_id = simconf->_id++;
    ff1->_Z = 54;
    ff1->_m = 131.0;
    ff1->_E = 1.0e4 * (3 - ((n * 3) / nions));
    ff1->_dir(0) = 1;
    ff1->_dir(1) = 0;
    ff1->_dir(2) = 0;
    ff1->_pos(0) = 0;
    ff1->_pos(1) = sample->w[1] / 2.0;
    ff1->_pos(2) = sample->w[2] / 2.0;
    ff1->setEf();
    recoils.push(ff1);
    while (!recoils.empty()) {
      IonBase* pka = recoils.front();
      recoils.pop();
      sample->averages(pka);
      const int layer1 = sample->lookupLayer(pka->_pos);
      if (pka->_gen > 0) sample->addAtomsToLayer(layer1, -1, pka->_Z);
      trim->trim(pka, recoils);
      const int layer2 = sample->lookupLayer(pka->_pos);
      if (pka->_pos(0) > 0) sample->addAtomsToLayer(layer2, 1, pka->_Z);
      std::printf("F %d %d %.12g %.12g %.12g %d\n", pka->_gen, pka->_Z, pka->_pos(0), pka->_pos(1), pka->_pos(2),
                  int(pka->_state));
      delete pka;
    }
  }
  for (unsigned int i = 0; i < sample->material.size(); ++i) {
    std::printf("L %.12g", sample->layerThickness[i]);
    for (unsigned int j = 0; j < sample->material[i]->_element.size(); ++j)
      std::printf(" %d %.12g", sample->material[i]->_element[j]._Z, sample->material[i]->_element[j]._t);
    std::printf("\n");
  }
  return 0;
}
