// ORACLE — test infrastructure only.
//
// The reference's runmytrim driver loop (apps/runmytrim.C:71-334) over the
// reference library and the reference's own output modules
// (apps/src/TrimVacCount.C, TrimVacEnergyCount.C, TrimRange.C), all compiled
// from the reference sources (oracle/Makefile target `ref`).  The reference
// reads its input with jsoncpp, which is not installed here, so this harness
// reads the same fields from a flat spec file instead (written by
// tests/golden/make_fixtures.py from the JSON deck):
//
//   type vaccount|vacenergycount|range     base PATH     seed N     threads N
//   scale L            (optional)          primaries_only 0|1  (optional)
//   ion Z mass energy number [final_energy]
//   layer thickness rho                    (then its elements:)
//   el Z mass fraction edisp|- elbind|-    ("-": not given in the deck)
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <list>
#include <queue>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "element.h"
#include "ion.h"
#include "material.h"
#include "sample_layers.h"
#include "simconf.h"

#include "apps/include/ThreadedTrimBase.h"
#include "apps/include/TrimRange.h"
#include "apps/include/TrimVacCount.h"
#include "apps/include/TrimVacEnergyCount.h"

using namespace MyTRIM_NS;

namespace {

struct El {
  int Z;
  double m, t;
  bool has_ed, has_eb;
  double ed, eb;
};
struct Layer {
  double thickness, rho;
  std::vector<El> el;
};

struct ThreadData {
  SimconfType _simconf;
  SampleLayers* _sample;
  ThreadedTrimBase* _trim;
  std::queue<IonBase*> _recoils;
  std::list<IonBase*> _pka;
};
std::vector<ThreadData> thread_data;

void computeThread(int tid) {
  auto& td = thread_data[tid];
  for (auto pka : td._pka) {
    td._simconf.seed(pka->_seed);
    td._recoils.push(pka);
    while (!td._recoils.empty()) {
      IonBase* recoil = td._recoils.front();
      td._recoils.pop();
      td._sample->averages(recoil);
      td._trim->trim(recoil, td._recoils);
      delete recoil;
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::cerr << "usage: runmytrim_ref SPECFILE\n";
    return 1;
  }
  std::ifstream in(argv[1]);
  std::string type = "vaccount", base = "mytrim", key;
  int seed = 0, ion_Z = 1, has_scale = 0, has_po = 0, po = 0, has_ef = 0;
  unsigned nthreads = 1, npka = 0;
  double scale = 1.0, ion_m = 1.0, ion_E = 1.0, ion_Ef = 3.0;
  std::vector<Layer> layers;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    if (!(ls >> key)) continue;
    if (key == "type") ls >> type;
    else if (key == "base") ls >> base;
    else if (key == "seed") ls >> seed;
    else if (key == "threads") ls >> nthreads;
    else if (key == "scale") { ls >> scale; has_scale = 1; }
    else if (key == "primaries_only") { ls >> po; has_po = 1; }
    else if (key == "ion") { ls >> ion_Z >> ion_m >> ion_E >> npka; if (ls >> ion_Ef) has_ef = 1; }
    else if (key == "layer") { Layer L; ls >> L.thickness >> L.rho; layers.push_back(L); }
    else if (key == "el") {
      El e; std::string a, b;
      ls >> e.Z >> e.m >> e.t >> a >> b;
      e.has_ed = a != "-"; e.has_eb = b != "-";
      e.ed = e.has_ed ? std::atof(a.c_str()) : 0.0;
      e.eb = e.has_eb ? std::atof(b.c_str()) : 0.0;
      layers.back().el.push_back(e);
    }
  }

  thread_data.resize(nthreads);
  for (auto& td : thread_data) {
    td._simconf.fullTraj = false;
    td._simconf.tmin = 0.2;
  }
  if (has_scale)
    for (auto& td : thread_data) td._simconf.setLengthScale(scale);
  double thickness = 0.0;
  for (auto& L : layers) thickness += L.thickness;
  for (auto& td : thread_data) td._sample = new SampleLayers(thickness, 100.0, 100.0);
  for (auto& td : thread_data) {
    if (type == "vaccount") td._trim = new TrimVacCount(&td._simconf, td._sample);
    else if (type == "vacenergycount") td._trim = new TrimVacEnergyCount(&td._simconf, td._sample);
    else td._trim = new TrimRange(&td._simconf, td._sample);
    td._trim->setBaseName(base);
    if (has_po) td._trim->_primaries_only = po != 0;
  }
  // one Element object for every element of every layer and thread, as in
  // apps/runmytrim.C:233-273
  Element element;
  for (auto& td : thread_data)
    for (auto& L : layers) {
      MaterialBase* material = new MaterialBase(&td._simconf, L.rho);
      for (auto& e : L.el) {
        element._Z = e.Z;
        element._m = e.m;
        element._t = e.t;
        if (e.has_ed) element._Edisp = e.ed;
        if (e.has_eb) element._Elbind = e.eb;
        material->_element.push_back(element);
      }
      material->prepare();
      td._sample->material.push_back(material);
      td._sample->layerThickness.push_back(L.thickness);
    }

  IonBase* tmpl = new IonBase;
  tmpl->_Z = ion_Z;
  tmpl->_m = ion_m;
  tmpl->_E = ion_E;
  if (has_ef) tmpl->_Ef = ion_Ef;
  thread_data[0]._simconf.seed(seed);
  for (unsigned n = 0; n < npka; n++) {
    auto& td = thread_data[n % nthreads];
    IonBase* pka = new IonBase(tmpl);
    pka->_gen = 0;
    pka->_dir = Point(1.0, 0.0, 0.0);
    pka->_pos = Point(0.0, td._sample->w[1] / 2.0, td._sample->w[2] / 2.0);
    pka->_seed = thread_data[0]._simconf.irand();
    td._pka.push_back(pka);
  }
  std::vector<std::thread> thread(nthreads);
  for (unsigned i = 0; i < nthreads; ++i) thread[i] = std::thread(computeThread, i);
  for (unsigned i = 0; i < nthreads; ++i) thread[i].join();
  for (unsigned i = 1; i < nthreads; ++i) {
    thread_data[0]._trim->threadJoin(*thread_data[i]._trim);
    thread_data[0]._simconf.vacancies_created += thread_data[i]._simconf.vacancies_created;
    thread_data[0]._simconf.EelTotal += thread_data[i]._simconf.EelTotal;
    thread_data[0]._simconf.EnucTotal += thread_data[i]._simconf.EnucTotal;
  }
  thread_data[0]._trim->writeOutput();
  std::cout.precision(17);
  std::cout << "vacancies " << thread_data[0]._simconf.vacancies_created << '\n'
            << "EelTotal " << thread_data[0]._simconf.EelTotal << '\n';
  return 0;
}
