// runmytrim: the JSON-driven MyTRIM front end (apps/runmytrim.C:95-334),
// built on the GPU drop-in API.
//
//   runmytrim < input.json            every flight through TrimBase::trim
//                                     (reference draw order, per-PKA seeds
//                                     from irand(), FIFO drain), so output
//                                     files match the reference's
//   runmytrim --batch < input.json    all cascades in one persistent-kernel
//                                     launch (TrimGPU::runCascades, Philox
//                                     streams); output type vaccount only
//
// Input: a "mytrim" object with "options" (seed, threads, scale), "sample"
// ("layers": thickness, rho, elements [Z, mass, fraction, edisp, elbind]),
// "ion" (Z, mass, energy, number, final_energy) and "output" (type
// vaccount | vacenergycount | range, base, primaries_only).  `//` and `/* */`
// comments are accepted.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <iterator>
#include <list>
#include <memory>
#include <queue>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "TrimRange.h"
#include "TrimVacCount.h"
#include "TrimVacEnergyCount.h"
#include "element.h"
#include "ion.h"
#include "json_lite.h"
#include "material.h"
#include "sample_layers.h"
#include "simconf.h"
#include "trim_gpu.h"

using namespace MyTRIM_NS;

#define mytrimError(msg)                            \
  do {                                              \
    std::cerr << "ERROR: " << msg << '\n' << std::flush; \
    return 1;                                       \
  } while (0)

namespace {

struct LayerSpec {
  Real thickness, rho;
  std::vector<Element> elements;
};

struct ThreadData {
  SimconfType simconf;
  std::unique_ptr<SampleLayers> sample;
  std::unique_ptr<ThreadedTrimBase> trim;
  std::queue<IonBase*> recoils;
  std::list<IonBase*> pka;
};

// FIFO drain of each primary's cascade (apps/runmytrim.C:71-93)
void compute(ThreadData* td, std::string* err) {
  try {
    for (IonBase* pka : td->pka) {
      td->simconf.seed(pka->_seed);
      td->recoils.push(pka);
      while (!td->recoils.empty()) {
        IonBase* r = td->recoils.front();
        td->recoils.pop();
        td->sample->averages(r);
        td->trim->trim(r, td->recoils);
        delete r;
      }
    }
  } catch (const std::exception& e) {
    *err = e.what();
    while (!td->recoils.empty()) {
      delete td->recoils.front();
      td->recoils.pop();
    }
  }
}

}  // namespace

int main(int argc, char* argv[]) {
  bool batch = false;
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "--batch") == 0)
      batch = true;
    else
      mytrimError("Please supply the input file via stdin (e.g. ./runmytrim < input.json`)");
  }

  json_lite::Value root;
  try {
    const std::string text((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());
    root = json_lite::parse(text);
  } catch (const std::exception& e) {
    mytrimError(e.what());
  }
  if (!root["mytrim"].isObject()) mytrimError("No 'mytrim' top level block found in input");
  const json_lite::Value& J = root["mytrim"];
  const json_lite::Value& opt = J["options"];

  unsigned int nthreads = 1;
  if (opt.isObject() && opt["threads"].isNumber()) {
    nthreads = unsigned(opt["threads"].asInt());
    if (nthreads < 1) mytrimError("options.threads must be positive");
    std::cerr << "Using " << nthreads << " threads\n";
  }

  int master_seed = 0;
  bool have_scale = false;
  Real scale = 1.0;
  if (opt.isObject()) {
    if (opt["seed"].isNumber()) {
      master_seed = opt["seed"].asInt();
      std::cerr << "Using provided master seed " << master_seed << '\n';
    } else {
      FILE* urand = std::fopen("/dev/random", "r");
      if (!urand || std::fread(&master_seed, sizeof(int), 1, urand) != 1) mytrimError("Unable to access /dev/random");
      std::fclose(urand);
    }
    if (opt["scale"].isNumber()) {
      have_scale = true;
      scale = opt["scale"].asDouble();
      std::cerr << "Using provided length scale " << scale << '\n';
    }
  }

  // sample
  if (!J["sample"].isObject()) mytrimError("Must specify a 'sample' block in the input file");
  const json_lite::Value& jl = J["sample"]["layers"];
  if (!jl.isArray()) mytrimError("sample.layers must be an array");
  const unsigned int nlayers = unsigned(jl.size());
  std::cerr << "Building " << nlayers << " layers\n";
  Real thickness = 0.0;
  for (unsigned int i = 0; i < nlayers; ++i) {
    if (!jl[i]["thickness"].isNumber()) mytrimError("No 'thickness' found for layer " << i);
    thickness += jl[i]["thickness"].asDouble();
  }

  // output
  const json_lite::Value& jo = J["output"];
  if (!jo.isObject()) mytrimError("Must specify an 'output' block in the input file");
  if (!jo["type"].isString()) mytrimError("output.type must be a string");
  const std::string type = jo["type"].asString();
  if (type != "vaccount" && type != "vacenergycount" && type != "range") mytrimError("Unknown output type " << type);
  if (batch && type != "vaccount") mytrimError("--batch supports output type vaccount only");

  // layers and their materials; one Element object is filled for every
  // element in turn, so an edisp / elbind given once carries over to the
  // elements after it (apps/runmytrim.C:233-273).  The reference repeats this
  // loop per thread with the same object, so its threads after the first can
  // inherit the last element's values; here every thread gets thread 0's.
  std::vector<LayerSpec> layers;
  Element element;
  for (unsigned int i = 0; i < nlayers; ++i) {
    if (!jl[i]["rho"].isNumber()) mytrimError("Missing 'rho' in layer " << i);
    const json_lite::Value& je = jl[i]["elements"];
    if (!je.isArray()) mytrimError("Missing 'elements' in layer " << i);
    LayerSpec L;
    L.thickness = jl[i]["thickness"].asDouble();
    L.rho = jl[i]["rho"].asDouble();
    for (unsigned int j = 0; j < je.size(); ++j) {
      if (!je[j]["Z"].isNumber()) mytrimError("Missing 'Z' in element " << j << " in layer " << i);
      element._Z = je[j]["Z"].asInt();
      if (!je[j]["mass"].isNumber()) mytrimError("Missing 'mass' in element " << j << " in layer " << i);
      element._m = je[j]["mass"].asDouble();
      if (!je[j]["fraction"].isNumber()) mytrimError("Missing 'fraction' in element " << j << " in layer " << i);
      element._t = je[j]["fraction"].asDouble();
      if (je[j]["edisp"].isNumber()) element._Edisp = je[j]["edisp"].asDouble();
      if (je[j]["elbind"].isNumber()) element._Elbind = je[j]["elbind"].asDouble();
      L.elements.push_back(element);
    }
    layers.push_back(L);
  }

  // ion
  const json_lite::Value& ji = J["ion"];
  if (!ji.isObject()) mytrimError("Must specify an 'ion' block in the input file");
  IonBase tmpl;
  if (!ji["Z"].isNumber()) mytrimError("Missing 'Z' in ion block");
  tmpl._Z = ji["Z"].asInt();
  if (!ji["mass"].isNumber()) mytrimError("Missing 'mass' in ion block");
  tmpl._m = ji["mass"].asDouble();
  if (!ji["energy"].isNumber()) mytrimError("Missing 'energy' in ion block");
  tmpl._E = ji["energy"].asDouble();
  if (!ji["number"].isNumber()) mytrimError("Missing 'number' in ion block");
  const unsigned int npka = unsigned(ji["number"].asInt());
  if (ji["final_energy"].isNumber()) tmpl._Ef = ji["final_energy"].asDouble();

  // per-thread state (apps/runmytrim.C:112-120, 181-279)
  std::vector<ThreadData> td(nthreads);
  const bool primaries_only = jo["primaries_only"].isBool() && jo["primaries_only"].asBool();
  try {
    for (auto& t : td) {
      t.simconf.fullTraj = false;
      t.simconf.tmin = 0.2;
      if (have_scale) t.simconf.setLengthScale(scale);
      t.sample.reset(new SampleLayers(thickness, 100.0, 100.0));
      if (type == "vaccount")
        t.trim.reset(new TrimVacCount(&t.simconf, t.sample.get()));
      else if (type == "vacenergycount")
        t.trim.reset(new TrimVacEnergyCount(&t.simconf, t.sample.get()));
      else
        t.trim.reset(new TrimRange(&t.simconf, t.sample.get()));
      if (jo["base"].isString()) t.trim->setBaseName(jo["base"].asString());
      if (jo["primaries_only"].isBool()) t.trim->_primaries_only = primaries_only;
      for (const LayerSpec& L : layers) {
        MaterialBase* m = new MaterialBase(&t.simconf, L.rho);
        m->_element = L.elements;
        m->prepare();
        t.sample->material.push_back(m);
        t.sample->layerThickness.push_back(L.thickness);
      }
    }
  } catch (const std::exception& e) {
    mytrimError(e.what());
  }

  if (batch) {
    // one launch for every cascade; vacancy / replacement depths binned on the
    // device from x = -1 in 1 A bins, merged into the reference's int(x) bins
    ThreadData& t = td[0];
    TrimGPU gpu(&t.simconf, t.sample.get());
    TrimGPU::Options o;
    o.follow_max_gen = primaries_only ? 1 : -1;
    o.vac_x0 = -1.0;
    o.vac_bin = 1.0;
    o.vac_bins = int(std::ceil(4.0 * thickness)) + 1024;
    IonBase pka(&tmpl);
    pka._gen = 0;
    pka._dir = Point(1.0, 0.0, 0.0);
    pka._pos = Point(0.0, t.sample->w[1] / 2.0, t.sample->w[2] / 2.0);
    TrimGPU::Tallies r;
    try {
      r = gpu.runCascades(pka, npka, uint32_t(master_seed), o);
    } catch (const std::exception& e) {
      mytrimError(e.what());
    }
    const size_t nb = r.vac_hist.size();  // [0] x < -1, [1] (-1, 0), [2] [0, 1), ..., [nb - 1] overflow
    if (r.vac_hist[nb - 1] || r.repl_hist[nb - 1])
      std::cerr << "WARNING: " << r.vac_hist[nb - 1] + r.repl_hist[nb - 1]
                << " vacancies/replacements deeper than the binned range\n";
    std::vector<unsigned int>& vb = static_cast<TrimVacCount*>(t.trim.get())->_vac_bin;
    std::vector<unsigned int>& rb = static_cast<TrimVacCount*>(t.trim.get())->_repl_bin;
    size_t last = 0;
    for (size_t i = 2; i + 1 < nb; ++i)
      if (r.vac_hist[i] || r.repl_hist[i]) last = i;
    const size_t n_out = last >= 2 ? last - 1 : ((r.vac_hist[1] || r.repl_hist[1]) ? 1 : 0);
    vb.assign(n_out, 0);
    rb.assign(n_out, 0);
    for (size_t x = 0; x < n_out; ++x) {
      vb[x] = unsigned(r.vac_hist[x + 2] + (x == 0 ? r.vac_hist[1] : 0));
      rb[x] = unsigned(r.repl_hist[x + 2] + (x == 0 ? r.repl_hist[1] : 0));
    }
    t.trim->writeOutput();
    std::cerr << "Vacancies/ion: " << Real(r.vacancies) / Real(npka) << '\n'
              << "Electronic energy loss/ion: " << r.Eel / Real(npka) << '\n'
              << "Collisions: " << r.collisions << ", kernel " << r.kernel_ms << " ms\n";
    return EXIT_SUCCESS;
  }

  // per-PKA seeds from thread 0's generator seeded with the master seed
  // (apps/runmytrim.C:282-304); primaries dealt round-robin to the threads
  td[0].simconf.seed(unsigned(master_seed));
  for (unsigned int n = 0; n < npka; ++n) {
    ThreadData& t = td[n % nthreads];
    IonBase* pka = new IonBase(&tmpl);
    pka->_gen = 0;
    pka->_dir = Point(1.0, 0.0, 0.0);
    pka->_pos = Point(0.0, t.sample->w[1] / 2.0, t.sample->w[2] / 2.0);
    pka->_seed = td[0].simconf.irand();
    t.pka.push_back(pka);
  }

  std::vector<std::string> errs(nthreads);
  std::vector<std::thread> threads;
  for (unsigned int i = 0; i < nthreads; ++i) threads.emplace_back(compute, &td[i], &errs[i]);
  for (auto& th : threads) th.join();
  for (unsigned int i = 0; i < nthreads; ++i)
    if (!errs[i].empty()) mytrimError("thread " << i << ": " << errs[i]);

  for (unsigned int i = 1; i < nthreads; ++i) {
    td[0].trim->threadJoin(*td[i].trim);
    td[0].simconf.vacancies_created += td[i].simconf.vacancies_created;
    td[0].simconf.EelTotal += td[i].simconf.EelTotal;
    td[0].simconf.EnucTotal += td[i].simconf.EnucTotal;
  }
  td[0].trim->writeOutput();
  std::cerr << "Vacancies/ion: " << Real(td[0].simconf.vacancies_created) / Real(npka) << '\n'
            << "Electronic energy loss/ion: " << td[0].simconf.EelTotal / Real(npka) << '\n';
  return EXIT_SUCCESS;
}
