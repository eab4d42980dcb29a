// Fission fragments in UO2 with Xe bubbles, run through TrimBase::trim on
// the GPU.  Same command line, seeding, sample set-up, drain loop and output
// files as the reference's apps/mytrim_uo2.C (PLAIN mode), so its golden
// files (tests/uo2/gold) apply unchanged:
//   MYTRIM_SEED=39172 mytrim_uo2 out 10 0.1 1
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <queue>

#include "element.h"
#include "functions.h"
#include "invert.h"
#include "ion.h"
#include "material.h"
#include "sample_clusters.h"
#include "simconf.h"
#include "trim.h"

using namespace MyTRIM_NS;

int main(int argc, char* argv[]) {
  if (argc != 5) {
    std::cerr << "syntax:\n" << argv[0] << " basename r Cbfactor Nev\n\n"
              << "r Bubble radius in Ang\nCbfactor=1 => 7e-4 bubbles/nm^3\n"
              << "Nev  number of fission events (two fragments each)\n";
    return 1;
  }
  SimconfType* simconf = new SimconfType;
  int seed;
  if (const char* s = std::getenv("MYTRIM_SEED")) {
    seed = std::atoi(s);
  } else {
    FILE* ur = std::fopen("/dev/urandom", "r");
    if (!ur || std::fread(&seed, sizeof(int), 1, ur) != 1) return 1;
    std::fclose(ur);
  }
  simconf->seed(seed < 0 ? -seed : seed);

  sampleClusters* sample = new sampleClusters(400.0, 400.0, 400.0);
  TrimBase* trim = new TrimBase(simconf, sample);
  const Real r = std::atof(argv[2]);
  const Real Cbf = std::atof(argv[3]);
  const int Nev = std::atoi(argv[4]);

  sample->initSpatialhash(int(sample->w[0] / r) - 1, int(sample->w[1] / r) - 1, int(sample->w[2] / r) - 1);
  const Real v_sam = sample->w[0] * sample->w[1] * sample->w[2];
  const int n_cl = int(v_sam * 7.0e-7 * Cbf);  // 7e-4 bubbles / nm^3 at Cbf = 1
  std::cerr << "adding " << n_cl << " clusters...\n";
  sample->addRandomClusters(n_cl, r, 25.0, simconf);

  char fname[256];
  std::snprintf(fname, sizeof(fname), "%s.clcoor", argv[1]);
  FILE* ccf = std::fopen(fname, "wt");
  for (int i = 0; i < sample->cn; ++i)
    std::fprintf(ccf, "%f %f %f %f %d\n", sample->c[0][i], sample->c[1][i], sample->c[2][i], sample->c[3][i], i);
  std::fclose(ccf);

  Element element;
  MaterialBase* uo2 = new MaterialBase(simconf, 10.0);
  element._Z = 92;
  element._m = 235.0;
  element._t = 1.0;
  uo2->_element.push_back(element);
  element._Z = 8;
  element._m = 16.0;
  element._t = 2.0;
  uo2->_element.push_back(element);
  uo2->prepare();
  sample->material.push_back(uo2);

  const int gas_z1 = 54;
  MaterialBase* xe = new MaterialBase(simconf, 3.5);
  element._Z = gas_z1;
  element._m = 132.0;
  element._t = 1.0;
  xe->_element.push_back(element);
  xe->prepare();
  sample->material.push_back(xe);

  std::queue<IonBase*> recoils;
  MassInverter* minv = new MassInverter;
  EnergyInverter* einv = new EnergyInverter;
  std::snprintf(fname, sizeof(fname), "%s.Erec", argv[1]);
  FILE* erec = std::fopen(fname, "wt");
  std::snprintf(fname, sizeof(fname), "%s.dist", argv[1]);
  FILE* rdist = std::fopen(fname, "wt");
  Real dif[3], pos1[3] = {0, 0, 0};

  for (int n = 0; n < Nev; ++n) {
    IonMDTag* ff1 = new IonMDTag;
    ff1->_gen = 0;
    ff1->_tag = -1;
    ff1->_md = 0;
    const Real A1 = minv->x(simconf->drand());
    const Real A2 = 235.0 - A1;
    einv->setMass(A1);
    const Real Etot = einv->x(simconf->drand());
    const Real E1 = Etot * A2 / (A1 + A2);
    const Real E2 = Etot - E1;
    const int Z1 = int(std::round((A1 * 92.0) / 235.0));
    const int Z2 = 92 - Z1;
    ff1->_Z = Z1;
    ff1->_m = A1;
    ff1->_E = E1 * 1.0e6;
    Real norm;
    do {
      for (int i = 0; i < 3; ++i) ff1->_dir(i) = 2.0 * simconf->drand() - 1.0;
      norm = ff1->_dir.norm_sq();
    } while (norm <= 0.0001 || norm > 1.0);
    ff1->_dir /= std::sqrt(norm);
    for (int i = 0; i < 3; ++i) ff1->_pos(i) = simconf->drand() * sample->w[i];
    ff1->setEf();
    recoils.push(ff1);

    IonMDTag* ff2 = new IonMDTag(*ff1);
    ff2->_dir = -ff2->_dir;
    ff2->_Z = Z2;
    ff2->_m = A2;
    ff2->_E = E2 * 1.0e6;
    ff2->_md = 0;
    ff2->setEf();
    recoils.push(ff2);
    std::cout << "A1=" << A1 << " Z1=" << Z1 << " (" << E1 << " MeV)\tA2=" << A2 << " Z2=" << Z2 << " (" << E2
              << " MeV)\n";
    const Real Efiss = ff1->_E + ff2->_E;

    while (!recoils.empty()) {
      IonMDTag* pka = dynamic_cast<IonMDTag*>(recoils.front());
      recoils.pop();
      sample->averages(pka);
      if (pka->_Z == gas_z1) {
        if (pka->_E > 200 && pka->_E < 12000 && pka->_md == 0) pka->_md = 1;
        if (pka->_gen > 0) std::fprintf(erec, "%f\t%d\t%d\n", pka->_E, pka->_gen, pka->_md);
        if (pka->_tag >= 0)
          for (int i = 0; i < 3; ++i) {
            dif[i] = sample->c[i][pka->_tag] - pka->_pos(i);
            if (sample->bc[i] == SampleBase::PBC) dif[i] -= std::round(dif[i] / sample->w[i]) * sample->w[i];
            pos1[i] = pka->_pos(i) + dif[i];
          }
      }
      trim->trim(pka, recoils);
      if (pka->_Z == gas_z1 && pka->_tag >= 0) {
        for (int i = 0; i < 3; ++i) dif[i] = pos1[i] - pka->_pos(i);
        std::fprintf(rdist, "%f %d %f %f %f\n", std::sqrt(v_dot(dif, dif)), pka->_md, pka->_pos(0), pka->_pos(1),
                     pka->_pos(2));
      }
      delete pka;
    }
    std::cout << simconf->EelTotal << '\n' << simconf->EnucTotal << '\n'
              << Efiss - (simconf->EelTotal + simconf->EnucTotal) << '\n';
    simconf->EelTotal = 0.0;
    simconf->EnucTotal = 0.0;
  }
  std::fclose(rdist);
  std::fclose(erec);
  return 0;
}
