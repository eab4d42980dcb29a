// ORACLE — test infrastructure only.
//
// The loop of the reference's runstopping (apps/runstopping.C:50-180) over the
// reference library compiled from its own sources (oracle/Makefile target
// `ref`).  jsoncpp is not installed here, so the deck's fields come from a
// flat spec file (tests/golden/make_fixtures.py writes it):
//
//   material rho        el Z mass fraction   (one line per element)
//   ion Z mass          energies E1 E2 ...   | step BEGIN END STEP | mult BEGIN END MULT
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "element.h"
#include "ion.h"
#include "material.h"
#include "simconf.h"

using namespace MyTRIM_NS;

int main(int argc, char** argv) {
  if (argc != 2) return 1;
  std::ifstream in(argv[1]);
  SimconfType* simconf = new SimconfType;
  simconf->fullTraj = false;
  simconf->tmin = 0.2;
  MaterialBase* material = nullptr;
  Element element;
  IonBase* pka = new IonBase;
  std::vector<Real> energies;
  std::string line, key;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    if (!(ls >> key)) continue;
    if (key == "material") {
      Real rho;
      ls >> rho;
      material = new MaterialBase(simconf, rho);
    } else if (key == "el") {
      ls >> element._Z >> element._m >> element._t;
      material->_element.push_back(element);
    } else if (key == "ion") {
      ls >> pka->_Z >> pka->_m;
    } else if (key == "energies") {
      Real E;
      while (ls >> E) energies.push_back(E);
    } else if (key == "step") {
      Real b, e, s;
      ls >> b >> e >> s;
      for (Real E = b; E <= e; E += s) energies.push_back(E);
    } else if (key == "mult") {
      Real b, e, m;
      ls >> b >> e >> m;
      for (Real E = b; E <= e; E *= m) energies.push_back(E);
    }
  }
  material->prepare();
  for (unsigned int n = 0; n < energies.size(); ++n) {
    pka->_E = energies[n];
    std::cout << pka->_E << ' ' << material->getrstop(pka) << '\n';
  }
  return 0;
}
