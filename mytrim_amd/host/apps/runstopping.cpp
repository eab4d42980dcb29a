// runstopping: electronic stopping of one ion in one material over a list of
// energies (apps/runstopping.C:50-180), evaluated on the GPU in one batch
// (MaterialBase::getrstop(pka, energies)).  Output lines "E S_e" in eV and
// eV/A with the stream's default formatting, as the reference prints them.
//
// Input (stdin): a "stopping" object with "material" (rho, elements [Z, mass,
// fraction]) and "ion" (Z, mass, energy), where energy is a number, an array,
// or {begin, end, step | mult}.  `//` and `/* */` comments are accepted.
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "element.h"
#include "ion.h"
#include "json_lite.h"
#include "material.h"
#include "simconf.h"

using namespace MyTRIM_NS;

#define mytrimError(msg)                                 \
  do {                                                   \
    std::cerr << "ERROR: " << msg << '\n' << std::flush; \
    return 1;                                            \
  } while (0)

int main(int argc, char* /*argv*/[]) {
  if (argc > 1) mytrimError("Please supply the input file via stdin (e.g. ./runstopping < input.json`)");
  json_lite::Value root;
  try {
    const std::string text((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());
    root = json_lite::parse(text);
  } catch (const std::exception& e) {
    mytrimError(e.what());
  }
  if (!root["stopping"].isObject()) mytrimError("No 'stopping' top level block found in input");
  const json_lite::Value& J = root["stopping"];

  try {
    SimconfType simconf;
    simconf.fullTraj = false;
    simconf.tmin = 0.2;

    const json_lite::Value& jm = J["material"];
    if (!jm.isObject()) mytrimError("Must specify a 'material' block in the input file");
    if (!jm["rho"].isNumber()) mytrimError("Missing 'rho'");
    const json_lite::Value& je = jm["elements"];
    if (!je.isArray()) mytrimError("Missing 'elements' in material");
    MaterialBase material(&simconf, jm["rho"].asDouble());
    Element element;
    for (unsigned int j = 0; j < je.size(); ++j) {
      if (!je[j]["Z"].isNumber()) mytrimError("Missing 'Z' in element " << j);
      element._Z = je[j]["Z"].asInt();
      if (!je[j]["mass"].isNumber()) mytrimError("Missing 'mass' in element " << j);
      element._m = je[j]["mass"].asDouble();
      if (!je[j]["fraction"].isNumber()) mytrimError("Missing 'fraction' in element " << j);
      element._t = je[j]["fraction"].asDouble();
      material._element.push_back(element);
    }
    material.prepare();

    const json_lite::Value& ji = J["ion"];
    if (!ji.isObject()) mytrimError("Must specify an 'ion' block in the input file");
    IonBase pka;
    if (!ji["Z"].isNumber()) mytrimError("Missing 'Z' in ion block");
    pka._Z = ji["Z"].asInt();
    if (!ji["mass"].isNumber()) mytrimError("Missing 'mass' in ion block");
    pka._m = ji["mass"].asDouble();

    // energies: number, array, or range with additive or multiplicative steps
    // (the loops accumulate in Real, as the reference's do)
    std::vector<Real> energies;
    const json_lite::Value& jE = ji["energy"];
    if (jE.isNumber()) {
      energies.push_back(jE.asDouble());
    } else if (jE.isArray()) {
      for (unsigned int j = 0; j < jE.size(); ++j) energies.push_back(jE[j].asDouble());
    } else if (jE.isObject()) {
      if (!jE["begin"].isNumber()) mytrimError("Missing 'begin' in energy block");
      const Real eb = jE["begin"].asDouble();
      if (!jE["end"].isNumber()) mytrimError("Missing 'end' in energy block");
      const Real ee = jE["end"].asDouble();
      const bool has_step = jE["step"].isNumber(), has_mult = jE["mult"].isNumber();
      if (has_step == has_mult) mytrimError("Specify either 'step' or 'mult' energy block");
      if (has_step) {
        const Real es = jE["step"].asDouble();
        if (!(es > 0.0)) mytrimError("'step' must be positive");
        for (Real E = eb; E <= ee; E += es) energies.push_back(E);
      } else {
        const Real em = jE["mult"].asDouble();
        if (em <= 1.0) mytrimError("'mult' must be larger than 1.0");
        for (Real E = eb; E <= ee; E *= em) energies.push_back(E);
      }
    } else {
      mytrimError("Missing or invalid 'energy' in ion block");
    }

    const std::vector<Real> se = material.getrstop(&pka, energies);
    for (size_t n = 0; n < energies.size(); ++n) std::cout << energies[n] << ' ' << se[n] << '\n';
  } catch (const std::exception& e) {
    mytrimError(e.what());
  }
  return EXIT_SUCCESS;
}
