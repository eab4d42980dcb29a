// The C++ API built as inside MOOSE/Magpie (MYTRIM_ENABLED, test doubles of
// the MOOSE and libMesh headers on the include path): Point and Real are
// libMesh's, and the API's errors reach mooseError, as the reference's do in
// that build (material.C:50-71, simconf.C:142-165).  Runs the paths that need
// no GPU and prints what mooseError received.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include "element.h"
#include "material.h"
#include "simconf.h"

using namespace MyTRIM_NS;

int main() {
  SimconfType simconf(1);
  const libMesh::Point p(1.0, 2.0, 2.0);
  std::printf("libMesh point norm %.1f, Cu mass %.3f\n", p.norm(), simconf.scoef[28].m1);
  try {
    MaterialBase m(&simconf, 8.9);
    Element e;
    e._Z = 29;
    e._m = 63.546;
    e._t = 0.0;
    m._element.push_back(e);
    m.prepare();
    std::printf("no error for an all-zero stoichiometry\n");
    return 1;
  } catch (const std::runtime_error& e) {
    std::printf("%s\n", e.what());
  }
  setenv("MYTRIM_DATADIR", "/nonexistent-mytrim-data", 1);
  try {
    SimconfType missing(1);
    std::printf("no error for a missing data directory\n");
    return 1;
  } catch (const std::runtime_error& e) {
    std::printf("%s\n", e.what());
  }
  return 0;
}
