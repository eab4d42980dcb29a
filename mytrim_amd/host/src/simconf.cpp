// SimconfType (simconf.C): defaults, seeding, length scale, tables.
#include "simconf.h"
#include "api_error.h"

#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "mytrim_gpu.h"

namespace MyTRIM_NS {

SimconfType::SimconfType(unsigned int s)
    : _id(0),
      scoef(_rows),
      _gen(new std::mt19937(s)),
      _dis_real(0.0, 1.0),
      _dis_int(0, 65535) {
  ed = 25.0;
  tmin = 0.2;
  tau = 0.0;
  da = 3.0;
  cw = 0.001;
  fullTraj = false;
  vacancies_created = 0;
  EelTotal = 0.0;
  EnucTotal = 0.0;
  setLengthScale(1.0);
  const char* dir = std::getenv("MYTRIM_DATADIR");
  if (dir) {
    readDataFiles(dir);
  } else {
    std::vector<double> rows(92 * 16);
    if (mt_default_tables(rows.data(), 92, 16) != MT_OK) api_error(mt_last_error());
    for (unsigned i = 0; i < _rows; ++i) {
      const double* v = &rows[16 * i];
      ScoefLine& c = scoef[i];
      c.mm1 = v[0];
      c.m1 = v[1];
      c.mnat = v[2];
      c.rho = v[3];
      c.atrho = v[4];
      c.vfermi = v[5];
      c.heat = v[6];
      for (int j = 0; j < 8; ++j) c.pcoef[j] = v[7 + j];
      c.lfctr = v[15];
    }
  }
}

void SimconfType::seed(unsigned int s) {
  _gen.reset(new std::mt19937(s));
  _dis_real.reset();
  _dis_int.reset();
}

// the reference sets the volume scale to l*l as well (simconf.C:74-79)
void SimconfType::setLengthScale(Real l) {
  _length_scale = l;
  _area_scale = l * l;
  _volume_scale = l * l;
}

std::vector<double> SimconfType::tableRows() const {
  std::vector<double> rows(92 * 16);
  for (unsigned i = 0; i < _rows; ++i) {
    double* v = &rows[16 * i];
    const ScoefLine& c = scoef[i];
    v[0] = c.mm1;
    v[1] = c.m1;
    v[2] = c.mnat;
    v[3] = c.rho;
    v[4] = c.atrho;
    v[5] = c.vfermi;
    v[6] = c.heat;
    for (int j = 0; j < 8; ++j) v[7 + j] = c.pcoef[j];
    v[15] = c.lfctr;
  }
  return rows;
}

// SCOEF.95A (2 header lines) and SLFCTR.dat (1 header line) token streams,
// as simconf.C:104-137 reads them
void SimconfType::readDataFiles(const std::string& dir) {
  auto open = [&](const char* name, int skip) {
    std::ifstream* in = new std::ifstream((dir + "/" + name).c_str());
    if (!*in) api_error("Error reading " + dir + "/" + name);
    std::string line;
    for (int i = 0; i < skip; ++i)
      if (!std::getline(*in, line)) api_error("Error reading file");
    return std::unique_ptr<std::ifstream>(in);
  };
  auto a = open("SCOEF.95A", 2);
  auto l = open("SLFCTR.dat", 1);
  for (unsigned i = 0; i < _rows; ++i) {
    int dummy;
    ScoefLine& c = scoef[i];
    *a >> dummy >> c.mm1 >> c.m1 >> c.mnat >> c.rho >> c.atrho >> c.vfermi >> c.heat;
    for (int j = 0; j < 8; ++j) *a >> c.pcoef[j];
    if (!*a) api_error("contents of SCOEF.95A");
    if (!(*l >> dummy >> c.lfctr)) api_error("contents of SLFCTR.dat");
  }
}

}  // namespace MyTRIM_NS
