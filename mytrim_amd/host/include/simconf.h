// SimconfType of the MyTRIM API (simconf.h): tables, knobs, the host random
// stream and the run counters.  The tables come from $MYTRIM_DATADIR
// (SCOEF.95A / SLFCTR.dat in the reference's format) when it is set, else
// from the tables built into libmytrim_gpu.
#ifndef MYTRIM_AMD_SIMCONF_H
#define MYTRIM_AMD_SIMCONF_H

#include <memory>
#include <random>
#include <string>
#include <vector>

#include "point.h"

namespace MyTRIM_NS {

class SimconfType {
 public:
  SimconfType(unsigned int seed = 12345678);

  Real drand() { return _dis_real(*_gen); }
  unsigned int irand() { return _dis_int(*_gen); }
  void seed(unsigned int seed);

  // length scale in Angstrom (1: A, 10: nm, 10000: um)
  void setLengthScale(Real l);
  Real lengthScale() { return _length_scale; }
  Real areaScale() { return _area_scale; }
  Real volumeScale() { return _volume_scale; }

  Real ed, tmin, tau, da, cw;
  int _id;

  struct ScoefLine {
    std::string sym, name;
    Real mm1, m1, mnat, rho, atrho, vfermi, heat, lfctr;
    std::vector<Real> pcoef;  // 8
    ScoefLine() : mm1(0), m1(0), mnat(0), rho(0), atrho(0), vfermi(0), heat(0), lfctr(0), pcoef(8) {}
  };
  static const unsigned int _rows = 92;
  std::vector<ScoefLine> scoef;

  bool fullTraj;
  Real EelTotal;
  Real EnucTotal;
  int vacancies_created;

  // the generator, for callers that must draw a known number of values ahead
  std::mt19937& generator() { return *_gen; }
  // 92 x 17 rows (mm1 m1 mnat rho atrho vfermi heat pcoef[8] lfctr)
  std::vector<double> tableRows() const;

 private:
  void readDataFiles(const std::string& dir);
  std::unique_ptr<std::mt19937> _gen;
  std::uniform_real_distribution<double> _dis_real;
  std::uniform_int_distribution<unsigned int> _dis_int;
  Real _length_scale, _area_scale, _volume_scale;
};

}  // namespace MyTRIM_NS

#endif
