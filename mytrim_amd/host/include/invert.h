// Fission-fragment samplers of the MyTRIM API (invert.h): inverse CDFs by
// bisection (H. R. Faust, Eur. Phys. J. A 14 (2002) 459).
#ifndef MYTRIM_AMD_INVERT_H
#define MYTRIM_AMD_INVERT_H

#include "point.h"

namespace MyTRIM_NS {

class Inverter {
 public:
  Inverter() : maxx(0.0), maxf(0.0), tol(1e-13) {}
  virtual ~Inverter() {}
  Real x(Real f) const;

 protected:
  virtual Real f(Real x) const = 0;
  Real maxx, maxf, tol;
};

class MassInverter : public Inverter {
 public:
  MassInverter();

 protected:
  virtual Real f(Real x) const;
};

class EnergyInverter : public Inverter {
 public:
  EnergyInverter();
  void setMass(Real A);

 protected:
  virtual Real f(Real x) const;

 private:
  Real _A;
};

}  // namespace MyTRIM_NS

#endif
