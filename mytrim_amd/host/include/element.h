// Element of the MyTRIM API (element.h).
#ifndef MYTRIM_AMD_ELEMENT_H
#define MYTRIM_AMD_ELEMENT_H

#include "point.h"

namespace MyTRIM_NS {

class Element {
 public:
  Element() : _Z(0), _m(0.0), _t(0.0), _Edisp(25.0), _Elbind(3.0), my(0), ec(0), ai(0), fi(0) {}
  int _Z;
  Real _m, _t;            // mass and relative amount
  Real _Edisp, _Elbind;   // displacement and lattice binding energy
  Real my, ec, ai, fi;    // set by MaterialBase::average
};

}  // namespace MyTRIM_NS

#endif
