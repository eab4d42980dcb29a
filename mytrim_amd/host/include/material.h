// MaterialBase of the MyTRIM API (material.h).  getrstop runs on the GPU
// (mt_stopping_batch); prepare() and average() are the host set-up.
#ifndef MYTRIM_AMD_MATERIAL_H
#define MYTRIM_AMD_MATERIAL_H

#include <vector>

#include "element.h"
#include "ion.h"
#include "simconf.h"

namespace MyTRIM_NS {

class MaterialBase {
 public:
  MaterialBase(SimconfType* simconf, Real rho);
  virtual ~MaterialBase() {}

  void prepare();
  void average(const IonBase* pka);
  Real getrstop(const IonBase* pka);
  // getrstop for one ion at many energies, one device batch
  std::vector<Real> getrstop(const IonBase* pka, const std::vector<Real>& energies);

  virtual const Element& getElement(unsigned int nn) { return _element[nn]; }

  Real _rho;
  Real _am, _az;
  Real _arho;
  Real mu;
  Real a, f, epsdg;
  Real fd, kd;
  Real pmax;
  int _tag;
  bool _dirty;
  std::vector<Element> _element;

 protected:
  SimconfType* _simconf;
};

}  // namespace MyTRIM_NS

#endif
