// MaterialBase (material.C): prepare/average on the host, getrstop on the GPU.
#include "material.h"
#include "api_error.h"

#include <cmath>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "device.h"
#include "mytrim_gpu.h"

namespace MyTRIM_NS {

MaterialBase::MaterialBase(SimconfType* simconf, Real rho) : _rho(rho), _tag(-1), _dirty(true), _simconf(simconf) {}

void MaterialBase::prepare() {  // material.C:37-74
  const unsigned end = unsigned(_element.size());
  Real tt = 0.0;
  for (unsigned i = 0; i < end; ++i) {
    if (_element[i]._t < 0.0) _element[i]._t = 0.0;
    tt += _element[i]._t;
  }
  if (tt == 0.0) api_error("Stoichiometry invalid, all elements zero.");
  for (unsigned i = 0; i < end; ++i) _element[i]._t /= tt;
  _am = 0.0;
  _az = 0.0;
  for (unsigned i = 0; i < end; ++i) {
    _am += _element[i]._m * _element[i]._t;
    _az += Real(_element[i]._Z) * _element[i]._t;
  }
  if (_am == 0.0) api_error("Stoichiometry invalid, atomic density is zero.");
  _arho = _rho * 0.6022 / _am;
}

void MaterialBase::average(const IonBase* pka) {  // material.C:78-110
  mu = pka->_m / _am;
  const Real fZ = Real(pka->_Z);
  const Real fZ023 = std::pow(fZ, 0.23);
  a = .5292 * .8853 / (fZ023 + std::pow(_az, 0.23));
  f = a * _am / (_az * fZ * 14.4 * (pka->_m + _am));
  epsdg = _simconf->tmin * f * Utility::pow<2>(1.0 + mu) / (4.0 * mu);
  fd = std::pow(0.01 * _az, -7.0 / 3.0);
  kd = std::pow(0.1334 * _az, 2.0 / 3.0) / std::sqrt(_am);
  for (auto& e : _element) {
    e.my = pka->_m / e._m;
    e.ec = 4.0 * e.my / Utility::pow<2>(1.0 + e.my);
    e.ai = .5292 * .8853 / (fZ023 + std::pow(double(e._Z), 0.23));
    e.fi = e.ai * e._m / (fZ * Real(e._Z) * 14.4 * (pka->_m + e._m));
  }
  _dirty = false;
}

Real MaterialBase::getrstop(const IonBase* pka) {
  return getrstop(pka, std::vector<Real>(1, pka->_E))[0];
}

std::vector<Real> MaterialBase::getrstop(const IonBase* pka, const std::vector<Real>& energies) {
  std::vector<Real> out(energies.size());
  if (energies.empty()) return out;
  device::StoppingContext& sc = device::stopping_context(_simconf);
  std::lock_guard<std::mutex> lock(sc.mu);
  sc.upload(*this);
  if (mt_stopping_batch(sc.ctx, pka->_Z, pka->_m, 0, energies.data(), (long long)energies.size(), out.data()) != MT_OK)
    api_error(std::string("getrstop: ") + mt_last_error());
  return out;
}

}  // namespace MyTRIM_NS
