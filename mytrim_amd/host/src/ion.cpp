// IonBase and variants (ion.C).
#include "ion.h"

namespace MyTRIM_NS {

IonBase::IonBase() : _seed(0), _tag(-1), _Ef(3.0), _state(MOVING) {}

IonBase::IonBase(IonBase* p)
    : _Z(p->_Z), _m(p->_m), _E(p->_E), _seed(0), _tag(-1), _Ef(p->_Ef), _state(MOVING) {}

IonBase::IonBase(int Z, Real m, Real E) : _Z(Z), _m(m), _E(E), _seed(0), _tag(-1), _Ef(3.0), _state(MOVING) {}

void IonBase::setEf() { _Ef = 3.0; }

void IonBase::parent(IonBase* p) {
  _gen = p->_gen + 1;
  _pos = p->_pos;
  _Ef = p->_Ef;
}

IonBase* IonBase::spawnRecoil() {
  IonBase* r = new IonBase;
  r->parent(this);
  return r;
}

bool IonBase::operator<(const IonBase& a) const { return (_Z < a._Z) || (_Z == a._Z && _m < a._m); }

std::ostream& operator<<(std::ostream& os, const IonBase& i) {
  os << i._pos(0) << ' ' << i._pos(1) << ' ' << i._pos(2) << ' ' << i._Z << ' ' << i._m << ' ' << i._E << ' '
     << i._id << ' ' << i._gen << ' ' << i._tag << ' ';
  return os;
}

IonBase* IonMDTag::spawnRecoil() {
  IonBase* r = new IonMDTag;
  r->parent(this);
  return r;
}

std::ostream& operator<<(std::ostream& os, const IonMDTag& i) {
  os << static_cast<const IonBase&>(i) << i._md << ' ';
  return os;
}

void IonClock::parent(IonBase* p) {
  IonBase::parent(p);
  IonClock* c = dynamic_cast<IonClock*>(p);
  _time = c ? c->_time : 0.0;
}

}  // namespace MyTRIM_NS
