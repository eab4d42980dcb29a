// IonBase and its recoil/tag variants of the MyTRIM API (ion.h).
#ifndef MYTRIM_AMD_ION_H
#define MYTRIM_AMD_ION_H

#include <iostream>

#include "point.h"

namespace MyTRIM_NS {

class IonBase {
 public:
  IonBase();
  IonBase(IonBase* prototype);
  IonBase(int Z, Real m, Real E);
  virtual ~IonBase() {}

  virtual void parent(IonBase* parent);
  virtual IonBase* spawnRecoil();
  void setEf();
  bool operator<(const IonBase&) const;

  int _Z;
  Real _m;
  Real _E;
  Point _dir, _pos;
  unsigned int _seed;
  int _gen, _id;
  int _tag;
  Real _Ef;
  enum StateType { MOVING, REPLACEMENT, SUBSTITUTIONAL, INTERSTITIAL, LOST, DELETE, VACANCY } _state;
};

std::ostream& operator<<(std::ostream& os, const IonBase& i);

class IonMDTag : public IonBase {
 public:
  IonMDTag() : IonBase(), _md(0) {}
  IonMDTag(IonMDTag* prototype) : IonBase(prototype), _md(prototype->_md) {}
  virtual IonBase* spawnRecoil();
  int _md;
};

std::ostream& operator<<(std::ostream& os, const IonMDTag& p);

class IonClock : public IonBase {
 public:
  IonClock() : IonBase(), _time(0.0) {}
  IonClock(IonClock* prototype) : IonBase(prototype), _time(prototype->_time) {}
  void parent(IonBase* parent);
  Real _time;
};

}  // namespace MyTRIM_NS

#endif
