// TrimBase and its hook subclasses of the MyTRIM API (trim.h).
//
// TrimBase::trim(pka, recoils) runs the flight on the GPU
// (mt_trim_flight), drawing its random numbers from the SimconfType stream in
// the reference's order, then replays the collision log through the
// protected virtual hooks in the reference's order, so subclasses see the
// same _pka / _recoil / _material / _element / _ls / _dee / _den and the same
// calls as with the CPU loop.  For many ions at once use TrimGPU (trim_gpu.h).
#ifndef MYTRIM_AMD_TRIM_H
#define MYTRIM_AMD_TRIM_H

#include <cmath>
#include <ostream>
#include <queue>
#include <string>
#include <vector>

#include "material.h"
#include "sample.h"
#include "simconf.h"

struct mt_ctx;

namespace MyTRIM_NS {

class TrimBase {
 public:
  TrimBase(SimconfType* simconf, SampleBase* sample);
  virtual ~TrimBase();

  void trim(IonBase* pka, std::queue<IonBase*>& recoils);
  void setBaseName(const std::string& name) { _base_name = name; }
  virtual void writeOutput() {}

  enum Potential { UNIVERSAL, MOLIERE, CKR };
  Potential _potential;

 protected:
  virtual bool followRecoil();
  virtual void vacancyCreation();
  virtual void replacementCollision() {}
  virtual void checkPKAState() {}
  virtual void dissipateRecoilEnergy() {}

  SimconfType* _simconf;
  SampleBase* _sample;
  IonBase *_pka, *_recoil;
  MaterialBase* _material;
  const Element* _element;
  std::queue<IonBase*>* recoil_queue_ptr;
  bool terminate;
  Real _ls;
  Real _dee;
  Real _den;
  std::string _base_name;

  // device context with the current sample uploaded
  mt_ctx* device();

 private:
  mt_ctx* _ctx;
  std::vector<double> _uploaded;  // signature of the uploaded target
};

// follow only the primary knock-ons, modified Kinchin-Pease below (trim.h:113-122)
class TrimPrimaries : public TrimBase {
 public:
  TrimPrimaries(SimconfType* simconf, SampleBase* sample) : TrimBase(simconf, sample) {}

 protected:
  virtual int maxGen() { return 1; }
  virtual bool followRecoil() { return (_recoil->_gen < maxGen()); }
  virtual void vacancyCreation();
};

// follow the first generation of recoils (trim.h:127-134)
class TrimRecoils : public TrimPrimaries {
 public:
  TrimRecoils(SimconfType* simconf, SampleBase* sample) : TrimPrimaries(simconf, sample) {}

 protected:
  virtual int maxGen() { return 2; }
};

// positions of the ion at every queued recoil (trim.h:139-154)
class TrimHistory : public TrimBase {
 public:
  TrimHistory(SimconfType* simconf, SampleBase* sample) : TrimBase(simconf, sample) {}
  const std::vector<Point>& getHistory() { return _pos_hist; }

 protected:
  virtual bool followRecoil() {
    _pos_hist.push_back(_pka->_pos);
    return true;
  }
  std::vector<Point> _pos_hist;
};

// vacancy / interstitial log (trim.h:159-175)
class TrimDefectLog : public TrimBase {
 public:
  TrimDefectLog(SimconfType* simconf, SampleBase* sample, std::ostream& os) : TrimBase(simconf, sample), _os(os) {}

 protected:
  std::ostream& _os;
  virtual void vacancyCreation();
  virtual void checkPKAState();
};

// vacancy map per species (trim.h:180-201)
class TrimVacMap : public TrimBase {
  static const int mx = 20, my = 20;

 public:
  TrimVacMap(SimconfType* simconf, SampleBase* sample, int z1, int z2, int z3 = -1);
  int vmap[mx][my][3];

 protected:
  int _z1, _z2, _z3;
  virtual void vacancyCreation();
};

// phonon energy log (trim.h:206-226)
class TrimPhononOut : public TrimBase {
 public:
  TrimPhononOut(SimconfType* simconf, SampleBase* sample, std::ostream& os) : TrimBase(simconf, sample), _os(os) {}

 protected:
  std::ostream& _os;
  virtual void checkPKAState();
  virtual void dissipateRecoilEnergy();
  virtual bool followRecoil();
};

}  // namespace MyTRIM_NS

#endif
