// TrimBase::trim on the device and the hook subclasses (trim.C).
#include "trim.h"

#include <cmath>
#include <iostream>
#include <stdexcept>

#include "device.h"
#include "functions.h"
#include "mytrim_gpu.h"

namespace MyTRIM_NS {

TrimBase::TrimBase(SimconfType* simconf, SampleBase* sample)
    : _potential(UNIVERSAL),
      _simconf(simconf),
      _sample(sample),
      _pka(nullptr),
      _recoil(nullptr),
      _material(nullptr),
      _element(nullptr),
      recoil_queue_ptr(nullptr),
      terminate(false),
      _ls(0),
      _dee(0),
      _den(0),
      _base_name("mytrim"),
      _ctx(nullptr) {}

TrimBase::~TrimBase() {
  if (_ctx) mt_ctx_destroy(_ctx);
}

mt_ctx* TrimBase::device() {
  if (!_ctx) _ctx = device::create_context(_simconf);
  std::vector<double> sig = device::target_signature(_sample, _simconf, int(_potential));
  if (sig != _uploaded) {
    device::upload_target(_ctx, _sample, _simconf, int(_potential));
    _uploaded.swap(sig);
  }
  return _ctx;
}

// One flight on the device with this SimconfType's stream, then the hooks in
// trim.C order for every logged collision (trim.C:275-424).
void TrimBase::trim(IonBase* pka, std::queue<IonBase*>& recoils) {
  _pka = pka;
  _pka->_state = IonBase::MOVING;
  recoil_queue_ptr = &recoils;
  mt_ctx* ctx = device();

  mt_ion ion;
  ion.Z = pka->_Z;
  ion.m = pka->_m;
  ion.E = pka->_E;
  for (int i = 0; i < 3; ++i) {
    ion.pos[i] = pka->_pos(i);
    ion.dir[i] = pka->_dir(i);
  }
  ion.gen = pka->_gen;
  ion.tag = pka->_tag;
  ion.state = 0;
  ion.Ef = pka->_Ef;

  // draw ahead from a copy of the generator; advance the real one by what
  // the flight consumed
  long long n_u = 1024;
  std::vector<double> u;
  std::vector<mt_event> ev(256);
  long long used = 0;
  int nev = 0;
  for (;;) {
    std::mt19937 g = _simconf->generator();
    std::uniform_real_distribution<double> dis(0.0, 1.0);
    u.resize(size_t(n_u));
    for (long long i = 0; i < n_u; ++i) u[size_t(i)] = dis(g);
    mt_ion io = ion;
    const int rc = mt_trim_flight(ctx, &io, u.data(), n_u, &used, ev.data(), int(ev.size()), &nev);
    if (rc == MT_OK) {
      ion = io;
      break;
    }
    // only exhaustion of the draws or of the event buffer is retried
    if (rc != MT_E_MORE) device::check(rc, "mt_trim_flight");
    if (used >= n_u)
      n_u *= 4;
    else
      ev.resize(ev.size() * 4);
  }
  for (long long i = 0; i < used; ++i) _simconf->drand();

  for (int k = 0; k < nev; ++k) {
    const mt_event& e = ev[size_t(k)];
    _material = _sample->material[size_t(e.material)];
    if (e.mat_tag >= 0) _material->_tag = e.mat_tag;  // sampleClusters::lookupMaterial side effect
    _element = &_material->getElement(unsigned(e.element));
    _ls = e.ls;
    _dee = e.dee;
    _den = e.den;
    _simconf->EelTotal += _dee;
    // the recoil starts where the collision happened, before the ion moved
    for (int i = 0; i < 3; ++i) _pka->_pos(i) = e.rec_pos[i];
    _recoil = _pka->spawnRecoil();
    for (int i = 0; i < 3; ++i) {
      _pka->_pos(i) = e.pka_pos[i];
      _pka->_dir(i) = e.pka_dir[i];
      _recoil->_dir(i) = e.rec_dir[i];
    }
    _pka->_E = e.pka_E;
    _recoil->_E = e.rec_E;
    _recoil->_m = _element->_m;
    _recoil->_Z = _element->_Z;
    if (e.fate == 2) {
      _pka->_state = IonBase::LOST;
    } else if (e.fate == 1) {
      if (followRecoil()) {
        v_norm(_recoil->_dir);
        _recoil->_tag = _material->_tag;
        _recoil->_id = _simconf->_id++;
        recoils.push(_recoil);
        if (_simconf->fullTraj) std::cout << "spawn " << _recoil->_id << ' ' << _pka->_id << '\n';
      } else {
        _recoil->_state = IonBase::DELETE;
      }
      if (e.vacancy) {
        vacancyCreation();
      } else {
        replacementCollision();
        _pka->_state = (_pka->_Z == _element->_Z) ? IonBase::REPLACEMENT : IonBase::SUBSTITUTIONAL;
      }
    } else {
      dissipateRecoilEnergy();
      _recoil->_state = IonBase::DELETE;
      if (_pka->_E < _pka->_Ef) _pka->_state = IonBase::INTERSTITIAL;
    }
    const bool dropped = e.fate == 2;
    if (_recoil->_state == IonBase::DELETE) {
      delete _recoil;
      _recoil = nullptr;
    }
    checkPKAState();
    if (_simconf->fullTraj) std::cout << _pka->_state << ' ' << *_pka << '\n';
    if (dropped) {
      delete _recoil;
      _recoil = nullptr;
    }
  }
  if (nev == 0) {  // ended in vacuum before the first collision
    for (int i = 0; i < 3; ++i) _pka->_pos(i) = ion.pos[i];
  }
}

bool TrimBase::followRecoil() { return true; }

void TrimBase::vacancyCreation() { _simconf->vacancies_created++; }

void TrimPrimaries::vacancyCreation() {  // trim.C:445-464
  _simconf->vacancies_created++;
  if (_recoil->_gen == maxGen()) {
    Real ed = 0.0115 * std::pow(_material->_az, -7.0 / 3.0) * _recoil->_E;
    Real g = 3.4008 * std::pow(ed, 1.0 / 6.0) + 0.40244 * std::pow(ed, 3.0 / 4.0) + ed;
    Real kd = 0.1337 * std::pow(_material->_az, 2.0 / 3.0) / std::sqrt(_material->_am);
    Real Ev = _recoil->_E / (1.0 + kd * g);
    _simconf->vacancies_created += int(0.8 * Ev / (2.0 * _element->_Edisp));
  }
}

void TrimDefectLog::vacancyCreation() { _os << "V " << *_recoil << '\n'; }

void TrimDefectLog::checkPKAState() {
  if (_pka->_state == IonBase::INTERSTITIAL)
    _os << "I " << *_pka << '\n';
  else if (_pka->_state == IonBase::SUBSTITUTIONAL)
    _os << "S " << *_pka << '\n';
  else if (_pka->_state == IonBase::REPLACEMENT)
    _os << "R " << *_pka << '\n';
}

TrimVacMap::TrimVacMap(SimconfType* simconf, SampleBase* sample, int z1, int z2, int z3)
    : TrimBase(simconf, sample), _z1(z1), _z2(z2), _z3(z3) {
  for (int e = 0; e < 3; ++e)
    for (int x = 0; x < mx; ++x)
      for (int y = 0; y < my; ++y) vmap[x][y][e] = 0;
}

void TrimVacMap::vacancyCreation() {  // trim.C:483-501
  int x = int((_recoil->_pos(0) * mx) / _sample->w[0]);
  int y = int((_recoil->_pos(1) * my) / _sample->w[1]);
  x -= int(x / mx) * mx;
  y -= int(y / my) * my;
  if (x < 0 || y < 0) return;  // the reference would index out of bounds here
  if (_recoil->_Z == _z1)
    vmap[x][y][0]++;
  else if (_recoil->_Z == _z2)
    vmap[x][y][1]++;
  else if (_recoil->_Z == _z3)
    vmap[x][y][2]++;
}

void TrimPhononOut::checkPKAState() {  // trim.C:503-527
  if (_pka->_state == IonBase::MOVING || _pka->_state == IonBase::LOST) return;
  _os << _pka->_E << ' ' << *_pka << '\n';
  _simconf->EnucTotal += _pka->_E;
}

void TrimPhononOut::dissipateRecoilEnergy() {
  Real Edep = _recoil->_E + _element->_Elbind;
  _os << Edep << ' ' << *_recoil << '\n';
  _simconf->EnucTotal += Edep;
}

bool TrimPhononOut::followRecoil() {
  _os << _element->_Elbind << ' ' << *_recoil << '\n';
  _simconf->EnucTotal += _element->_Elbind;
  return true;
}

}  // namespace MyTRIM_NS
