// ORACLE — test infrastructure only.
//
// Fixture writer linked against the reference library compiled from its own
// sources (oracle/Makefile target `ref`).  It only drives the reference's
// public/protected API (TrimBase hooks, MaterialBase, samples); every number
// it prints is computed by the reference.  Runs only where the reference is
// mounted; its output is committed under tests/golden/.
//
//   ref_harness stop Z1 m1 rho nel Z m t ...   (energies on stdin)  -> getrstop
//   ref_harness avg  Z1 m1 tmin rho nel Z m t ...                   -> average()
//   ref_harness trace SPECFILE h max_rows                           -> per collision
//   ref_harness tally SPECFILE                                      -> tallies
//   ref_harness log SPECFILE defects|phonons|history|vacmap [z1 z2 z3]
//       -> the stream of TrimDefectLog / TrimPhononOut (17 digits), the
//          positions TrimHistory stores, or TrimVacMap's map; "H h" before
//          each history
//   ref_harness dynamic  -> SampleDynamic::addAtomsToLayer on a fixed sequence
//          (layer thicknesses and compositions after each step)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <queue>
#include <sstream>
#include <string>
#include <vector>

#include <iomanip>

#include "element.h"
#include "ion.h"
#include "material.h"
#include "sample_clusters.h"
#include "sample_layers.h"
#include "sample_burried_wire.h"
#include "sample_wire.h"
#include "sample_dynamic.h"
#include "sample_solid.h"
#include "simconf.h"
#include "trim.h"

#include "src/philox_ref.h"

using namespace MyTRIM_NS;

namespace {

struct MatSpec {
  double rho;
  std::vector<int> Z;
  std::vector<double> m, t, Ed, Eb;
};

struct Spec {
  std::string geo = "solid";
  double w[3] = {10000, 10000, 10000};
  int bc[3] = {0, 0, 0};
  std::vector<MatSpec> mats;
  std::vector<double> layers;
  int kn[3] = {1, 1, 1};
  std::vector<double> cl;  // x y z r
  double tmin = 0.2, tau = 0.0, cw = 0.001, scale = 1.0, Ef = 3.0;
  int potential = 0;
  int pZ = 1;
  double pm = 1, pE = 1, ppos[3] = {0, 0, 0}, pdir[3] = {1, 0, 0};
  // per-history primaries (pkafile): Z m E x y z dx dy dz, one row per
  // global history id, from history 0
  std::vector<double> prow;
  long long nhist = 1, hist0 = 0;
  unsigned seed = 1;
  int follow = -1, kp = 0;
  int perhist = 0;  // tally: also print per-history arrays
  double range_x0 = 0, range_bw = 10, vac_x0 = 0, vac_bw = 10, lat_y0 = 0, lat_z0 = 0, lat_bw = 10;
  int range_nb = 100, vac_nb = 100, lat_nb = 100;
};

Spec read_spec(const char* path) {
  Spec s;
  std::ifstream in(path);
  if (!in) {
    std::cerr << "cannot open " << path << "\n";
    exit(2);
  }
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string k;
    if (!(ss >> k) || k[0] == '#') continue;
    if (k == "geo") ss >> s.geo;
    else if (k == "w") ss >> s.w[0] >> s.w[1] >> s.w[2];
    else if (k == "bc") ss >> s.bc[0] >> s.bc[1] >> s.bc[2];
    else if (k == "mat") {
      MatSpec m;
      int n;
      ss >> m.rho >> n;
      for (int i = 0; i < n; ++i) {
        int z;
        double mm, t, ed, eb;
        ss >> z >> mm >> t >> ed >> eb;
        m.Z.push_back(z);
        m.m.push_back(mm);
        m.t.push_back(t);
        m.Ed.push_back(ed);
        m.Eb.push_back(eb);
      }
      s.mats.push_back(m);
    } else if (k == "layers") {
      double t;
      while (ss >> t) s.layers.push_back(t);
    } else if (k == "kn") ss >> s.kn[0] >> s.kn[1] >> s.kn[2];
    else if (k == "cluster") {
      double v;
      for (int i = 0; i < 4; ++i) {
        ss >> v;
        s.cl.push_back(v);
      }
    } else if (k == "tmin") ss >> s.tmin;
    else if (k == "tau") ss >> s.tau;
    else if (k == "cw") ss >> s.cw;
    else if (k == "scale") ss >> s.scale;
    else if (k == "Ef") ss >> s.Ef;
    else if (k == "potential") ss >> s.potential;
    else if (k == "pka")
      ss >> s.pZ >> s.pm >> s.pE >> s.ppos[0] >> s.ppos[1] >> s.ppos[2] >> s.pdir[0] >> s.pdir[1] >>
          s.pdir[2];
    else if (k == "pkafile") {
      std::string path;
      ss >> path;
      std::ifstream pf(path.c_str());
      if (!pf) {
        std::cerr << "cannot open " << path << "\n";
        exit(2);
      }
      double v;
      while (pf >> v) s.prow.push_back(v);
      if (s.prow.size() % 9) {
        std::cerr << "pkafile: rows of 9 numbers expected\n";
        exit(2);
      }
    } else if (k == "nhist") ss >> s.nhist;
    else if (k == "hist0") ss >> s.hist0;
    else if (k == "seed") ss >> s.seed;
    else if (k == "follow") ss >> s.follow;
    else if (k == "kp") ss >> s.kp;
    else if (k == "perhist") ss >> s.perhist;
    else if (k == "range") ss >> s.range_x0 >> s.range_bw >> s.range_nb;
    else if (k == "vac") ss >> s.vac_x0 >> s.vac_bw >> s.vac_nb;
    else if (k == "lat") ss >> s.lat_y0 >> s.lat_z0 >> s.lat_bw >> s.lat_nb;
  }
  return s;
}

SampleBase* build(const Spec& s, SimconfType* sc) {
  SampleBase* smp;
  if (s.geo == "solid") {
    smp = new SampleSolid(s.w[0], s.w[1], s.w[2]);
  } else if (s.geo == "layers") {
    SampleLayers* sl = new SampleLayers(s.w[0], s.w[1], s.w[2]);
    for (double t : s.layers) sl->layerThickness.push_back(t);
    smp = sl;
  } else if (s.geo == "wire") {
    smp = new SampleWire(s.w[0], s.w[1], s.w[2]);
  } else if (s.geo == "burriedwire") {
    smp = new SampleBurriedWire(s.w[0], s.w[1], s.w[2]);
  } else {
    sampleClusters* c = new sampleClusters(s.w[0], s.w[1], s.w[2]);
    for (int i = 0; i < 3; ++i) c->bc[i] = SampleBase::sampleBoundary(s.bc[i]);
    c->initSpatialhash(s.kn[0], s.kn[1], s.kn[2]);
    for (size_t i = 0; i + 3 < s.cl.size(); i += 4) c->addCluster(s.cl[i], s.cl[i + 1], s.cl[i + 2], s.cl[i + 3]);
    smp = c;
  }
  for (int i = 0; i < 3; ++i) smp->bc[i] = SampleBase::sampleBoundary(s.bc[i]);
  for (const auto& ms : s.mats) {
    MaterialBase* m = new MaterialBase(sc, ms.rho);
    for (size_t i = 0; i < ms.Z.size(); ++i) {
      Element e;
      e._Z = ms.Z[i];
      e._m = ms.m[i];
      e._t = ms.t[i];
      e._Edisp = ms.Ed[i];
      e._Elbind = ms.Eb[i];
      m->_element.push_back(e);
    }
    m->prepare();
    smp->material.push_back(m);
  }
  return smp;
}

int bin_of(double x, double x0, double bw, int nb) {
  const double t = (x - x0) / bw;
  if (!(t >= 0.0)) return 0;
  if (t >= double(nb)) return nb + 1;
  return 1 + int(t);
}

struct Tallies {
  long long n_coll = 0, n_flight = 0, vac = 0, repl = 0, subst = 0, inter = 0, lost = 0, vacuum = 0,
            disp = 0;
  double E0 = 0, Eel = 0, Enuc = 0, Eunf = 0, Elost = 0, Edrop = 0, Evac = 0;
  long long range_n = 0;
  double range_sum = 0, range_sum2 = 0;
  std::vector<long long> range_hist, lat_y, lat_z, vac_hist, repl_hist;
};

// Tally definitions identical to oracle/src/runs.cpp (TallyHooks)
class TallyTrim : public TrimBase {
 public:
  TallyTrim(SimconfType* sc, SampleBase* smp, const Spec& s, Tallies& t)
      : TrimBase(sc, smp), _s(s), _t(t) {}

 protected:
  bool followRecoil() override {
    _t.disp++;
    const bool f = _s.follow < 0 || _recoil->_gen < _s.follow;
    if (f)
      _t.Enuc += _element->_Elbind;
    else
      _t.Eunf += _den;
    return f;
  }
  void vacancyCreation() override {
    _t.vac++;
    _t.vac_hist[bin_of(_recoil->_pos(0), _s.vac_x0, _s.vac_bw, _s.vac_nb)]++;
    if (_s.kp && _recoil->_gen == _s.follow) {
      const Real ed = 0.0115 * std::pow(_material->_az, -7.0 / 3.0) * _recoil->_E;
      const Real g = 3.4008 * std::pow(ed, 1.0 / 6.0) + 0.40244 * std::pow(ed, 3.0 / 4.0) + ed;
      const Real kd = 0.1337 * std::pow(_material->_az, 2.0 / 3.0) / std::sqrt(_material->_am);
      const Real Ev = _recoil->_E / (1.0 + kd * g);
      _t.vac += int(0.8 * Ev / (2.0 * _element->_Edisp));
    }
  }
  void replacementCollision() override {
    _t.repl_hist[bin_of(_recoil->_pos(0), _s.vac_x0, _s.vac_bw, _s.vac_nb)]++;
  }
  void dissipateRecoilEnergy() override { _t.Enuc += _recoil->_E + _element->_Elbind; }
  void checkPKAState() override {
    _t.n_coll++;
    _t.Eel += _dee;
    switch (_pka->_state) {
      case IonBase::LOST:
        _t.lost++;
        _t.Elost += _pka->_E;
        _t.Edrop += _den;
        break;
      case IonBase::REPLACEMENT:
        _t.repl++;
        _t.Enuc += _pka->_E;
        break;
      case IonBase::SUBSTITUTIONAL:
        _t.subst++;
        _t.Enuc += _pka->_E;
        break;
      case IonBase::INTERSTITIAL:
        _t.inter++;
        _t.Enuc += _pka->_E;
        break;
      default:
        break;
    }
  }
  const Spec& _s;
  Tallies& _t;
};

// Per-collision trace in the row layout of oracle/src/runs.cpp (orc_trace)
class TraceTrim : public TrimBase {
 public:
  TraceTrim(SimconfType* sc, SampleBase* smp) : TrimBase(sc, smp) {}
  std::vector<double> rows;
  int flight = 0, ic = 0;

 protected:
  double _rE = 0, _rdir[3] = {0, 0, 0};
  void grab() {
    _rE = _recoil->_E;
    for (int i = 0; i < 3; ++i) _rdir[i] = _recoil->_dir(i);
  }
  bool followRecoil() override {
    grab();
    return true;
  }
  void dissipateRecoilEnergy() override { grab(); }
  void checkPKAState() override {
    ++ic;
    if (_pka->_state == IonBase::LOST) grab();
    const double r[21] = {double(flight), double(ic), double(_pka->_Z), _pka->_E, _pka->_pos(0),
                          _pka->_pos(1), _pka->_pos(2), _pka->_dir(0), _pka->_dir(1), _pka->_dir(2),
                          double(_pka->_state), _ls, _dee, _den, double(_element->_Z), _rE,
                          _rdir[0], _rdir[1], _rdir[2], double(_pka->_gen), 0.0};
    rows.insert(rows.end(), r, r + 21);
  }
};

// the primary ion of global history h: the spec's single species, or row h
// of its pkafile
IonBase* primary(const Spec& s, long long h) {
  IonBase* p = new IonBase;
  if (!s.prow.empty()) {
    if ((h + 1) * 9 > (long long)s.prow.size()) {
      std::cerr << "pkafile holds no row for history " << h << "\n";
      exit(2);
    }
    const double* r = &s.prow[h * 9];
    p->_Z = int(r[0]);
    p->_m = r[1];
    p->_E = r[2];
    p->_pos = Point(r[3], r[4], r[5]);
    p->_dir = Point(r[6], r[7], r[8]);
  } else {
    p->_Z = s.pZ;
    p->_m = s.pm;
    p->_E = s.pE;
    p->_pos = Point(s.ppos[0], s.ppos[1], s.ppos[2]);
    p->_dir = Point(s.pdir[0], s.pdir[1], s.pdir[2]);
  }
  p->_gen = 0;
  p->_Ef = s.Ef;
  return p;
}

unsigned history_seed(unsigned master, long long h) {
  return unsigned(orc::flight_key_child(master, unsigned(h)));
}

void print_vec(const char* name, const std::vector<long long>& v) {
  printf("%s", name);
  for (long long x : v) printf(" %lld", x);
  printf("\n");
}

void print_dvec(const char* name, const std::vector<double>& v) {
  printf("%s", name);
  for (double x : v) printf(" %.17g", x);
  printf("\n");
}

int cmd_tally(const Spec& s) {
  SimconfType sc;
  sc.tmin = s.tmin;
  sc.tau = s.tau;
  sc.cw = s.cw;
  sc.setLengthScale(s.scale);
  SampleBase* smp = build(s, &sc);
  Tallies t;
  t.range_hist.assign(s.range_nb + 2, 0);
  t.lat_y.assign(s.lat_nb + 2, 0);
  t.lat_z.assign(s.lat_nb + 2, 0);
  t.vac_hist.assign(s.vac_nb + 2, 0);
  t.repl_hist.assign(s.vac_nb + 2, 0);
  TallyTrim trim(&sc, smp, s, t);
  trim._potential = TrimBase::Potential(s.potential);
  std::vector<long long> hvac, hcoll;
  // per history: Eel, Enuc, E0, and the primary's end (x y z state; state -1
  // for a vacuum exit, as the batched kernels report it)
  std::vector<double> heel, henuc, he0, hend;
  for (long long h = s.hist0; h < s.hist0 + s.nhist; ++h) {
    const long long vac0 = t.vac, coll0 = t.n_coll;
    const double eel0 = t.Eel, enuc0 = t.Enuc;
    double end[4] = {0, 0, 0, 0};
    sc.seed(history_seed(s.seed, h));
    std::queue<IonBase*> q;
    IonBase* p = primary(s, h);
    t.E0 += p->_E;
    he0.push_back(p->_E);
    q.push(p);
    while (!q.empty()) {
      IonBase* ion = q.front();
      q.pop();
      smp->averages(ion);
      trim.trim(ion, q);
      t.n_flight++;
      const bool vacuum = ion->_state == IonBase::MOVING;
      if (vacuum) {
        t.vacuum++;
        t.Evac += ion->_E;
      }
      if (ion->_gen == 0 && !vacuum && ion->_state != IonBase::LOST) {
        t.range_n++;
        t.range_sum += ion->_pos(0);
        t.range_sum2 += ion->_pos(0) * ion->_pos(0);
        t.range_hist[bin_of(ion->_pos(0), s.range_x0, s.range_bw, s.range_nb)]++;
        t.lat_y[bin_of(ion->_pos(1) - s.lat_y0, -0.5 * s.lat_nb * s.lat_bw, s.lat_bw, s.lat_nb)]++;
        t.lat_z[bin_of(ion->_pos(2) - s.lat_z0, -0.5 * s.lat_nb * s.lat_bw, s.lat_bw, s.lat_nb)]++;
      }
      if (ion->_gen == 0) {
        for (int i = 0; i < 3; ++i) end[i] = ion->_pos(i);
        end[3] = vacuum ? -1.0 : double(ion->_state);
      }
      delete ion;
    }
    hvac.push_back(t.vac - vac0);
    hcoll.push_back(t.n_coll - coll0);
    heel.push_back(t.Eel - eel0);
    henuc.push_back(t.Enuc - enuc0);
    hend.insert(hend.end(), end, end + 4);
  }
  printf("n_coll %lld\nn_flight %lld\nvac %lld\nrepl %lld\nsubst %lld\ninter %lld\nlost %lld\n",
         t.n_coll, t.n_flight, t.vac, t.repl, t.subst, t.inter, t.lost);
  printf("vacuum %lld\ndisp %lld\n", t.vacuum, t.disp);
  printf("E0 %.17g\nEel %.17g\nEnuc %.17g\nEunf %.17g\nElost %.17g\nEdrop %.17g\nEvac %.17g\n", t.E0,
         t.Eel, t.Enuc, t.Eunf, t.Elost, t.Edrop, t.Evac);
  printf("range_n %lld\nrange_sum %.17g\nrange_sum2 %.17g\n", t.range_n, t.range_sum, t.range_sum2);
  print_vec("range_hist", t.range_hist);
  print_vec("lat_y_hist", t.lat_y);
  print_vec("lat_z_hist", t.lat_z);
  print_vec("vac_hist", t.vac_hist);
  print_vec("repl_hist", t.repl_hist);
  print_vec("h_vac", hvac);
  if (s.perhist) {
    print_vec("h_coll", hcoll);
    print_dvec("h_Eel", heel);
    print_dvec("h_Enuc", henuc);
    print_dvec("h_E0", he0);
    print_dvec("h_end", hend);
  }
  return 0;
}

int cmd_trace(const Spec& s, long long h, int max_rows) {
  SimconfType sc;
  sc.tmin = s.tmin;
  sc.tau = s.tau;
  sc.cw = s.cw;
  sc.setLengthScale(s.scale);
  SampleBase* smp = build(s, &sc);
  TraceTrim trim(&sc, smp);
  trim._potential = TrimBase::Potential(s.potential);
  sc.seed(history_seed(s.seed, h));
  std::queue<IonBase*> q;
  q.push(primary(s, h));
  while (!q.empty() && int(trim.rows.size() / 21) < max_rows) {
    IonBase* ion = q.front();
    q.pop();
    smp->averages(ion);
    trim.trim(ion, q);
    trim.flight++;
    trim.ic = 0;
    delete ion;
  }
  const int n = std::min<int>(max_rows, int(trim.rows.size() / 21));
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < 21; ++j) printf(j ? " %.17g" : "%.17g", trim.rows[21 * i + j]);
    printf("\n");
  }
  return 0;
}

}  // namespace

// SampleDynamic::addAtomsToLayer (sample_dynamic.C:64-119) on a fixed sequence
// of additions and removals; the same sequence runs through the C++ API in
// tests/native/dynamic_check.cpp
int cmd_dynamic() {
  SimconfType sc;
  SampleDynamic smp(&sc, 1000.0, 100.0, 100.0);
  MaterialBase* si = new MaterialBase(&sc, 2.33);
  Element e;
  e._Z = 14;
  e._m = 28.086;
  e._t = 1.0;
  si->_element.push_back(e);
  si->prepare();
  MaterialBase* ox = new MaterialBase(&sc, 2.2);
  ox->_element.push_back(e);
  e._Z = 8;
  e._m = 15.999;
  e._t = 2.0;
  ox->_element.push_back(e);
  ox->prepare();
  smp.material.push_back(si);
  smp.material.push_back(ox);
  smp.layerThickness.push_back(500.0);
  smp.layerThickness.push_back(10000.0);
  const int steps[][3] = {{0, 1000, 5}, {0, -10, 14}, {1, 500, 5}, {0, 2000, 32}, {1, -100000000, 8}, {1, 7, 8}};
  for (const auto& st : steps) {
    smp.addAtomsToLayer(st[0], st[1], st[2]);
    printf("T %.17g %.17g\n", smp.layerThickness[0], smp.layerThickness[1]);
    for (MaterialBase* m : smp.material) {
      printf("M %.17g %.17g %.17g %.17g", m->_rho, m->_arho, m->_am, m->_az);
      for (const Element& x : m->_element) printf(" %d %.17g %.17g", x._Z, x._m, x._t);
      printf("\n");
    }
  }
  return 0;
}

// the stream-writing and recording TrimBase subclasses of trim.h:139-227
int cmd_log(const Spec& s, const std::string& mode, int z1, int z2, int z3) {
  SimconfType sc;
  sc.tmin = s.tmin;
  sc.tau = s.tau;
  sc.cw = s.cw;
  sc.setLengthScale(s.scale);
  SampleBase* smp = build(s, &sc);
  std::ostringstream os;
  os << std::setprecision(17);
  TrimBase* trim;
  TrimHistory* th = nullptr;
  TrimVacMap* tv = nullptr;
  if (mode == "defects") {
    trim = new TrimDefectLog(&sc, smp, os);
  } else if (mode == "phonons") {
    trim = new TrimPhononOut(&sc, smp, os);
  } else if (mode == "history") {
    trim = th = new TrimHistory(&sc, smp);
  } else if (mode == "vacmap") {
    trim = tv = new TrimVacMap(&sc, smp, z1, z2, z3);
  } else if (mode == "transport") {
    trim = new TrimBase(&sc, smp);
  } else {
    fprintf(stderr, "unknown log mode %s\n", mode.c_str());
    return 2;
  }
  trim->_potential = TrimBase::Potential(s.potential);
  // transport: the layer bookkeeping of apps/mytrim_dynamic.C:160-178 with the
  // composition held fixed (per layer and Z: -1 where a recoil starts, +1
  // where a flight ending at x > 0 stops)
  SampleLayers* layers = dynamic_cast<SampleLayers*>(smp);
  std::map<std::pair<int, int>, long long> moves;
  const bool transport = mode == "transport";
  if (transport && !layers) {
    fprintf(stderr, "transport needs a layers sample\n");
    return 2;
  }
  size_t nh = 0;
  for (long long h = s.hist0; h < s.hist0 + s.nhist; ++h) {
    sc.seed(history_seed(s.seed, h));
    std::queue<IonBase*> q;
    IonBase* p = primary(s, h);
    p->_id = 0;
    q.push(p);
    while (!q.empty()) {
      IonBase* ion = q.front();
      q.pop();
      smp->averages(ion);
      if (transport && ion->_gen > 0) moves[{layers->lookupLayer(ion->_pos), ion->_Z}] -= 1;
      trim->trim(ion, q);
      if (transport && ion->_pos(0) > 0) moves[{layers->lookupLayer(ion->_pos), ion->_Z}] += 1;
      delete ion;
    }
    printf("H %lld\n", h);
    fputs(os.str().c_str(), stdout);
    os.str("");
    if (th) {
      const std::vector<Point>& v = th->getHistory();
      for (; nh < v.size(); ++nh) printf("%.17g %.17g %.17g\n", v[nh](0), v[nh](1), v[nh](2));
    }
  }
  for (const auto& m : moves) printf("T %d %d %lld\n", m.first.first, m.first.second, m.second);
  if (tv) {
    printf("M");
    for (int x = 0; x < 20; ++x)
      for (int y = 0; y < 20; ++y)
        for (int e = 0; e < 3; ++e) printf(" %d", tv->vmap[x][y][e]);
    printf("\n");
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: ref_harness stop|avg|trace|tally ...\n");
    return 2;
  }
  const std::string cmd = argv[1];
  if (cmd == "stop" || cmd == "avg") {
    SimconfType sc;
    int a = 2;
    const int Z1 = atoi(argv[a++]);
    const double m1 = atof(argv[a++]);
    double tmin = 0.2;
    if (cmd == "avg") tmin = atof(argv[a++]);
    const double rho = atof(argv[a++]);
    const int nel = atoi(argv[a++]);
    MaterialBase mat(&sc, rho);
    for (int i = 0; i < nel; ++i) {
      Element e;
      e._Z = atoi(argv[a++]);
      e._m = atof(argv[a++]);
      e._t = atof(argv[a++]);
      mat._element.push_back(e);
    }
    mat.prepare();
    IonBase ion(Z1, m1, 0.0);
    if (cmd == "stop") {
      double E;
      while (std::cin >> E) {
        ion._E = E;
        printf("%.17g\n", mat.getrstop(&ion));
      }
    } else {
      sc.tmin = tmin;
      mat.average(&ion);
      printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g", mat._am, mat._az, mat._arho,
             mat.mu, mat.a, mat.f, mat.epsdg, mat.fd, mat.kd);
      for (const auto& e : mat._element) printf(" %.17g %.17g %.17g %.17g", e.my, e.ec, e.ai, e.fi);
      printf("\n");
    }
    return 0;
  }
  if (cmd == "trace") return cmd_trace(read_spec(argv[2]), atoll(argv[3]), atoi(argv[4]));
  if (cmd == "tally") return cmd_tally(read_spec(argv[2]));
  if (cmd == "dynamic") return cmd_dynamic();
  if (cmd == "log")
    return cmd_log(read_spec(argv[2]), argv[3], argc > 4 ? atoi(argv[4]) : -1, argc > 5 ? atoi(argv[5]) : -1,
                   argc > 6 ? atoi(argv[6]) : -1);
  fprintf(stderr, "unknown command %s\n", argv[1]);
  return 2;
}
