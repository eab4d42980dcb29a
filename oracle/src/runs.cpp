// ORACLE — test infrastructure only.  Cascade drivers with tallies and the
// C interface used from Python (tests, bench.py cpu_baseline).
//
// History loop: apps/runmytrim.C:71-93 (FIFO drain, averages() before each
// trim()).  Tally definitions (shared with the HIP path):
//   vacancies / replacements   TrimBase::vacancyCreation (trim.C:439-443),
//                              TrimVacCount depth bins (apps/src/TrimVacCount.C:31-53)
//   energy ledger              TrimPhononOut (trim.C:503-527) plus the energy
//                              that leaves the ledger: LOST ions and the recoil
//                              dropped on that step (trim.C:344-357), flights
//                              ending in vacuum (trim.C:80-82) and recoils that
//                              are not followed
//   range / lateral            final position of the gen-0 ion after trim()
//                              (apps/mytrim_layers.C:171-186)
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <thread>

#include "counted.h"
#include "oracle.h"
#include "oracle_capi.h"

namespace orc {
namespace {

std::unique_ptr<Sample> build_sample(const orc_run_spec& s, std::vector<std::unique_ptr<Material>>& mats) {
  mats.clear();
  for (int k = 0; k < s.n_mat; ++k) {
    std::unique_ptr<Material> m(new Material);
    m->rho = s.mat_rho[k];
    for (int e = s.mat_el_off[k]; e < s.mat_el_off[k + 1]; ++e) {
      Element el;
      el.Z = s.el_Z[e];
      el.m = s.el_m[e];
      el.t = s.el_t[e];
      el.Edisp = s.el_Edisp[e];
      el.Elbind = s.el_Elbind[e];
      m->el.push_back(el);
    }
    m->prepare();
    mats.push_back(std::move(m));
  }
  std::unique_ptr<Sample> smp;
  if (s.geo == 0) {
    smp.reset(new SampleSolid(s.w[0], s.w[1], s.w[2]));
  } else if (s.geo == 1) {
    SampleLayers* sl = new SampleLayers(s.w[0], s.w[1], s.w[2]);
    for (int i = 0; i < s.n_layers; ++i) sl->layerThickness.push_back(s.layer_thick[i]);
    smp.reset(sl);
  } else if (s.geo == 3) {
    smp.reset(new SampleWire(s.w[0], s.w[1], s.w[2]));
  } else if (s.geo == 4) {
    smp.reset(new SampleBurriedWire(s.w[0], s.w[1], s.w[2]));
  } else {
    SampleClusters* sc = new SampleClusters(s.w[0], s.w[1], s.w[2]);
    for (int i = 0; i < 3; ++i) sc->bc[i] = s.bc[i];
    sc->initSpatialhash(s.kn[0], s.kn[1], s.kn[2]);
    for (int i = 0; i < s.n_cl; ++i) sc->addCluster(s.cl_x[i], s.cl_y[i], s.cl_z[i], s.cl_r[i]);
    smp.reset(sc);
  }
  for (int i = 0; i < 3; ++i) smp->bc[i] = s.bc[i];
  for (auto& m : mats) smp->material.push_back(m.get());
  return smp;
}

inline int bin_of(double x, double x0, double bw, int nb) {
  const double t = (x - x0) / bw;
  if (!(t >= 0.0)) return 0;
  if (t >= double(nb)) return nb + 1;
  return 1 + int(t);
}

struct Partial {
  long long n_coll = 0, n_flight = 0, vac = 0, repl = 0, subst = 0, inter = 0, lost = 0,
            vacuum = 0, rutherford = 0, newton = 0, disp = 0;
  double E0 = 0, Eel = 0, Enuc = 0, Eunf = 0, Elost = 0, Edrop = 0, Evac = 0;
  long long range_n = 0;
  double range_sum = 0, range_sum2 = 0;
  std::vector<long long> range_hist, lat_y, lat_z, vac_hist, repl_hist;
};

struct TallyHooks : Hooks {
  const orc_run_spec& s;
  Partial& p;
  const orc_run_out* out;
  long long hidx = 0;  // history index - hist0
  TallyHooks(const orc_run_spec& spec, Partial& part, const orc_run_out* o) : s(spec), p(part), out(o) {}

  bool followRecoil() override {
    p.disp++;
    const bool f = s.follow_max_gen < 0 || recoil->gen < s.follow_max_gen;
    if (f)
      p.Enuc += element->Elbind;
    else
      p.Eunf += den;
    return f;
  }
  void vacancyCreation() override {  // trim.C:439-443, 445-464
    p.vac++;
    if (out->h_vac) out->h_vac[hidx]++;
    p.vac_hist[bin_of(recoil->pos[0], s.vac_x0, s.vac_bw, s.vac_nb)]++;
    if (s.kp && recoil->gen == s.follow_max_gen) {
      const double ed = 0.0115 * std::pow(material->az, -7.0 / 3.0) * recoil->E;
      const double g = 3.4008 * std::pow(ed, 1.0 / 6.0) + 0.40244 * std::pow(ed, 3.0 / 4.0) + ed;
      const double kd = 0.1337 * std::pow(material->az, 2.0 / 3.0) / std::sqrt(material->am);
      const double Ev = recoil->E / (1.0 + kd * g);
      const int extra = int(0.8 * Ev / (2.0 * element->Edisp));
      p.vac += extra;
      if (out->h_vac) out->h_vac[hidx] += extra;
    }
  }
  void replacementCollision() override {
    p.repl_hist[bin_of(recoil->pos[0], s.vac_x0, s.vac_bw, s.vac_nb)]++;
  }
  void dissipateRecoilEnergy() override { p.Enuc += recoil->E + element->Elbind; }
  void checkPKAState() override {
    p.n_coll++;
    if (out->h_coll) out->h_coll[hidx]++;
    if (out->h_Eel) out->h_Eel[hidx] += dee;
    p.Eel += dee;
    if (newton_iters == 0)
      p.rutherford++;
    else
      p.newton += newton_iters;
    switch (pka->state) {
      case LOST:
        p.lost++;
        p.Elost += pka->E;
        p.Edrop += den;
        break;
      case REPLACEMENT:
        p.repl++;
        p.Enuc += pka->E;
        break;
      case SUBSTITUTIONAL:
        p.subst++;
        p.Enuc += pka->E;
        break;
      case INTERSTITIAL:
        p.inter++;
        p.Enuc += pka->E;
        break;
      default:
        break;
    }
  }
  void flightEnd(bool vacuum) override {
    p.n_flight++;
    if (out->h_flights) out->h_flights[hidx]++;
    if (vacuum) {
      p.vacuum++;
      p.Evac += pka->E;
    }
    if (pka->gen == 0) {
      if (out->h_x) out->h_x[hidx] = pka->pos[0];
      if (out->h_y) out->h_y[hidx] = pka->pos[1];
      if (out->h_z) out->h_z[hidx] = pka->pos[2];
      if (out->h_state) out->h_state[hidx] = vacuum ? -1 : pka->state;
      if (!vacuum && pka->state != LOST) {
        p.range_n++;
        p.range_sum += pka->pos[0];
        p.range_sum2 += pka->pos[0] * pka->pos[0];
        p.range_hist[bin_of(pka->pos[0], s.range_x0, s.range_bw, s.range_nb)]++;
        p.lat_y[bin_of(pka->pos[1] - s.lat_y0, -0.5 * s.lat_nb * s.lat_bw, s.lat_bw, s.lat_nb)]++;
        p.lat_z[bin_of(pka->pos[2] - s.lat_z0, -0.5 * s.lat_nb * s.lat_bw, s.lat_bw, s.lat_nb)]++;
      }
    }
  }
};

// Event records of the reference's stream-writing TrimBase subclasses, one
// per line they print (trim.C:466-527, trim.h:139-201):
//   1 V  TrimDefectLog::vacancyCreation, the recoil     (trim.C:466-469)
//   2 I, 3 S, 4 R  TrimDefectLog::checkPKAState, the ion at rest (trim.C:471-477)
//   5    TrimHistory::followRecoil, the ion's position  (trim.h:139-155)
//   6    TrimPhononOut::followRecoil: Elbind, the recoil (trim.C:520-526)
//   7    TrimPhononOut::dissipateRecoilEnergy: recoil E + Elbind, the recoil (trim.C:512-518)
//   8    TrimPhononOut::checkPKAState: the ion's E, the ion at rest (trim.C:503-510)
// and TrimVacMap's 20 x 20 map of vacancies per (x, y) cell for three recoil
// Z (trim.C:479-501), and the layer transport of apps/mytrim_dynamic.C:160-178
// (mask 16): per (layer, Z), +1 for every flight that ends at x > 0, in the
// layer it ends in, and -1 for every recoil flight (gen > 0), in the layer it
// starts in: the atom moves SampleDynamic::addAtomsToLayer applies.  Records are written in the reference's order; the
// phonon records of a recoil carry the id and tag it has when they are
// printed (not yet assigned: 0 here, uninitialised in the reference, and -1).
struct LogHooks : TallyHooks {
  int mask;
  const int* vz;
  std::vector<orc_record>& rec;
  int* vacmap;
  long long hist = 0;
  long long* transport = nullptr;  // [material][93]
  LogHooks(const orc_run_spec& spec, Partial& part, const orc_run_out* o, int m, const int* z,
           std::vector<orc_record>& r, int* vm)
      : TallyHooks(spec, part, o), mask(m), vz(z), rec(r), vacmap(vm) {}
  // the layer index lookupLayer() returns (material index; -1: vacuum)
  static int layer_of(Sample& smp, const double pos[3]) {
    const Material* m = smp.lookupMaterial(pos);
    for (size_t k = 0; k < smp.material.size(); ++k)
      if (smp.material[k] == m) return int(k);
    return -1;
  }
  void beforeTrim(const Ion& ion, Sample& smp) override {  // apps/mytrim_dynamic.C:160-164
    if (!transport || ion.gen <= 0) return;
    const int k = layer_of(smp, ion.pos);
    if (k >= 0) transport[k * 93 + ion.Z] -= 1;
  }
  void afterTrim(const Ion& ion, Sample& smp) override {  // apps/mytrim_dynamic.C:174-178
    if (!transport || !(ion.pos[0] > 0.0)) return;
    const int k = layer_of(smp, ion.pos);
    if (k >= 0) transport[k * 93 + ion.Z] += 1;
  }
  void put(int kind, const Ion& i, double value, int id, int tag) {
    orc_record r{};
    for (int k = 0; k < 3; ++k) r.pos[k] = i.pos[k];
    r.E = i.E;
    r.value = value;
    r.m = i.m;
    r.id = (unsigned long long)(unsigned)id;
    r.hist = unsigned(hist);
    r.Z = i.Z;
    r.gen = i.gen;
    r.tag = tag;
    r.kind = kind;
    rec.push_back(r);
  }
  bool followRecoil() override {
    const bool f = TallyHooks::followRecoil();
    if (mask & 2) put(5, *pka, 0.0, pka->id, pka->tag);
    if ((mask & 4) && f) put(6, *recoil, element->Elbind, 0, -1);
    return f;
  }
  void vacancyCreation() override {
    TallyHooks::vacancyCreation();
    if (mask & 1) put(1, *recoil, 0.0, recoil->id, recoil->tag);
    if (vacmap) {  // trim.C:481-500
      int x = int((recoil->pos[0] * 20) / sample_w[0]);
      int y = int((recoil->pos[1] * 20) / sample_w[1]);
      x -= int(x / 20) * 20;
      y -= int(y / 20) * 20;
      const int c = recoil->Z == vz[0] ? 0 : recoil->Z == vz[1] ? 1 : recoil->Z == vz[2] ? 2 : -1;
      if (c >= 0 && x >= 0 && y >= 0) vacmap[(x * 20 + y) * 3 + c]++;
    }
  }
  void dissipateRecoilEnergy() override {
    TallyHooks::dissipateRecoilEnergy();
    if (mask & 4) put(7, *recoil, recoil->E + element->Elbind, 0, -1);
  }
  void checkPKAState() override {
    TallyHooks::checkPKAState();
    const int st = pka->state;
    if ((mask & 1) && (st == INTERSTITIAL || st == SUBSTITUTIONAL || st == REPLACEMENT))
      put(st == INTERSTITIAL ? 2 : st == SUBSTITUTIONAL ? 3 : 4, *pka, 0.0, pka->id, pka->tag);
    if ((mask & 4) && st != MOVING && st != LOST) put(8, *pka, pka->E, pka->id, pka->tag);
  }
  double sample_w[3] = {1.0, 1.0, 1.0};
};

Ion* make_primary(const orc_run_spec& s, long long h) {
  const long long j = s.pka_uniform ? 0 : (h - s.hist0);
  Ion* ion = new Ion;
  ion->Z = s.pka_Z[j];
  ion->m = s.pka_m[j];
  ion->E = s.pka_E[j];
  for (int i = 0; i < 3; ++i) {
    ion->pos[i] = s.pka_pos[3 * j + i];
    ion->dir[i] = s.pka_dir[3 * j + i];
  }
  ion->gen = 0;
  ion->Ef = s.Ef;
  ion->hist = uint32_t(h);
  ion->key = 0;
  return ion;
}

// Operation count of the reference's collision loop (trim.C:74-424).
// Stopping and scattering are counted by evaluating them with the counting
// scalar on the collision's own inputs; the straight-line code of the loop is
// counted here statement by statement (add/sub, mul, div, cmp, sqrt, ...):
//   v_norm(dir)                    mul 6, add 2, sqrt 1, div 1      (trim.C:85)
//   eps, eeg, pmax, ls             mul 6, add 2, div 2, sqrt 2, pow 1 (trim.C:88-92)
//   first step r1*min(ls, cw)      cmp 1, mul 1                     (trim.C:93-94)
//   p = pmax sqrt(r2)              mul 1, sqrt 1                    (trim.C:144)
//   element pick                   add 1, cmp 1 per element tried   (trim.C:148-154)
//   eps, b                         mul 1, div 1                     (trim.C:159-160)
//   dee, den, clamps, E, p1, p2    mul 7, add 3, cmp 2, sqrt 2      (trim.C:170, 275-296)
//   move + recoil dir and E        mul 9, add 5                     (trim.C:310-316)
//   azimuth, per unit-ball try     mul 6, add 5, cmp 1              (trim.C:324-328)
//   azimuth, once                  mul 9, add 5, sqrt 1, cmp 1, div 3 (trim.C:330-333)
//   psi, new dir, recoil dir       add 7, mul 9, atan2, cos, sin    (trim.C:336-341)
//   CUT test                       cmp 2 per CUT axis               (trim.C:344-352)
//   fate                           add 1, cmp 2                     (trim.C:359-406)
//   queued recoil v_norm           mul 6, add 2, sqrt 1, div 1      (trim.C:364)
//   layer lookup                   add 1, cmp 1 per layer tried     (sample_layers.C:32-37)
// Uniform draws are counted apart (they are integer work on the device).
struct WorkHooks : Hooks {
  const orc_run_spec& s;
  OpCount ops;
  long long coll = 0, uniforms = 0, flights = 0;
  MtDraws* draws = nullptr;
  Sample* sample = nullptr;
  WorkHooks(const orc_run_spec& spec) : s(spec) {}
  bool followRecoil() override {
    ops.mul += 6;
    ops.add += 2;
    ops.sqrt_ += 1;
    ops.div += 1;
    return true;
  }
  void checkPKAState() override {
    count_physics(*material, pka->Z, pka->m, E_in, eps_in, b_in, s.potential, ops);
    ++coll;
    ops.mul += 6 + 6 + 1 + 1 + 7 + 9 + 9 + 9;
    ops.add += 2 + 2 + 3 + 5 + 5 + 7 + 1;
    ops.div += 1 + 2 + 1 + 3;
    ops.sqrt_ += 1 + 2 + 1 + 2 + 1;
    ops.pow_ += 1;
    ops.cmp += 2 + 1 + 2;
    ops.atan2_ += 1;
    ops.cos_ += 1;
    ops.sin_ += 1;
    const long long tried = (element - &material->el[0]) + 1;
    ops.add += tried;
    ops.cmp += tried;
    const int t = draws ? draws->last_tries : 1;
    ops.mul += 6 * t;
    ops.add += 5 * t;
    ops.cmp += t;
    for (int i = 0; i < 3; ++i)
      if (s.bc[i] == CUT) ops.cmp += 2;
    if (s.geo == 1) {
      int k = 0;
      while (k < int(sample->material.size()) && sample->material[k] != material) ++k;
      ops.add += k + 1;
      ops.cmp += k + 1;
    }
    uniforms += 2 + 3 * t;
  }
  void flightEnd(bool) override {
    ++flights;
    ++uniforms;  // r1
  }
};

}  // namespace

// per-history std::mt19937 seed in the MT scheme (unique per history, unlike
// the 16-bit irand() seeds of apps/runmytrim.C:302)
uint32_t mt_history_seed(uint32_t master, long long h) {
  return uint32_t(flight_key_child(master, uint32_t(h)));
}

template <class Draws, class H>
static void run_history(const orc_run_spec& s, long long h, Sample& smp, Draws& draws, H& hk,
                        Partial& p) {
  const Config cfg{s.tmin, s.tau, s.cw, s.scale, s.potential};
  hk.hidx = h - s.hist0;
  std::deque<Ion*> q;
  Ion* pka = make_primary(s, h);
  p.E0 += pka->E;
  q.push_back(pka);
  while (!q.empty()) {
    Ion* ion = q.front();
    q.pop_front();
    smp.averages(*ion, cfg.tmin);
    hk.beforeTrim(*ion, smp);
    trim(ion, q, smp, cfg, draws, hk);
    hk.afterTrim(*ion, smp);
    delete ion;
  }
}

}  // namespace orc

using namespace orc;

extern "C" {

int orc_load_tables(const char* path) { return load_tables(path); }
void orc_plant_fault(int f) { orc::set_planted_fault(f); }
double orc_rpstop(int z2, double e) { return rpstop(z2, e); }
double orc_rstop(int z1, double m1, double E, int z2) { return rstop(z1, m1, E, z2); }

double orc_getrstop(int Z1, double m1, double E, double rho, int n_el, const int* Z2,
                    const double* m2, const double* t) {
  Material m;
  m.rho = rho;
  for (int i = 0; i < n_el; ++i) {
    Element e;
    e.Z = Z2[i];
    e.m = m2[i];
    e.t = t[i];
    m.el.push_back(e);
  }
  m.prepare();
  return m.getrstop(Z1, m1, E);
}

void orc_average(int Z1, double m1, double tmin, double rho, int n_el, const int* Z2,
                 const double* m2, const double* t, double* out) {
  Material m;
  m.rho = rho;
  for (int i = 0; i < n_el; ++i) {
    Element e;
    e.Z = Z2[i];
    e.m = m2[i];
    e.t = t[i];
    m.el.push_back(e);
  }
  m.prepare();
  m.average(Z1, m1, tmin);
  const double head[9] = {m.am, m.az, m.arho, m.mu, m.a, m.f, m.epsdg, m.fd, m.kd};
  for (int i = 0; i < 9; ++i) out[i] = head[i];
  for (int i = 0; i < n_el; ++i) {
    out[9 + 4 * i + 0] = m.el[i].my;
    out[9 + 4 * i + 1] = m.el[i].ec;
    out[9 + 4 * i + 2] = m.el[i].ai;
    out[9 + 4 * i + 3] = m.el[i].fi;
  }
}

int orc_scatter(double eps, double b, int potential, double* out3) {
  const Scatter s = scatter(eps, b, potential);
  out3[0] = s.s2;
  out3[1] = s.ct;
  out3[2] = s.st;
  return s.iters;
}

void orc_philox(const unsigned* c, const unsigned* k, unsigned* out4) {
  const Philox4 r = philox4x32_10(c[0], c[1], c[2], c[3], k[0], k[1]);
  for (int i = 0; i < 4; ++i) out4[i] = r.v[i];
}

// n uniforms of SimconfType::drand() after seed(seed), skipping `skip` draws
void orc_mt_stream(unsigned seed, long long skip, long long n, double* out) {
  MtRng rng;
  rng.seed(seed);
  for (long long i = 0; i < skip; ++i) rng.drand();
  for (long long i = 0; i < n; ++i) out[i] = rng.drand();
}

unsigned orc_mt_history_seed(unsigned master, long long h) { return mt_history_seed(master, h); }

unsigned long long orc_flight_key_child(unsigned long long parent, unsigned ic) {
  return flight_key_child(parent, ic);
}

unsigned long long orc_first_draw(unsigned long long key, unsigned seed, unsigned hist) {
  return first_draw(key, seed, hist);
}

static int copy_out(const std::string& s, char* buf, int len) {
  if (int(s.size()) + 1 > len) return -int(s.size()) - 1;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return int(s.size());
}

int orc_uo2(int seed, double r, double cbf, int nev, char* erec, int erec_len, char* clcoor,
            int clcoor_len, char* dist, int dist_len) {
  const Uo2Out o = run_uo2(seed, r, cbf, nev);
  if (copy_out(o.erec, erec, erec_len) < 0) return -1;
  if (copy_out(o.clcoor, clcoor, clcoor_len) < 0) return -2;
  if (copy_out(o.dist, dist, dist_len) < 0) return -3;
  return 0;
}

int orc_run(const orc_run_spec* sp, orc_run_out* out) {
  const orc_run_spec& s = *sp;
  const auto t0 = std::chrono::steady_clock::now();
  const int nt = s.nthreads > 0 ? s.nthreads : 1;
  std::vector<Partial> parts(nt);
  std::atomic<long long> next(s.hist0);
  const long long hend = s.hist0 + s.n_hist;
  auto worker = [&](int tid) {
    Partial& p = parts[tid];
    p.range_hist.assign(s.range_nb + 2, 0);
    p.lat_y.assign(s.lat_nb + 2, 0);
    p.lat_z.assign(s.lat_nb + 2, 0);
    p.vac_hist.assign(s.vac_nb + 2, 0);
    p.repl_hist.assign(s.vac_nb + 2, 0);
    std::vector<std::unique_ptr<Material>> mats;
    std::unique_ptr<Sample> smp = build_sample(s, mats);
    TallyHooks hk(s, p, out);
    MtRng rng;
    MtDraws mtd(rng);
    PhiloxDraws phd(s.seed);
    while (true) {
      const long long h = next.fetch_add(1);
      if (h >= hend) break;
      if (s.rng == 0) {
        rng.seed(mt_history_seed(s.seed, h));
        run_history(s, h, *smp, mtd, hk, p);
      } else {
        run_history(s, h, *smp, phd, hk, p);
      }
    }
  };
  if (out->h_vac) std::memset(out->h_vac, 0, sizeof(int) * s.n_hist);
  if (out->h_coll) std::memset(out->h_coll, 0, sizeof(long long) * s.n_hist);
  if (out->h_flights) std::memset(out->h_flights, 0, sizeof(int) * s.n_hist);
  if (out->h_Eel) std::memset(out->h_Eel, 0, sizeof(double) * s.n_hist);
  if (nt == 1) {
    worker(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(worker, t);
    for (auto& t : th) t.join();
  }
  Partial tot;
  tot.range_hist.assign(s.range_nb + 2, 0);
  tot.lat_y.assign(s.lat_nb + 2, 0);
  tot.lat_z.assign(s.lat_nb + 2, 0);
  tot.vac_hist.assign(s.vac_nb + 2, 0);
  tot.repl_hist.assign(s.vac_nb + 2, 0);
  for (auto& p : parts) {
    tot.n_coll += p.n_coll;
    tot.n_flight += p.n_flight;
    tot.vac += p.vac;
    tot.repl += p.repl;
    tot.subst += p.subst;
    tot.inter += p.inter;
    tot.lost += p.lost;
    tot.vacuum += p.vacuum;
    tot.rutherford += p.rutherford;
    tot.newton += p.newton;
    tot.disp += p.disp;
    tot.E0 += p.E0;
    tot.Eel += p.Eel;
    tot.Enuc += p.Enuc;
    tot.Eunf += p.Eunf;
    tot.Elost += p.Elost;
    tot.Edrop += p.Edrop;
    tot.Evac += p.Evac;
    tot.range_n += p.range_n;
    tot.range_sum += p.range_sum;
    tot.range_sum2 += p.range_sum2;
    for (size_t i = 0; i < p.range_hist.size(); ++i) tot.range_hist[i] += p.range_hist[i];
    for (size_t i = 0; i < p.lat_y.size(); ++i) {
      tot.lat_y[i] += p.lat_y[i];
      tot.lat_z[i] += p.lat_z[i];
    }
    for (size_t i = 0; i < p.vac_hist.size(); ++i) {
      tot.vac_hist[i] += p.vac_hist[i];
      tot.repl_hist[i] += p.repl_hist[i];
    }
  }
  out->n_coll = tot.n_coll;
  out->n_flight = tot.n_flight;
  out->vac = tot.vac;
  out->repl = tot.repl;
  out->subst = tot.subst;
  out->inter = tot.inter;
  out->lost = tot.lost;
  out->vacuum = tot.vacuum;
  out->rutherford = tot.rutherford;
  out->newton = tot.newton;
  out->disp = tot.disp;
  out->E0 = tot.E0;
  out->Eel = tot.Eel;
  out->Enuc = tot.Enuc;
  out->Eunf = tot.Eunf;
  out->Elost = tot.Elost;
  out->Edrop = tot.Edrop;
  out->Evac = tot.Evac;
  out->range_n = tot.range_n;
  out->range_sum = tot.range_sum;
  out->range_sum2 = tot.range_sum2;
  if (out->range_hist) std::memcpy(out->range_hist, tot.range_hist.data(), 8 * tot.range_hist.size());
  if (out->lat_y_hist) std::memcpy(out->lat_y_hist, tot.lat_y.data(), 8 * tot.lat_y.size());
  if (out->lat_z_hist) std::memcpy(out->lat_z_hist, tot.lat_z.data(), 8 * tot.lat_z.size());
  if (out->vac_hist) std::memcpy(out->vac_hist, tot.vac_hist.data(), 8 * tot.vac_hist.size());
  if (out->repl_hist) std::memcpy(out->repl_hist, tot.repl_hist.data(), 8 * tot.repl_hist.size());
  out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// Trace row (21 doubles): flight, ic, pka Z, E, pos[3], dir[3], state, ls,
// dee, den, element Z, recoil E, recoil dir[3] (before normalisation), gen
namespace {
struct TraceHooks : Hooks {
  double* rows;
  int max_rows, n = 0, flight = 0, ic = 0;
  double rE = 0.0, rdir[3] = {0, 0, 0};
  bool followRecoil() override {
    grab();
    return true;
  }
  void dissipateRecoilEnergy() override { grab(); }
  void grab() {
    rE = recoil->E;
    for (int i = 0; i < 3; ++i) rdir[i] = recoil->dir[i];
  }
  void checkPKAState() override {
    ++ic;
    if (pka->state == LOST && recoil) grab();
    if (n < max_rows) {
      double* r = rows + 21 * n;
      r[0] = flight;
      r[1] = ic;
      r[2] = pka->Z;
      r[3] = pka->E;
      for (int i = 0; i < 3; ++i) r[4 + i] = pka->pos[i];
      for (int i = 0; i < 3; ++i) r[7 + i] = pka->dir[i];
      r[10] = pka->state;
      r[11] = ls;
      r[12] = dee;
      r[13] = den;
      r[14] = element->Z;
      r[15] = rE;
      for (int i = 0; i < 3; ++i) r[16 + i] = rdir[i];
      r[19] = pka->gen;
      r[20] = 0.0;
    }
    ++n;
  }
  void flightEnd(bool) override {
    ++flight;
    ic = 0;
  }
};
}  // namespace

long long orc_run_log(const orc_run_spec* sp, int mask, const int* vacmap_Z, orc_record* records,
                      long long max_records, int* vacmap, long long* transport) {
  const orc_run_spec& s = *sp;
  Partial p;
  p.range_hist.assign(s.range_nb + 2, 0);
  p.lat_y.assign(s.lat_nb + 2, 0);
  p.lat_z.assign(s.lat_nb + 2, 0);
  p.vac_hist.assign(s.vac_nb + 2, 0);
  p.repl_hist.assign(s.vac_nb + 2, 0);
  std::vector<std::unique_ptr<Material>> mats;
  std::unique_ptr<Sample> smp = build_sample(s, mats);
  orc_run_out none{};
  std::vector<orc_record> rec;
  const int zero[3] = {0, 0, 0};
  LogHooks hk(s, p, &none, mask, vacmap_Z ? vacmap_Z : zero, rec, vacmap);
  for (int i = 0; i < 3; ++i) hk.sample_w[i] = s.w[i];
  if (mask & 16) hk.transport = transport;
  MtRng rng;
  MtDraws mtd(rng);
  PhiloxDraws phd(s.seed);
  for (long long h = s.hist0; h < s.hist0 + s.n_hist; ++h) {  // in order, one thread
    hk.hist = h;
    if (s.rng == 0) {
      rng.seed(mt_history_seed(s.seed, h));
      run_history(s, h, *smp, mtd, hk, p);
    } else {
      run_history(s, h, *smp, phd, hk, p);
    }
  }
  const long long n = (long long)rec.size();
  for (long long i = 0; i < n && i < max_records; ++i) records[i] = rec[size_t(i)];
  return n;
}

int orc_trace(const orc_run_spec* sp, long long h, double* rows, int max_rows) {
  const orc_run_spec& s = *sp;
  std::vector<std::unique_ptr<Material>> mats;
  std::unique_ptr<Sample> smp = build_sample(s, mats);
  TraceHooks hk;
  hk.rows = rows;
  hk.max_rows = max_rows;
  MtRng rng;
  MtDraws mtd(rng);
  PhiloxDraws phd(s.seed);
  const Config cfg{s.tmin, s.tau, s.cw, s.scale, s.potential};
  std::deque<Ion*> q;
  Ion* pka = make_primary(s, h);
  q.push_back(pka);
  if (s.rng == 0) rng.seed(mt_history_seed(s.seed, h));
  while (!q.empty() && hk.n < max_rows) {
    Ion* ion = q.front();
    q.pop_front();
    smp->averages(*ion, cfg.tmin);
    if (s.rng == 0)
      trim(ion, q, *smp, cfg, mtd, hk);
    else
      trim(ion, q, *smp, cfg, phd, hk);
    delete ion;
  }
  for (Ion* ion : q) delete ion;
  return hk.n < max_rows ? hk.n : max_rows;
}

// Scattering inputs of every collision of histories [hist0, hist0 + n_hist)
// (std::mt19937 draws): rows of (eps, b, E before the collision, target Z).
namespace {
struct ScatterInputHooks : Hooks {
  double* rows;
  long long max_rows, n = 0;
  void checkPKAState() override {
    if (n < max_rows) {
      double* r = rows + 4 * n;
      r[0] = eps_in;
      r[1] = b_in;
      r[2] = E_in;
      r[3] = element->Z;
    }
    ++n;
  }
};
}  // namespace

long long orc_scatter_inputs(const orc_run_spec* sp, double* rows, long long max_rows) {
  const orc_run_spec& s = *sp;
  std::vector<std::unique_ptr<Material>> mats;
  std::unique_ptr<Sample> smp = build_sample(s, mats);
  ScatterInputHooks hk;
  hk.rows = rows;
  hk.max_rows = max_rows;
  MtRng rng;
  MtDraws mtd(rng);
  const Config cfg{s.tmin, s.tau, s.cw, s.scale, s.potential};
  for (long long h = s.hist0; h < s.hist0 + s.n_hist && hk.n < max_rows; ++h) {
    std::deque<Ion*> q;
    q.push_back(make_primary(s, h));
    rng.seed(mt_history_seed(s.seed, h));
    while (!q.empty()) {
      Ion* ion = q.front();
      q.pop_front();
      smp->averages(*ion, cfg.tmin);
      trim(ion, q, *smp, cfg, mtd, hk);
      delete ion;
    }
  }
  return hk.n < max_rows ? hk.n : max_rows;
}

// Algorithmic work per collision of the reference's loop on `spec`
// (std::mt19937 draws): out = flop_eq, collisions, flights, uniforms, then
// add mul div cmp sqrt exp log pow cbrt sin cos atan2 counts per collision.
int orc_workcount(const orc_run_spec* sp, double* out) {
  const orc_run_spec& s = *sp;
  std::vector<std::unique_ptr<Material>> mats;
  std::unique_ptr<Sample> smp = build_sample(s, mats);
  WorkHooks hk(s);
  MtRng rng;
  MtDraws mtd(rng);
  hk.draws = &mtd;
  hk.sample = smp.get();
  const Config cfg{s.tmin, s.tau, s.cw, s.scale, s.potential};
  for (long long h = s.hist0; h < s.hist0 + s.n_hist; ++h) {
    rng.seed(mt_history_seed(s.seed, h));
    std::deque<Ion*> q;
    q.push_back(make_primary(s, h));
    while (!q.empty()) {
      Ion* ion = q.front();
      q.pop_front();
      smp->averages(*ion, cfg.tmin);
      trim(ion, q, *smp, cfg, mtd, hk);
      delete ion;
    }
  }
  const double n = double(hk.coll > 0 ? hk.coll : 1);
  const OpCount& o = hk.ops;
  out[0] = o.flop_eq() / n;
  out[1] = double(hk.coll);
  out[2] = double(hk.flights);
  out[3] = double(hk.uniforms) / n;
  const long long c[12] = {o.add, o.mul, o.div, o.cmp, o.sqrt_, o.exp_, o.log_, o.pow_, o.cbrt_, o.sin_, o.cos_, o.atan2_};
  for (int i = 0; i < 12; ++i) out[4 + i] = double(c[i]) / n;
  return 0;
}

}  // extern "C"
