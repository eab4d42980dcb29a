// ORACLE — test infrastructure only.
//
// A CPU restatement, in FP64, of the MyTRIM binary-collision cascade path
// (reference: trim.C:35-527, material.C:37-282, simconf.C/h, ion.C,
// element.C, sample*.C, functions.h, shim/point.C).  It follows the
// reference's arithmetic order and its random-number draw order so that, fed
// by the same std::mt19937 stream, it reproduces the reference bit for bit
// (pinned by the reference's own tests/uo2 gold files and by fixtures written
// by the compiled reference, see tests/golden/).  A second draw scheme
// ("philox") restates the device path's counter-based streams so that the
// HIP kernels can be compared history by history.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this code, and only as the checker or the timed CPU baseline; the
// product (mytrim_amd/) never links or calls it.
#pragma once
#include <cmath>
#include <cstdint>
#include <deque>
#include <random>
#include <string>
#include <vector>

#include "philox_ref.h"

namespace orc {

// ---------------------------------------------------------------- tables
// One row of SCOEF.95A + SLFCTR.dat (simconf.h:71-90); only the columns the
// stopping code reads.
struct Coef {
  double mm1, m1, mnat, rho, atrho, vfermi, heat;
  double pcoef[8];
  double lfctr;
};
extern Coef g_coef[92];
int load_tables(const char* path);  // 0 on success

// ---------------------------------------------------------------- target
struct Element {  // element.h:29-41, element.C:25
  int Z = 0;
  double m = 0.0, t = 0.0;
  double Edisp = 25.0, Elbind = 3.0;
  double my = 0.0, ec = 0.0, ai = 0.0, fi = 0.0;
};

struct Material {  // material.h:34-75
  double rho = 0.0;
  double am = 0.0, az = 0.0, arho = 0.0;
  double mu = 0.0, a = 0.0, f = 0.0, epsdg = 0.0, fd = 0.0, kd = 0.0;
  double pmax = 0.0;
  int tag = -1;
  std::vector<Element> el;

  void prepare();                                    // material.C:37-74
  void average(int Z, double m, double tmin);        // material.C:78-110
  double getrstop(int Z, double m, double E) const;  // material.C:114-122
};

double rpstop(int z2p, double e);                   // material.C:134-158
double rstop(int z1, double m, double E, int z2);   // material.C:161-282

// ---------------------------------------------------------------- ions
enum State { MOVING = 0, REPLACEMENT, SUBSTITUTIONAL, INTERSTITIAL, LOST, DELETE, VACANCY };
enum Bc { PBC = 0, INF = 1, CUT = 2 };

struct Ion {  // ion.h:30-85 (+ IonMDTag::_md, + device-scheme stream ids)
  int Z = 0;
  double m = 0.0, E = 0.0;
  double dir[3] = {0.0, 0.0, 0.0};
  double pos[3] = {0.0, 0.0, 0.0};
  unsigned seed = 0;
  int gen = 0, id = 0, tag = -1;
  double Ef = 3.0;
  int state = MOVING;
  int md = 0;
  uint64_t key = 0;   // flight key (philox scheme)
  uint32_t hist = 0;  // history index (philox scheme)
};

// ---------------------------------------------------------------- samples
struct Sample {  // sample.h:33-58
  double w[3];
  int bc[3];
  std::vector<Material*> material;
  Sample(double x, double y, double z);
  virtual ~Sample() {}
  void averages(const Ion& ion, double tmin);
  virtual Material* lookupMaterial(const double pos[3]) = 0;
};

struct SampleSolid : Sample {  // sample_solid.C:25-29
  SampleSolid(double x, double y, double z) : Sample(x, y, z) {}
  Material* lookupMaterial(const double*) override { return material[0]; }
};

struct SampleLayers : Sample {  // sample_layers.C:26-49
  std::vector<double> layerThickness;
  SampleLayers(double x, double y, double z) : Sample(x, y, z) {}
  int lookupLayer(const double pos[3]) const;
  Material* lookupMaterial(const double pos[3]) override { return material[lookupLayer(pos)]; }
};

struct SampleWire : Sample {  // sample_wire.C:30-46
  SampleWire(double x, double y, double z) : Sample(x, y, z) {
    bc[0] = CUT;
    bc[1] = CUT;
  }
  Material* lookupMaterial(const double pos[3]) override;
};

struct SampleBurriedWire : SampleWire {  // sample_burried_wire.C:29-55
  SampleBurriedWire(double x, double y, double z) : SampleWire(x, y, z) { bc[0] = bc[1] = bc[2] = INF; }
  Material* lookupMaterial(const double pos[3]) override;
};

struct SampleClusters : Sample {  // sample_clusters.C:31-252
  double sd = 0.0, kd[3] = {0, 0, 0};
  int kn[3] = {0, 0, 0};
  std::vector<int> sh;   // cell -> first cluster (-1)
  std::vector<int> cl;   // cluster -> next cluster in the same cell (-1)
  std::vector<double> c[4];
  int cn = 0;
  double cmr = 0.0;
  SampleClusters(double x, double y, double z) : Sample(x, y, z) {}
  Material* lookupMaterial(const double pos[3]) override;
  int lookupCluster(const double pos[3], double dr) const;
  void initSpatialhash(int x, int y, int z);
  void addCluster(double x, double y, double z, double r);
  template <class Rng>
  void addRandomClusters(unsigned n, double r, double dr, Rng& rng) {
    for (unsigned i = 0; i < n; ++i) {
      while (true) {
        double npos[3];
        for (int j = 0; j < 3; ++j) npos[j] = rng.drand() * w[j];
        if (lookupCluster(npos, dr + r) == -1) {
          addCluster(npos[0], npos[1], npos[2], r);
          break;
        }
      }
    }
  }
};

// ---------------------------------------------------------------- config
struct Config {  // simconf.C:44-49 defaults
  double tmin = 0.2, tau = 0.0, cw = 0.001, scale = 1.0;
  int potential = 0;  // 0 universal, 1 Moliere, 2 C-Kr (trim.h:63-68)
};

// std::mt19937 + uniform_real_distribution<double>(0,1): SimconfType::drand
struct MtRng {  // simconf.h:52-54, simconf.C:36-43, 66-71
  std::mt19937 gen;
  std::uniform_real_distribution<double> dis{0.0, 1.0};
  std::uniform_int_distribution<unsigned> idis{0, 65535};
  explicit MtRng(unsigned s = 12345678u) : gen(s) {}
  void seed(unsigned s) {
    gen.seed(s);
    dis.reset();
    idis.reset();
  }
  double drand() { return dis(gen); }
  unsigned irand() { return idis(gen); }
};

// ---------------------------------------------------------------- scattering
struct Scatter {
  double s2, ct, st;
  int iters;  // Newton iterations (0 for the Rutherford branch)
};
// trim.C:172-272: CM scattering for reduced energy eps and impact parameter b
Scatter scatter(double eps, double b, int potential);

// ---------------------------------------------------------------- hooks
// Mirrors the protected virtual hooks of TrimBase (trim.h:72-84) and the
// state they can read (trim.h:86-107).
struct Hooks {
  Ion* pka = nullptr;
  Ion* recoil = nullptr;
  Material* material = nullptr;
  const Element* element = nullptr;
  double ls = 0.0, dee = 0.0, den = 0.0;
  int newton_iters = 0;  // of the last collision (0 = Rutherford branch)
  // inputs of the last collision's physics (ion energy before it, reduced
  // energy and impact parameter), recorded for operation counting
  double E_in = 0.0, eps_in = 0.0, b_in = 0.0;
  long long vacancies_created = 0;
  double EelTotal = 0.0, EnucTotal = 0.0;
  int next_id = 0;

  virtual ~Hooks() {}
  virtual bool followRecoil() { return true; }               // trim.C:427-437
  virtual void vacancyCreation() { vacancies_created++; }    // trim.C:439-443
  virtual void replacementCollision() {}
  virtual void checkPKAState() {}
  virtual void dissipateRecoilEnergy() {}
  // not a reference hook: called once after the collision loop ends
  virtual void flightEnd(bool /*vacuum*/) {}
  // not reference hooks: the drain loop calls these before and after each
  // trim() (the places apps/mytrim_dynamic.C:160-178 updates SampleDynamic)
  virtual void beforeTrim(const Ion&, Sample&) {}
  virtual void afterTrim(const Ion&, Sample&) {}
};

// ---------------------------------------------------------------- draws
// Reference draw order: r1 at flight start, then per collision r2, hh and
// the unit-ball rejection loop for the azimuth (trim.C:71, 143, 147, 322-333)
struct MtDraws {
  MtRng& rng;
  int last_tries = 0;  // unit-ball draws of the last azimuth (operation counting)
  explicit MtDraws(MtRng& r) : rng(r) {}
  double first(const Ion&) { return rng.drand(); }
  void collision(const Ion&, int) {}
  double r2() { return rng.drand(); }
  double hh() { return rng.drand(); }
  void perp(const double d[3], double out[3]);
};

// Deliberate error planted in the device draw scheme for tests of the tests
// (0: none; 1: azimuth over [0, pi) instead of [0, 2 pi)).  Test use only.
int planted_fault();
void set_planted_fault(int f);

// Device scheme: one Philox4x32-10 block per collision, keyed by
// (seed, history) and counted by (collision index, flight key); the azimuth
// about the flight direction is drawn directly.
struct PhiloxDraws {
  uint32_t seed;
  Philox4 blk;
  explicit PhiloxDraws(uint32_t s) : seed(s) {}
  double first(const Ion& ion);
  void collision(const Ion& ion, int ic);
  double r2() const { return u53(blk.v[0], blk.v[1]); }
  double hh() const { return u32d(blk.v[2]); }
  void perp(const double d[3], double out[3]);
};

// ---------------------------------------------------------------- flight
// TrimBase::trim (trim.C:35-425).  Recoils are appended to `recoils`.
template <class Draws>
void trim(Ion* pka, std::deque<Ion*>& recoils, Sample& sample, const Config& cfg, Draws& draws,
          Hooks& h);

// ---------------------------------------------------------------- drivers
// Operations of one collision's stopping and scattering evaluated on the
// same inputs with the counting scalar (oracle/src/counted.h).
struct OpCount;
void count_physics(const Material& mat, int Z, double m, double E, double eps, double b,
                   int potential, OpCount& ops);

// apps/mytrim_uo2.C:48-358 in PLAIN mode; returns the three output files
struct Uo2Out {
  std::string erec, clcoor, dist;
  double EelTotal = 0.0, EnucTotal = 0.0;
};
Uo2Out run_uo2(int seed, double r, double cbf, int nev);

}  // namespace orc
