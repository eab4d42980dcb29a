// Fission fragments slowing down in UO2 that holds Xe bubbles, driven
// through the drop-in C++ API (every TrimBase::trim() flight runs on the GPU
// with the SimconfType stream; the hooks replay on the host).
//
// This is the scenario of MyTRIM's regression test tests/uo2 (the PLAIN mode
// of apps/mytrim_uo2.C run as `mytrim_uo2 out 10 0.1 1` with
// MYTRIM_SEED=39172).  The gold files fix the order of the random draws and
// the output formats:
//   1. bubbles: sampleClusters::addRandomClusters (surfaces >= 25 A apart);
//   2. per fission event: fragment mass, total kinetic energy, an isotropic
//      direction by rejection in the unit ball, a uniform origin;
//   3. the two fragments and every recoil they make, drained first in first
//      out, each flight from its own trim() call.
//
//   fission_bubbles BASE RADIUS BUBBLE_FACTOR EVENTS      (seed: $MYTRIM_SEED)
//
// writes BASE.clcoor (bubble centres, radii, tags), BASE.Erec (energy,
// generation and MD mark of every Xe recoil) and BASE.dist (for Xe atoms
// that start in a bubble: distance from the bubble's centre where they stop,
// MD mark, stopping point).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <queue>
#include <string>
#include <utility>

#include "invert.h"
#include "ion.h"
#include "material.h"
#include "sample_clusters.h"
#include "simconf.h"
#include "trim.h"

using namespace MyTRIM_NS;

namespace {

constexpr int kXe = 54;
constexpr double kBoxSide = 400.0;        // A, periodic
constexpr double kBubblesPerA3 = 7.0e-7;  // at BUBBLE_FACTOR 1 (7e-4 per nm^3)
constexpr double kBubbleGap = 25.0;       // A between bubble surfaces

struct File {
  explicit File(const std::string& path) : f(std::fopen(path.c_str(), "wt")) {}
  ~File() {
    if (f) std::fclose(f);
  }
  std::FILE* f;
};

// UO2 matrix (material 0) and Xe gas (material 1)
void add_materials(sampleClusters& s, SimconfType* sc) {
  auto add = [&](double rho, std::initializer_list<std::pair<int, std::pair<double, double>>> els) {
    MaterialBase* m = new MaterialBase(sc, rho);
    for (const auto& e : els) {
      Element el;
      el._Z = e.first;
      el._m = e.second.first;
      el._t = e.second.second;
      m->_element.push_back(el);
    }
    m->prepare();
    s.material.push_back(m);
  };
  add(10.0, {{92, {235.0, 1.0}}, {8, {16.0, 2.0}}});
  add(3.5, {{kXe, {132.0, 1.0}}});
}

// The two fragments of one fission of U-235, back to back from a uniform
// point of the box.
class FissionSource {
 public:
  std::pair<IonMDTag*, IonMDTag*> next(SimconfType* sc, const sampleClusters& s) {
    const Real a1 = masses_.x(sc->drand());
    const Real a2 = 235.0 - a1;
    energies_.setMass(a1);
    const Real e_total = energies_.x(sc->drand());  // MeV
    // momentum balance: the energies split inversely to the masses
    const Real e1 = e_total * a2 / (a1 + a2);
    Real d[3], n2;
    do {
      for (Real& c : d) c = 2.0 * sc->drand() - 1.0;
      n2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    } while (n2 <= 0.0001 || n2 > 1.0);
    // unit vector by a division per component (bit for bit the direction
    // the gold trajectories start from)
    const Real norm = std::sqrt(n2);
    for (Real& c : d) c /= norm;
    Real origin[3];
    for (int i = 0; i < 3; ++i) origin[i] = sc->drand() * s.w[i];
    const int z1 = int(std::round(a1 * 92.0 / 235.0));
    IonMDTag* f1 = fragment(z1, a1, e1, origin, d, false);
    IonMDTag* f2 = fragment(92 - z1, a2, e_total - e1, origin, d, true);
    return {f1, f2};
  }

 private:
  static IonMDTag* fragment(int Z, Real A, Real e_mev, const Real* origin, const Real* d, bool reverse) {
    IonMDTag* f = new IonMDTag;
    f->_gen = 0;
    f->_tag = -1;
    f->_md = 0;
    f->_Z = Z;
    f->_m = A;
    f->_E = e_mev * 1.0e6;
    for (int i = 0; i < 3; ++i) {
      f->_dir(i) = reverse ? -d[i] : d[i];
      f->_pos(i) = origin[i];
    }
    f->setEf();
    return f;
  }
  MassInverter masses_;
  EnergyInverter energies_;
};

// What the scenario records of Xe flights, before and after each trim().
class XeRecorder {
 public:
  XeRecorder(const std::string& base, const sampleClusters& s)
      : s_(s), erec_(base + ".Erec"), dist_(base + ".dist") {}
  bool ok() const { return erec_.f && dist_.f; }

  void before(IonMDTag& ion) {
    if (ion._Z != kXe) return;
    // an Xe atom that starts its flight in the MD energy window (200 eV to
    // 12 keV) is marked 1; recoils start unmarked (IonMDTag::spawnRecoil)
    if (ion._md == 0 && ion._E > 200.0 && ion._E < 12000.0) ion._md = 1;
    if (ion._gen > 0) std::fprintf(erec_.f, "%f\t%d\t%d\n", ion._E, ion._gen, ion._md);
    if (ion._tag >= 0) {
      // centre of the bubble the atom starts in, as the periodic image
      // nearest to the atom
      for (int i = 0; i < 3; ++i) {
        Real d = s_.c[i][ion._tag] - ion._pos(i);
        if (s_.bc[i] == SampleBase::PBC) d -= std::round(d / s_.w[i]) * s_.w[i];
        centre_[i] = ion._pos(i) + d;
      }
    }
  }

  void after(const IonMDTag& ion) {
    if (ion._Z != kXe || ion._tag < 0) return;
    Real r2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      const Real d = centre_[i] - ion._pos(i);
      r2 += d * d;
    }
    std::fprintf(dist_.f, "%f %d %f %f %f\n", std::sqrt(r2), ion._md, ion._pos(0), ion._pos(1), ion._pos(2));
  }

 private:
  const sampleClusters& s_;
  File erec_, dist_;
  Real centre_[3] = {0.0, 0.0, 0.0};
};

}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr,
                 "usage: %s BASE RADIUS BUBBLE_FACTOR EVENTS\n"
                 "  fission fragments in UO2 with Xe bubbles of RADIUS A, %g bubbles per A^3\n"
                 "  times BUBBLE_FACTOR; the seed is taken from $MYTRIM_SEED\n",
                 argv[0], kBubblesPerA3);
    return 1;
  }
  const std::string base = argv[1];
  const Real radius = std::atof(argv[2]);
  const Real factor = std::atof(argv[3]);
  const int events = std::atoi(argv[4]);
  const char* seed_env = std::getenv("MYTRIM_SEED");
  if (!seed_env) {
    std::fprintf(stderr, "set MYTRIM_SEED\n");
    return 1;
  }
  const int seed = std::atoi(seed_env);

  std::unique_ptr<SimconfType> sc(new SimconfType);
  sc->seed(unsigned(seed < 0 ? -seed : seed));
  sampleClusters sample(kBoxSide, kBoxSide, kBoxSide);
  // hash cells about one radius wide
  sample.initSpatialhash(int(sample.w[0] / radius) - 1, int(sample.w[1] / radius) - 1,
                         int(sample.w[2] / radius) - 1);
  const int n_bubbles = int(sample.w[0] * sample.w[1] * sample.w[2] * kBubblesPerA3 * factor);
  sample.addRandomClusters(unsigned(n_bubbles), radius, kBubbleGap, sc.get());
  {
    File cl(base + ".clcoor");
    if (!cl.f) return 1;
    for (int i = 0; i < sample.cn; ++i)
      std::fprintf(cl.f, "%f %f %f %f %d\n", sample.c[0][i], sample.c[1][i], sample.c[2][i], sample.c[3][i], i);
  }
  add_materials(sample, sc.get());

  TrimBase trim(sc.get(), &sample);
  FissionSource source;
  XeRecorder rec(base, sample);
  if (!rec.ok()) return 1;
  std::queue<IonBase*> pending;
  long long flights = 0;
  for (int ev = 0; ev < events; ++ev) {
    const auto frags = source.next(sc.get(), sample);
    pending.push(frags.first);
    pending.push(frags.second);
    while (!pending.empty()) {
      IonMDTag* ion = dynamic_cast<IonMDTag*>(pending.front());
      pending.pop();
      sample.averages(ion);
      rec.before(*ion);
      trim.trim(ion, pending);
      rec.after(*ion);
      delete ion;
      ++flights;
    }
  }
  std::printf("%d bubbles, %d fission events, %lld flights\n", sample.cn, events, flights);
  return 0;
}
