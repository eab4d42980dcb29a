// TrimGPU::runCascades: the batched device path (mt_run_cascades).
#include "trim_gpu.h"
#include "api_error.h"

#include <climits>
#include <ostream>
#include <stdexcept>

#include "device.h"
#include "mytrim_gpu.h"

namespace MyTRIM_NS {

namespace {

struct Launch {
  mt_primaries pk;
  mt_run_opts ro;
};

Launch launch_args(long long n, bool uniform, const std::vector<int>& Z, const std::vector<double>& m,
                   const std::vector<double>& E, const std::vector<double>& pos, const std::vector<double>& dir,
                   double Ef, uint32_t seed, long long hist0, const TrimGPU::Options& o) {
  Launch L;
  mt_primaries& pk = L.pk;
  pk = {};
  pk.n_hist = n;
  pk.hist0 = hist0;
  pk.pka_uniform = uniform ? 1 : 0;
  pk.Z = Z.data();
  pk.m = m.data();
  pk.E = E.data();
  pk.pos = pos.data();
  pk.dir = dir.data();
  pk.Ef = Ef;
  pk.seed = seed;
  mt_run_opts& ro = L.ro;
  ro = {};
  ro.follow_max_gen = o.follow_max_gen;
  ro.kp = o.kinchin_pease ? 1 : 0;
  ro.arith = int(o.arith);
  ro.range_x0 = o.range_x0;
  ro.range_bw = o.range_bin;
  ro.range_nb = o.range_bins;
  ro.lat_y0 = o.lat_y0;
  ro.lat_z0 = o.lat_z0;
  ro.lat_bw = o.lat_bin;
  ro.lat_nb = o.lat_bins;
  ro.vac_x0 = o.vac_x0;
  ro.vac_bw = o.vac_bin;
  ro.vac_nb = o.vac_bins;
  ro.per_history = o.per_history ? 1 : 0;
  return L;
}

TrimGPU::Tallies run(mt_ctx* ctx, long long n, bool uniform, const std::vector<int>& Z,
                     const std::vector<double>& m, const std::vector<double>& E, const std::vector<double>& pos,
                     const std::vector<double>& dir, double Ef, uint32_t seed, long long hist0,
                     const TrimGPU::Options& o, size_t n_mat) {
  Launch L = launch_args(n, uniform, Z, m, E, pos, dir, Ef, seed, hist0, o);
  TrimGPU::Tallies t;
  t.range_hist.assign(size_t(o.range_bins + 2), 0);
  t.lat_y_hist.assign(size_t(o.lat_bins + 2), 0);
  t.lat_z_hist.assign(size_t(o.lat_bins + 2), 0);
  t.vac_hist.assign(size_t(o.vac_bins + 2), 0);
  t.repl_hist.assign(size_t(o.vac_bins + 2), 0);
  mt_tallies out = {};
  out.range_hist = t.range_hist.data();
  out.lat_y_hist = t.lat_y_hist.data();
  out.lat_z_hist = t.lat_z_hist.data();
  out.vac_hist = t.vac_hist.data();
  out.repl_hist = t.repl_hist.data();
  if (o.per_history) {
    t.h_vac.assign(size_t(n), 0);
    t.h_coll.assign(size_t(n), 0);
    t.h_flights.assign(size_t(n), 0);
    t.h_state.assign(size_t(n), 0);
    t.h_x.assign(size_t(n), 0.0);
    t.h_y.assign(size_t(n), 0.0);
    t.h_z.assign(size_t(n), 0.0);
    t.h_Eel.assign(size_t(n), 0.0);
    out.h_vac = t.h_vac.data();
    out.h_coll = t.h_coll.data();
    out.h_flights = t.h_flights.data();
    out.h_state = t.h_state.data();
    out.h_x = t.h_x.data();
    out.h_y = t.h_y.data();
    out.h_z = t.h_z.data();
    out.h_Eel = t.h_Eel.data();
  }
  if (o.log_mask) {
    static_assert(sizeof(TrimGPU::Record) == sizeof(mt_record), "record layout");
    mt_log lg = {};
    lg.mask = o.log_mask;
    t.records.resize(size_t(std::max<long long>(o.max_records, 0)));
    lg.records = reinterpret_cast<mt_record*>(t.records.data());
    lg.max_records = (long long)t.records.size();
    for (int i = 0; i < 3; ++i) lg.vacmap_Z[i] = o.vacmap_Z[i];
    t.vacmap.assign(20 * 20 * 3, 0);
    lg.vacmap = t.vacmap.data();
    if (o.log_mask & MT_LOG_TRANSPORT) {
      t.transport.assign(n_mat * 93, 0);
      lg.transport = t.transport.data();
    }
    device::check(mt_run_cascades_log(ctx, &L.pk, &L.ro, &lg, &out), "mt_run_cascades_log");
    t.n_records = lg.n_records;
    t.records.resize(size_t(std::min<long long>(lg.n_records, lg.max_records)));
  } else {
    device::check(mt_run_cascades(ctx, &L.pk, &L.ro, &out), "mt_run_cascades");
  }
  t.collisions = out.n_coll;
  t.flights = out.n_flight;
  t.vacancies = out.vac;
  t.replacements = out.repl;
  t.substitutions = out.subst;
  t.interstitials = out.inter;
  t.lost = out.lost;
  t.vacuum = out.vacuum;
  t.displacements = out.disp;
  t.E0 = out.E0;
  t.Eel = out.Eel;
  t.Enuc = out.Enuc;
  t.Eunfollowed = out.Eunf;
  t.Elost = out.Elost;
  t.Edropped = out.Edrop;
  t.Evacuum = out.Evac;
  t.range_n = out.range_n;
  t.range_sum = out.range_sum;
  t.range_sum2 = out.range_sum2;
  t.kernel_ms = out.kernel_ms;
  return t;
}

}  // namespace

TrimGPU::Tallies TrimGPU::runCascades(const IonBase& p, long long n, uint32_t seed, const Options& o,
                                      long long hist0) {
  std::vector<int> Z{p._Z};
  std::vector<double> m{p._m}, E{p._E}, pos{p._pos(0), p._pos(1), p._pos(2)}, dir{p._dir(0), p._dir(1), p._dir(2)};
  Tallies t = run(device(), n, true, Z, m, E, pos, dir, p._Ef, seed, hist0, o, _sample->material.size());
  _simconf->vacancies_created += int(t.vacancies);
  _simconf->EelTotal += t.Eel;
  _simconf->EnucTotal += t.Enuc;
  return t;
}

TrimGPU::Tallies TrimGPU::runCascades(const std::vector<IonBase*>& pkas, uint32_t seed, const Options& o,
                                      long long hist0) {
  std::vector<int> Z;
  std::vector<double> m, E, pos, dir;
  double Ef = pkas.empty() ? 3.0 : pkas[0]->_Ef;
  for (IonBase* p : pkas) {
    if (p->_Ef != Ef) api_error("runCascades: all primaries must share _Ef");
    Z.push_back(p->_Z);
    m.push_back(p->_m);
    E.push_back(p->_E);
    for (int i = 0; i < 3; ++i) {
      pos.push_back(p->_pos(i));
      dir.push_back(p->_dir(i));
    }
  }
  Tallies t = run(device(), (long long)pkas.size(), false, Z, m, E, pos, dir, Ef, seed, hist0, o,
                  _sample->material.size());
  _simconf->vacancies_created += int(t.vacancies);
  _simconf->EelTotal += t.Eel;
  _simconf->EnucTotal += t.Enuc;
  return t;
}

void TrimGPU::applyTransport(SampleDynamic& sample, const Tallies& t) {
  const size_t nl = sample.material.size();
  if (t.transport.size() != nl * 93) api_error("applyTransport: no transport tally for this sample");
  for (size_t k = 0; k < nl; ++k)
    for (int Z = 1; Z <= 92; ++Z) {
      const long long n = t.transport[k * 93 + size_t(Z)];
      if (n == 0) continue;
      if (n > INT_MAX || n < INT_MIN) api_error("applyTransport: count out of int range");
      sample.addAtomsToLayer(int(k), int(n), Z);
    }
}

void TrimGPU::writeRecords(std::ostream& os, const std::vector<Record>& records, bool phonons) {
  static const char kDefect[] = " VISR";
  for (const Record& r : records) {
    if (phonons ? (r.kind < 6 || r.kind > 8) : (r.kind < 1 || r.kind > 4)) continue;
    if (phonons)
      os << r.value << ' ';
    else
      os << kDefect[r.kind] << ' ';
    os << r.pos[0] << ' ' << r.pos[1] << ' ' << r.pos[2] << ' ' << r.Z << ' ' << r.m << ' ' << r.E << ' '
       << r.id << ' ' << r.gen << ' ' << r.tag << ' ' << '\n';
  }
}

double TrimGPU::autotune(const IonBase& p, long long sample, const Options& o, int reps) {
  std::vector<int> Z{p._Z};
  std::vector<double> m{p._m}, E{p._E}, pos{p._pos(0), p._pos(1), p._pos(2)}, dir{p._dir(0), p._dir(1), p._dir(2)};
  Launch L = launch_args(sample, true, Z, m, E, pos, dir, p._Ef, 1u, 0, o);
  mt_tuning best = {};
  double rate = 0.0;
  device::check(mt_autotune(device(), &L.pk, &L.ro, sample, reps, &best, &rate), "mt_autotune");
  return rate;
}

}  // namespace MyTRIM_NS
