// TrimGPU: the batched path.  Runs the full cascades of many primaries in one
// persistent-kernel launch (mt_run_cascades) with the tally hooks built in:
//   follow every recoil (TrimBase) or recoils with gen < maxGen
//   (TrimPrimaries / TrimRecoils, optionally with modified Kinchin-Pease),
//   vacancy and replacement depth bins (TrimVacCount), the phonon energy
//   ledger (TrimPhononOut), range and lateral histograms of the primaries.
// Each history draws from its own counter-based stream (seed, history), so
// results do not depend on scheduling; they are statistically equivalent to
// the host-stream path of TrimBase::trim.
#ifndef MYTRIM_AMD_TRIM_GPU_H
#define MYTRIM_AMD_TRIM_GPU_H

#include <cstdint>
#include <iosfwd>
#include <vector>

#include "sample_dynamic.h"
#include "trim.h"

namespace MyTRIM_NS {

class TrimGPU : public TrimBase {
 public:
  TrimGPU(SimconfType* simconf, SampleBase* sample) : TrimBase(simconf, sample) {}

  struct Options {
    int follow_max_gen = -1;  // < 0: every recoil
    bool kinchin_pease = false;
    Real range_x0 = 0.0, range_bin = 10.0;
    int range_bins = 1000;
    Real lat_y0 = 0.0, lat_z0 = 0.0, lat_bin = 10.0;
    int lat_bins = 400;
    Real vac_x0 = 0.0, vac_bin = 1.0;
    int vac_bins = 10000;
    bool per_history = false;
    // FP64 everywhere; FP32 physics with an FP64 ledger and positions; FP32
    // with the learned scattering surrogate (installed on the device context)
    enum Arith { FP64 = 0, FP32 = 1, FP32_SURROGATE = 2 } arith = FP64;
    // event streams (the records TrimDefectLog, TrimHistory and
    // TrimPhononOut write, trim.C:466-527): bits DEFECTS 1, HISTORY 2,
    // PHONONS 4, VACMAP 8; at most max_records records are kept.  Bit
    // TRANSPORT 16 (layers samples): the atom moves of a SampleDynamic drain
    // loop (apps/mytrim_dynamic.C:160-178) with the layers held fixed
    int log_mask = 0;
    long long max_records = 1 << 20;
    int vacmap_Z[3] = {-1, -1, -1};  // TrimVacMap z1 z2 z3
  };

  // one event record (mt_record of include/mytrim_gpu.h): kind 1 V, 2 I,
  // 3 S, 4 R (TrimDefectLog), 5 TrimHistory, 6/7/8 TrimPhononOut (recoil
  // spawned, recoil dissipated, ion at rest) with `value` the energy it prints
  struct Record {
    Real pos[3];
    Real E, value, m;
    unsigned long long id;
    unsigned int hist;
    int Z, gen, tag, kind, pad;
  };

  struct Tallies {
    long long collisions = 0, flights = 0, vacancies = 0, replacements = 0, substitutions = 0,
              interstitials = 0, lost = 0, vacuum = 0, displacements = 0;
    Real E0 = 0, Eel = 0, Enuc = 0, Eunfollowed = 0, Elost = 0, Edropped = 0, Evacuum = 0;
    long long range_n = 0;
    Real range_sum = 0, range_sum2 = 0;
    std::vector<long long> range_hist, lat_y_hist, lat_z_hist, vac_hist, repl_hist;
    std::vector<int> h_vac, h_flights, h_state;
    std::vector<long long> h_coll;
    std::vector<Real> h_x, h_y, h_z, h_Eel;
    double kernel_ms = 0;
    std::vector<Record> records;  // log_mask != 0
    long long n_records = 0;      // produced (records past max_records are dropped)
    std::vector<int> vacmap;      // [20][20][3] (LOG VACMAP)
    // [layer][Z] net atoms of element Z that came to rest in the layer (+1
    // per flight ending at x > 0) minus those whose recoil flight started
    // there (-1 per flight with gen > 0) (LOG TRANSPORT; Z = 0..92)
    std::vector<long long> transport;
  };

  // the composition update between batches: addAtomsToLayer(layer, n, Z) for
  // every non-zero entry of t.transport, layer by layer and Z by Z
  static void applyTransport(SampleDynamic& sample, const Tallies& t);

  // the lines TrimDefectLog (kinds 1-4) or TrimPhononOut (kinds 6-8) would
  // print for `records` (IonBase's operator<<, ion.C:74-79), in record order
  static void writeRecords(std::ostream& os, const std::vector<Record>& records, bool phonons);

  // n copies of `pka` (histories hist0 .. hist0+n-1 of stream `seed`)
  Tallies runCascades(const IonBase& pka, long long n, uint32_t seed, const Options& opt,
                      long long hist0 = 0);
  // one primary per history
  Tallies runCascades(const std::vector<IonBase*>& pkas, uint32_t seed, const Options& opt,
                      long long hist0 = 0);
  // time candidate launch shapes on `sample` copies of `pka`, keep the fastest
  // for later runCascades calls; returns its collisions per kernel second
  double autotune(const IonBase& pka, long long sample, const Options& opt, int reps = 1);
};

}  // namespace MyTRIM_NS

#endif
