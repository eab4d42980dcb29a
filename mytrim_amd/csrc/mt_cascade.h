// Launch parameters of the persistent cascade kernel (shared by the host
// launcher and the device code).
#pragma once
#include <stdint.h>

#include "mt_mlp.h"
#include "mt_tables.h"

namespace mt {

// HBM recoil queue: structure-of-arrays ring.  Every field is written and
// read with agent-scope relaxed atomics (coherent across the XCDs without
// L2 write-back / invalidate fences); the ready word of a slot holds the lap
// (position / capacity + 1) in its high half and a checksum of the entry in
// its low half, so a reader that sees the lap but stale fields re-reads.
// After copying its entry the reader stamps the slot consumed_word(lap).
struct Queue {
  uint64_t *px, *py, *pz, *dx, *dy, *dz, *E;  // FP64 bit patterns
  uint64_t* key;
  uint64_t* mt;     // meta << 32 | tag, meta = species | (gen << 8)
  uint32_t* hist;
  uint64_t* ready;
  // the ring is split into nq shards of 2^sub_log2 slots; shard s owns slots
  // [s << sub_log2, (s + 1) << sub_log2) and its own head / tail words
  uint64_t sub_mask;
  int sub_log2, nq;
};
constexpr int kMaxShards = 16;
constexpr int kQueueArrays = 11;
constexpr int kQueueReady = 10;  // index of the ready words among the arrays

// control words (one HBM allocation); the words waves contend on sit on
// cache lines (128 B = 16 words) of their own
enum Ctrl {
  kState = 0,           // epoch << 32 | live waves << 16 | idle waves
  kPkaNext = 16,        // next primary to start
  kLogCount = 32,       // event records produced (diagnostic kernels)
  kError = 64,          // error bits
  kQueueOps = 65,       // wave-level queue transfers
  // steady-state throughput of a launch, read by the launch-shape tuner:
  // wall-clock stamp of the first wave's start, of the admission of the last
  // primary, and the flights ended by then (kFlightsEnded at that moment; each
  // wave adds its ended flights in batches of up to 256)
  kT0 = 66,
  kTAdmit = 67,
  kFlightsAtAdmit = 68,
  // the waves' shader cycles and wall-clock ticks from start to exit, summed:
  // their ratio is the launch's mean shader clock
  kClkSum = 69,
  kWallSum = 70,
  // wall-clock stamp of the admission of the middle primary (index
  // n_hist / 2) and the flights ended by then: the launch's ramp-up is over
  kTHalf = 71,
  kFlightsAtHalf = 72,
  kQueuePeak = 80,      // largest tail - head seen at a push
  kLive = 96,           // admitted cascades not yet complete
  kCascDone = 97,       // completed cascades
  kFlightsEnded = 112,
  kQBase = 128,         // shard s: next position to pop at kQBase + 32 s, to push at + 16
  kNumCtrl = kQBase + 32 * kMaxShards
};
__host__ __device__ constexpr int kQHead(int s) { return kQBase + 32 * s; }
__host__ __device__ constexpr int kQTail(int s) { return kQBase + 32 * s + 16; }
// ready word of a slot whose lap-`lap` entry has been read out: the lap with
// its top bit set (a real lap never has it), checksum half zero.  A pusher of
// lap L writes only over the marker of lap L - 1 (or the zero of a fresh run).
__host__ __device__ constexpr uint64_t consumed_word(uint32_t lap) {
  return (uint64_t)(lap | 0x80000000u) << 32;
}

enum ErrBits {
  kErrQueueFull = 1,
  kErrRunaway = 2,    // flight exceeded max collisions
  kErrBadEntry = 4,   // queue entry failed validation
  kErrTimeout = 8,    // a wait exceeded the wall-clock limit
  kErrNewton = 16,    // Newton solve hit its iteration cap
  kErrNonFinite = 32  // a flight's energy or position became NaN or infinite
};

enum Counter {
  cColl = 0, cFlight, cVac, cRepl, cSubst, cInter, cLost, cVacuum, cRuth, cNewton, cDisp,
  cRangeN, cSurr, cCollPrimary, kNumCounters = 16
};

enum Energy { eEel = 0, eEnuc, eEunf, eElost, eEdrop, eEvac, eRangeSum, eRangeSum2, kNumEnergy = 8 };

// Event records of the diagnostic (per-history) kernels, one per line the
// reference's stream-writing TrimBase subclasses print (the layout of
// mt_record): kinds 1 V (recoil at vacancyCreation), 2 I / 3 S / 4 R (ion at
// rest), 5 TrimHistory (ion when a recoil is followed), 6/7/8 TrimPhononOut
// (spawned recoil: Elbind; dissipated recoil: E + Elbind; ion at rest: E).
struct LogRecord {
  double pos[3];
  double E, value, m;
  unsigned long long id;
  uint32_t hist;
  int Z, gen, tag, kind, pad;
};
constexpr int kLogDefects = 1, kLogHistory = 2, kLogPhonons = 4, kLogVacmap = 8, kLogTransport = 16;
constexpr int kTransportZ = 93;  // transport columns per layer: Z = 0..92
constexpr int kVacMapN = 20;  // TrimVacMap cells per axis (trim.h:178)
struct LogDev {
  int mask;                // kLog* bits; 0: no records
  LogRecord* rec;          // cap records; the count runs on past cap
  unsigned long long cap;
  int vz[3];               // TrimVacMap z1 z2 z3
  int* vacmap;             // [20][20][3]
  unsigned long long* transport;  // [layer][kTransportZ], two's-complement counts
};

struct Params {
  const unsigned char* tables;  // table blob (TableHeader + arrays)
  int table_bytes;              // multiple of 16
  int stack_cap;                // LDS recoil stack entries per wave
  Geometry geo;
  double tau, cw, invscale, Ef;
  int follow_max_gen, kp;
  // primaries
  long long n_hist, hist0;
  int pka_uniform;
  const int* pka_species;
  const double* pka_E;
  const double* pka_pos;
  const double* pka_dir;
  uint32_t seed;
  // work distribution
  unsigned long long* ctrl;
  Queue q;
  long long admit_limit;        // no new primaries while the queue holds more
  // cascade budget: admitted-but-incomplete cascades times the flights per
  // cascade (f_init until 64 cascades completed, then flights ended per
  // completed cascade, an overestimate) stays below `budget` queue entries;
  // at least c_floor cascades may always be live
  double budget, f_init;
  long long c_floor;
  int* h_out;                   // per history: flights queued or running
  int primary_waves64;          // waves per 64 that feed primaries (0: all waves do everything)
  long long max_ic;             // collisions per flight before kErrRunaway
  unsigned long long max_wait_ticks;  // wall-clock ticks before kErrTimeout
  // tallies
  unsigned long long* counters;   // kNumCounters
  double* wave_energy;            // per wave: 2 * kNumEnergy (sum, compensation)
  double* energy;                 // kNumEnergy, rare events (atomic)
  unsigned long long *range_hist, *lat_y_hist, *lat_z_hist, *vac_hist, *repl_hist;
  double range_x0, range_bw;
  int range_nb;
  double lat_y0, lat_z0, lat_bw;
  int lat_nb;
  double vac_x0, vac_bw, vac_ibw;  // vac_ibw = 1 / vac_bw
  int vac_nb;
  int lds_hist;                   // leading vac/repl depth bins counted in LDS (flushed at exit)
  // per history (index hist - hist0)
  int* h_vac;
  unsigned long long* h_coll;
  int* h_flights;
  double *h_x, *h_y, *h_z;
  int* h_state;
  double* h_Eel;
  // scattering surrogate (FP32 + surrogate mode)
  MlpDev mlp;
  // event records and vacancy map (per-history kernels only)
  LogDev log;
};

// launch bound of the persistent kernel (threads per workgroup)
constexpr int kCascadeMaxBlock = 256;
// resident waves per CU when the tuning leaves it open (two per SIMD)
constexpr int kDefaultWavesPerCu = 8;

// device spans zeroed before a run
constexpr int kMaxSpans = 16;
// the ion species of a run, for mapping primaries to species indices on the
// device (primary_species)
struct SpeciesList {
  int n;
  int Z[kMaxSpecies];
  double m[kMaxSpecies];
};

struct ZeroSpans {
  void* p[kMaxSpans];
  size_t bytes[kMaxSpans];
  int n;
};

// LDS bytes of one wave's recoil stack entry
constexpr int kStackEntryBytes = 7 * 8 + 8 + 4 + 4 + 4 + 4;  // doubles, key, meta, tag, hist, pad

}  // namespace mt
