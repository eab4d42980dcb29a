// Persistent cascade kernel: every lane of every resident wave runs one ion
// or recoil flight (TrimBase::trim, trim.C:35-425) one collision per loop
// iteration; finished lanes refill from the wave's LDS recoil stack, then
// from the HBM queue, then from the primaries (the FIFO drain of
// apps/runmytrim.C:71-93, reordered: every flight's result depends only on
// its own counter-based random stream, not on when it runs).
//
// Work distribution
//   * recoils pushed by a wave go to its LDS stack (LIFO, depth-first); what
//     does not fit goes to one of up to 16 HBM queue shards (structure-of-
//     arrays rings, head/tail words on their own cache lines; a wave pushes
//     to the home shard of its XCD and steals from others when that is empty)
//   * an entry is handed over without cache write-back/invalidate fences:
//     its fields are stored and loaded at agent scope and a ready word
//     (lap << 32 | checksum of the fields) tells a reader that all of them
//     have landed
//   * primaries are claimed in wave batches from a counter under a cascade
//     budget: live cascades x expected flights per cascade stay within half
//     the queue capacity
//   * termination: a wave that finds no work registers idle in a packed
//     (epoch | live | idle) word; every wave leaves once all live waves are
//     idle, every shard is empty and the primaries are used up, observed
//     twice with the same epoch.  Waits are bounded by a wall-clock limit.
//
// Tallies: per-lane 32-bit event counters summed over the wave at exit,
// per-lane compensated FP64 energy sums, histogram bins in LDS (or HBM
// atomics) flushed at exit.  The per-history instantiations (PERHIST) add
// per-history results, the records of the stream-writing TrimBase subclasses
// (trim.C:466-527) and the layer transport of a SampleDynamic drain loop;
// the hot instantiations carry none of that code.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "mt_cascade.h"
#include "mt_geometry.h"
#include "mt_mlp.h"
#include "mt_philox.h"
#include "mt_qsum.h"
#include "mt_physics.h"
#include "mt_scatter32.h"

namespace mt {

namespace {

__device__ __forceinline__ int lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t bcast64(uint64_t v, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t uni64(uint64_t v) {  // lane 0's value, in scalar registers
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// relaxed agent-scope load, made wave-uniform
__device__ __forceinline__ uint64_t ld_relaxed(unsigned long long* p) {
  return uni64(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Neumaier-compensated accumulation: value = s + e
__device__ __forceinline__ void kadd(double& s, double& e, double x) {
  const double t = s + x;
  e += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
  s = t;
}

__device__ __forceinline__ int bin_of(double x, double x0, double bw, int nb) {
  const double t = (x - x0) / bw;
  if (!(t >= 0.0)) return 0;
  if (t >= (double)nb) return nb + 1;
  return 1 + (int)t;
}

// bin_of with the quotient from the stored reciprocal ibw = 1 / bw: d ibw is
// within two ulps of the correctly rounded d / bw, so both have the same floor
// and the same comparisons with 0 and nb unless d ibw lies within a few ulps
// of an integer; there (a lane in some 1e12) the division decides
__device__ __forceinline__ int bin_of_r(double x, double x0, double bw, double ibw, int nb) {
  const double d = x - x0;
  double t = d * ibw;
  if (fabs(t - rint(t)) <= 1e-12 * fmax(fabs(t), 1.0)) t = d / bw;
  if (!(t >= 0.0)) return 0;
  if (t >= (double)nb) return nb + 1;
  return 1 + (int)t;
}

// one flight's state, held in registers
struct Flight {
  double px, py, pz, dx, dy, dz, E;
  int sp, gen, tag;
  uint32_t hist;
  uint64_t key;
  CellCache cc;  // clusters samples: the occupancy cell of the last lookup
  int ic;
  double eel;  // electronic loss of this flight (per-history tally)
};

struct StackView {  // one wave's LDS stack, structure of arrays
  double *px, *py, *pz, *dx, *dy, *dz, *E;
  uint64_t* key;
  uint32_t* meta;
  int* tag;
  uint32_t* hist;
  __device__ void bind(unsigned char* base, int cap) {
    px = (double*)base;
    py = px + cap;
    pz = py + cap;
    dx = pz + cap;
    dy = dx + cap;
    dz = dy + cap;
    E = dz + cap;
    key = (uint64_t*)(E + cap);
    meta = (uint32_t*)(key + cap);
    tag = (int*)(meta + cap);
    hist = (uint32_t*)(tag + cap);
  }
};

__host__ __device__ constexpr int stack_bytes(int cap) { return ((kStackEntryBytes - 4) * cap + 15) & ~15; }

__device__ __forceinline__ void st_agent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


// Kernel arguments are read from the argument segment where they are used
// (scalar loads, which the scalar cache serves) rather than held in scalar
// registers across the loop: the argument block (~150 dwords) does not fit
// beside the loop state in the 102 scalar registers, and every use of a
// spilled one costs a v_readlane, a vector instruction.  This holds for the
// rare paths and for the per-collision constants (seed, cw, tau, boundaries,
// follow rule, bins).  The empty asm makes the pointer opaque, so the loads
// stay at their uses instead of being hoisted out of the loop.
typedef const __attribute__((address_space(4))) Params* ArgPtr;
__device__ __forceinline__ ArgPtr cold_args() {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t k = (uint64_t)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(k));
  return (ArgPtr)k;
#else
  return nullptr;
#endif
}

// Appends one record per lane where `pred` holds (called by the whole wave:
// one counter atomic per wave; records past the capacity are counted, not
// stored)
__device__ __forceinline__ void log_put(LogRecord* rec, unsigned long long cap, unsigned long long* count, bool pred,
                                        const LogRecord& r) {
  const uint64_t m = __ballot(pred);
  if (m == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)m) - 1;
  uint64_t base = 0;
  if (lane == leader) base = atomicAdd(count, (unsigned long long)__popcll(m));
  base = bcast64(base, leader);
  if (pred) {
    const uint64_t i = base + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
    if (i < cap) rec[i] = r;
  }
}

// Free path, impact parameter, target atom, electronic loss and scattering
// of one collision (trim.C:88-272) in precision R.  FP64 follows the
// reference's expressions; FP32 evaluates the same quantities in single
// precision (scatter_f32 for the angle) and hands back doubles.
template <int ARITH, int POT>
__device__ __forceinline__ void physics(const TableView& T, int sp, int k, double E, int ic, uint32_t klo,
                                        uint32_t khi, uint32_t seed, uint32_t hist, const P4& blk, double cw,
                                        const MlpDev& mlp, int& nn, double& ls, double& dee, double& s2,
                                        double& ct, double& st, int& it) {
  const MatConst& mc = T.mat[sp * T.n_mat + k];
  const int eb = sp * T.n_el + mc.el_begin;
  double hh = u32d(blk.z);
  for (nn = 0; nn < mc.n_el; ++nn) {
    hh -= T.el[eb + nn].t;
    if (hh <= 0.0) break;
  }
  if (nn >= mc.n_el) nn = mc.n_el - 1;
  const ElemConst& ec = T.el[eb + nn];
  if constexpr (ARITH == 0) {
    // free path and maximum impact parameter (trim.C:88-95); log E serves
    // eeg^0.1, the stopping fits and the apsis' starting point
    const double lE = m_log(E);
    double eps = E * mc.f;
    const double eeg = m_sqrt(eps * mc.epsdg);
    const double pmax = mc.a * m_rcp(eeg + m_sqrt(eeg) + 0.125 * m_exp(0.05 * (lE + mc.ln_f_epsdg)));
    ls = m_rcp(M_PI * (pmax * pmax) * mc.arho);
    if (ic == 1) {
      const uint64_t f = first_draw(((uint64_t)khi << 32) | klo, seed, hist);
      ls = u53((uint32_t)(f >> 32), (uint32_t)f) * mn(ls, cw);
    }
    // impact parameter and target atom (trim.C:143-160)
    const double p = pmax * m_sqrt(u53(blk.x, blk.y));
    eps = ec.fi * E;
    const double b = p * m_rcp(ec.ai);
    // electronic loss along ls (trim.C:166-170) and scattering (trim.C:172-272)
    dee = ls * getrstop<double>(T, sp, k, E, lE);
    scatter<double, POT>(eps, b, lE + ec.ln_fi, s2, ct, st, it);
  } else {
    const float Ef = (float)E;
    const float lEf = m_log(Ef);
    float eps = Ef * (float)mc.f;
    const float eeg = m_sqrt(eps * (float)mc.epsdg);
    const float pmax = (float)mc.a * m_rcp(eeg + m_sqrt(eeg) + 0.125f * m_exp(0.05f * (lEf + (float)mc.ln_f_epsdg)));
    float lsf = m_rcp(3.14159265358979f * (pmax * pmax) * (float)mc.arho);
    if (ic == 1) {
      const uint64_t f = first_draw(((uint64_t)khi << 32) | klo, seed, hist);
      lsf = u24f((uint32_t)(f >> 32)) * mn(lsf, (float)cw);
    }
    const float p = pmax * m_sqrt(u24f(blk.x));
    eps = (float)ec.fi * Ef;
    const float b = p * m_rcp((float)ec.ai);
    ls = (double)lsf;
    dee = ls * (double)getrstop<float>(T, sp, k, Ef, lEf);
    float s2f, ctf, stf;
    scatter_single<ARITH, POT>(mlp, eps, b, lEf + (float)ec.ln_fi, s2f, ctf, stf, it);
    s2 = s2f;
    ct = ctf;
    st = stf;
  }
}

}  // namespace

// Three waves per SIMD (168 VGPRs): the FP64 and FP32 kernels fit without
// spills (the per-history FP64 variants spill a few dwords); the surrogate
// kernels, whose network needs more registers, keep two.
template <int ARITH, int POT, int GEO, bool PERHIST>
__global__ void __launch_bounds__(kCascadeMaxBlock, (ARITH == 2 ? 2 : 3)) cascade_kernel(Params P) {
  extern __shared__ __align__(16) unsigned char lds[];
  {
    const uint4* src = (const uint4*)P.tables;
    uint4* dst = (uint4*)lds;
    for (int i = threadIdx.x; i < P.table_bytes / 16; i += blockDim.x) dst[i] = src[i];
  }
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nwb = blockDim.x >> 6;
  const int cap = P.stack_cap;
  StackView S;
  S.bind(lds + P.table_bytes + wid * stack_bytes(cap), cap);
  // depth bins [0, lds_hist) of the vacancy and replacement histograms
  uint32_t* lvac = (uint32_t*)(lds + P.table_bytes + nwb * stack_bytes(cap));
  uint32_t* lrepl = lvac + P.lds_hist;
  for (int i = threadIdx.x; i < 2 * P.lds_hist; i += blockDim.x) lvac[i] = 0u;
  __syncthreads();

  const TableHeader* H = (const TableHeader*)lds;
  TableView T;
  T.n_species = H->n_species;
  T.n_mat = H->n_mat;
  T.n_el = H->n_el;
  T.potential = H->potential;
  T.mat = (const MatConst*)(lds + H->off_mat);
  T.el = (const ElemConst*)(lds + H->off_el);
  T.stop = (const StopConst*)(lds + H->off_stop);
  const Geometry& G = P.geo;

  unsigned long long* ctrl = P.ctrl;
  const unsigned long long clk0 = (unsigned long long)clock64();
  const unsigned long long wall0 = (unsigned long long)wall_clock64();
  if (lane == 0) {
    atomicAdd(&ctrl[kState], 1ull << 16);  // live += 1
    // the first wave to start stamps the launch (the words start at zero)
    atomicCAS(&ctrl[kT0], 0ull, wall0);
  }

  // wave-uniform state
  int scount = 0;          // entries in this wave's LDS stack
  bool wave_idle = false;  // registered idle
  int backoff = 1;         // idle sleep rounds
  // home queue shard: waves of one XCD (blocks are dealt to the 8 XCDs in
  // turn) share their shards
  const int per_xcd = P.q.nq >= 8 ? P.q.nq / 8 : 1;
  const int home = (P.q.nq >= 8 ? (int)(blockIdx.x & 7u) * per_xcd + wid % per_xcd
                                : (int)(blockIdx.x * nwb + wid)) & (P.q.nq - 1);
  int victim = home;       // last shard this wave stole from
  // wave roles (P.primary_waves64 > 0): feeder waves admit the primaries in
  // bulk and send every recoil to the HBM queue; the others run what the
  // queue holds and admit primaries only for their idle lanes when it is
  // empty.  A feeder that may not admit (primaries used up, cascade budget
  // reached) works like any other wave until it may again.
  const bool roles = P.primary_waves64 > 0;
  const bool feeder_wave = roles && ((int)((blockIdx.x * nwb + wid) & 63u) < P.primary_waves64);
  bool feeding = feeder_wave;  // wave-uniform: admitted primaries at its last refill
  int poll_skip = 0;       // iterations before the shared queue is polled again
  long long peak_seen = 0; // largest queue fill this wave has reported
  unsigned long long fe_pending = 0;  // ended flights not yet added to kFlightsEnded
  // per lane: outstanding-flight count of the history before this lane's last
  // flight-end decrement, not yet examined (1 means the cascade closed)
  int pend_old = 0;
  bool pend = false;
  auto close_pending = [&]() {
    const uint64_t closed_m = __ballot(pend && pend_old == 1);
    pend = false;
    if (lane == 0 && closed_m) {
      atomicAdd(&ctrl[kCascDone], (unsigned long long)__popcll(closed_m));
      atomicAdd(&ctrl[kLive], (unsigned long long)(-(long long)__popcll(closed_m)));
    }
  };
  unsigned long long idle_since = 0;
  unsigned long long c_qops = 0;
  // per-lane accumulators
  // (32-bit per lane: a lane runs at most a few 1e7 collisions per launch)
  // Event counts are kept per lane (VGPRs) rather than per wave: a wave-wide
  // count of every event in every iteration would hold a dozen 64-bit scalar
  // registers and cost a ballot and a population count each.  Collisions are
  // counted when the flight ends (its collision index), displacements are
  // vacancies + replacements + substitutions.
  unsigned l_newton32 = 0, l_kpvac32 = 0, l_coll32 = 0, l_ruth32 = 0, l_surr32 = 0, l_vac32 = 0;
  unsigned l_flight32 = 0, l_fate32[4] = {0u, 0u, 0u, 0u}, l_vacuum32 = 0, l_rangen32 = 0, l_coll0_32 = 0;
  double eel_s = 0.0, eel_e = 0.0, enuc_s = 0.0, enuc_e = 0.0;
  unsigned err = 0;

  bool active = false;
  Flight F;
  F.ic = 0;

  while (true) {
    // ------------------------------------------------------------ refill
    // Idle lanes take flights from the wave's LDS stack.  When the stack
    // cannot serve them, the wave first tops it up in bulk (up to half its
    // capacity) from the HBM queue and then from the primaries, so the
    // shared control words are touched once per batch, not per flight.
    uint64_t idle_m = __ballot(!active);
    if (idle_m) {
      const int n_idle = __popcll(idle_m);
      if (poll_skip > 0) --poll_skip;
      // a wave whose last poll found nothing waits a few iterations before
      // polling again, unless all its lanes are idle
      if (scount < n_idle && (poll_skip == 0 || n_idle == 64)) {
        const ArgPtr A = cold_args();
        // one lane loads one entry, so a batch holds at most 64
        const int want = min(64, min(cap - scount, max(n_idle, cap / 2) - scount));
        // cascade budget (see Params): cascades that may be admitted now
        auto budget_room = [&]() -> long long {
          const uint64_t live = ld_relaxed(&ctrl[kLive]);
          const uint64_t done = ld_relaxed(&ctrl[kCascDone]);
          const double fhat =
              done >= 64ull ? (double)ld_relaxed(&ctrl[kFlightsEnded]) / (double)done : A->f_init;
          const long long cmax = max((long long)(A->budget / fmax(fhat, 1.0)), A->c_floor);
          return cmax - (long long)live;
        };
        const bool pka_left = ld_relaxed(&ctrl[kPkaNext]) < (unsigned long long)A->n_hist;
        long long n_adm = 0;
        // a feeder wave admits primaries while they last and the budget lets
        // it, leaving the shared queue to the others; otherwise it works like
        // any other wave
        if (feeder_wave) {
          if (pka_left) n_adm = min((long long)want, budget_room());
          feeding = n_adm > 0;
        }
        // the home shard first, then up to three others (rotating victim)
        int qs = home;
        uint64_t h0 = 0, t0 = 0;
        if (!feeding) {
          h0 = ld_relaxed(&ctrl[kQHead(qs)]);
          t0 = ld_relaxed(&ctrl[kQTail(qs)]);
          for (int v = 0; v < 3 && t0 <= h0 && A->q.nq > 1; ++v) {
            victim = (victim + 1) & (A->q.nq - 1);
            if (victim == home) victim = (victim + 1) & (A->q.nq - 1);
            qs = victim;
            h0 = ld_relaxed(&ctrl[kQHead(qs)]);
            t0 = ld_relaxed(&ctrl[kQTail(qs)]);
          }
        }
        const bool may_pop = t0 > h0;
        // a wave that is not feeding admits primaries when the queue has
        // nothing for it; with roles, only as many as it has idle lanes
        if (!feeding && !may_pop && pka_left) {
          const int w = roles ? min(want, n_idle - scount) : want;
          n_adm = min((long long)w, budget_room());
        }
        const bool may_pka = n_adm > 0;
        if (wave_idle && (may_pop || may_pka)) {
          if (lane == 0) atomicAdd(&ctrl[kState], (1ull << 32) - 1ull);  // epoch += 1, idle -= 1
          wave_idle = false;
        }
        if (may_pop) {
          // claim up to `want` queue positions with one CAS
          uint64_t got = 0, base = 0;
          if (lane == 0) {
            uint64_t h = h0, t = t0;
            for (int tries = 0; tries < 4 && t > h; ++tries) {
              const uint64_t k = min((uint64_t)want, t - h);
              unsigned long long expect = h;
              if (__hip_atomic_compare_exchange_strong(&ctrl[kQHead(qs)], &expect, h + k, __ATOMIC_RELAXED,
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                got = k;
                base = h;
                break;
              }
              h = expect;  // the head another wave moved it to
              t = ld_relaxed(&ctrl[kQTail(qs)]);
            }
          }
          got = bcast64(got, 0);
          base = bcast64(base, 0);
          if (got > 0) {
            ++c_qops;
            bool mine = lane < (int)got;
            uint64_t slot = 0;
            if (mine) {
              const uint64_t pos = base + (uint64_t)lane;
              slot = ((uint64_t)qs << A->q.sub_log2) | (pos & A->q.sub_mask);
              const uint32_t lap = (uint32_t)(pos >> A->q.sub_log2) + 1u;
              const unsigned long long t_start = (unsigned long long)wall_clock64();
              const int j = scount + lane;
              // wait for the lap, then for fields that match the checksum;
              // fields go straight to this lane's stack slot
              while (true) {
                const uint64_t rw = ld_agent(&A->q.ready[slot]);
                if ((uint32_t)(rw >> 32) == lap) {
                  QSum h = sum_init(lap);
                  uint64_t w;
                  w = ld_agent(&A->q.px[slot]); S.px[j] = __longlong_as_double((long long)w); h = sum_add(h, w);
                  w = ld_agent(&A->q.py[slot]); S.py[j] = __longlong_as_double((long long)w); h = sum_add(h, w);
                  w = ld_agent(&A->q.pz[slot]); S.pz[j] = __longlong_as_double((long long)w); h = sum_add(h, w);
                  w = ld_agent(&A->q.dx[slot]); S.dx[j] = __longlong_as_double((long long)w); h = sum_add(h, w);
                  w = ld_agent(&A->q.dy[slot]); S.dy[j] = __longlong_as_double((long long)w); h = sum_add(h, w);
                  w = ld_agent(&A->q.dz[slot]); S.dz[j] = __longlong_as_double((long long)w); h = sum_add(h, w);
                  w = ld_agent(&A->q.E[slot]); S.E[j] = __longlong_as_double((long long)w); h = sum_add(h, w);
                  w = ld_agent(&A->q.key[slot]); S.key[j] = w; h = sum_add(h, w);
                  w = ld_agent(&A->q.mt[slot]); h = sum_add(h, w);
                  const uint32_t m = (uint32_t)(w >> 32);
                  S.meta[j] = (int)(m & 0xffu) < T.n_species ? m : 0u;
                  S.tag[j] = (int)(uint32_t)w;
                  const uint32_t wh = ld_agent(&A->q.hist[slot]);
                  S.hist[j] = wh;
                  h = sum_add(h, wh);
                  if (sum_end(h) == (uint32_t)rw) {
                    if ((int)(m & 0xffu) >= T.n_species) err |= kErrBadEntry;
                    // every field has been loaded (the checksum used each):
                    // the slot may take the next lap
                    st_agent(&A->q.ready[slot], consumed_word(lap));
                    break;
                  }
                }
                __builtin_amdgcn_s_sleep(1);
                if ((unsigned long long)wall_clock64() - t_start > A->max_wait_ticks) {
                  err |= kErrTimeout;
                  mine = false;
                  break;
                }
                if (ld_relaxed(&ctrl[kError]) != 0ull) {  // another wave aborted the run
                  mine = false;
                  break;
                }
              }
            }
            const uint64_t bad = __ballot((err & (kErrTimeout | kErrBadEntry)) != 0u);
            if (bad == 0ull) {
              scount += (int)got;
            } else if (lane == __ffsll((long long)bad) - 1) {  // abort the run: every wave leaves
              atomicOr((unsigned int*)&ctrl[kError], err & (kErrTimeout | kErrBadEntry));
            }
          }
        } else if (may_pka) {
          // admit primaries into the stack
          uint64_t first = 0;
          if (lane == 0) first = atomicAdd(&ctrl[kPkaNext], (unsigned long long)n_adm);
          first = bcast64(first, 0);
          const long long got = min(n_adm, max(0ll, A->n_hist - (long long)first));
          if (lane == 0 && got > 0) {
            atomicAdd(&ctrl[kLive], (unsigned long long)got);
            const long long mid = A->n_hist / 2;
            if ((long long)first <= mid && (long long)first + got > mid) {  // the middle primary
              __hip_atomic_store(&ctrl[kFlightsAtHalf], (unsigned long long)ld_relaxed(&ctrl[kFlightsEnded]),
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              __hip_atomic_store(&ctrl[kTHalf], (unsigned long long)wall_clock64(), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
            }
            if ((long long)first + got == A->n_hist) {  // the last primary: the drain starts
              __hip_atomic_store(&ctrl[kFlightsAtAdmit], (unsigned long long)ld_relaxed(&ctrl[kFlightsEnded]),
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              __hip_atomic_store(&ctrl[kTAdmit], (unsigned long long)wall_clock64(), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
            }
          }
          const long long jh = (long long)first + lane;
          if (lane < got) {
            __hip_atomic_store(&A->h_out[jh], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const long long e = A->pka_uniform ? 0 : jh;
            const int j = scount + lane;
            S.px[j] = A->pka_pos[3 * e + 0];
            S.py[j] = A->pka_pos[3 * e + 1];
            S.pz[j] = A->pka_pos[3 * e + 2];
            S.dx[j] = A->pka_dir[3 * e + 0];
            S.dy[j] = A->pka_dir[3 * e + 1];
            S.dz[j] = A->pka_dir[3 * e + 2];
            S.E[j] = A->pka_E[e];
            S.key[j] = 0ull;
            S.meta[j] = (uint32_t)A->pka_species[e];
            S.tag[j] = -1;
            S.hist[j] = (uint32_t)(A->hist0 + jh);
          }
          scount += (int)max(got, 0ll);
        }
        if (!may_pop && !may_pka) poll_skip = 4;
        __builtin_amdgcn_wave_barrier();
      }
      const int take = min(n_idle, scount);
      if (take > 0) {
        const int r = lane_rank(idle_m);
        if (!active && r < take) {
          const int j = scount - 1 - r;
          F.px = S.px[j];
          F.py = S.py[j];
          F.pz = S.pz[j];
          F.dx = S.dx[j];
          F.dy = S.dy[j];
          F.dz = S.dz[j];
          F.E = S.E[j];
          F.key = S.key[j];
          const uint32_t m = S.meta[j];
          F.sp = (int)(m & 0xffu);
          F.gen = (int)(m >> 8);
          F.tag = S.tag[j];
          F.hist = S.hist[j];
          F.ic = 0;
          F.eel = 0.0;
          F.cc.cell = kNoCell;
          // every index a flight uses is checked here, where all flights start
          if (F.sp < T.n_species && F.hist - (uint32_t)P.hist0 < (uint32_t)P.n_hist)
            active = true;
          else
            err |= kErrBadEntry;
        }
        scount -= take;
        const uint64_t bad = __ballot((err & kErrBadEntry) != 0u);
        if (bad && lane == __ffsll((long long)bad) - 1) atomicOr((unsigned int*)&ctrl[kError], (unsigned)kErrBadEntry);
      }
      if (__ballot(active) != 0ull) backoff = 1;
      if (__ballot(active) == 0ull) {
        // nothing to run: idle protocol and termination test
        const ArgPtr A = cold_args();
        if (!wave_idle) {
          if (lane == 0) atomicAdd(&ctrl[kState], 1ull);
          wave_idle = true;
          idle_since = (unsigned long long)wall_clock64();
        }
        const uint64_t s1 = ld_relaxed(&ctrl[kState]);
        if ((s1 & 0xffffull) == ((s1 >> 16) & 0xffffull)) {
          const uint64_t pn = ld_relaxed(&ctrl[kPkaNext]);
          bool empty = true;
          for (int q = 0; q < A->q.nq; ++q) empty = empty && ld_relaxed(&ctrl[kQHead(q)]) == ld_relaxed(&ctrl[kQTail(q)]);
          const uint64_t s2 = ld_relaxed(&ctrl[kState]);
          if (s1 == s2 && empty && pn >= (unsigned long long)A->n_hist) break;
        }
        if ((unsigned long long)wall_clock64() - idle_since > A->max_wait_ticks) {
          err |= kErrTimeout;
          break;
        }
        if (ld_relaxed(&ctrl[kError]) != 0ull) break;  // a wave aborted the run
        close_pending();
        // back off (up to ~30 us) so idle waves do not crowd the control words
        for (int z = 0; z < backoff; ++z) __builtin_amdgcn_s_sleep(127);
        backoff = min(2 * backoff, 8);
        continue;
      }
    }

    // ------------------------------------------------------------ one collision
    bool ended = false, vacuum = false, push = false, ruth = false, surr = false;
    int state = 0;
    double rx = 0.0, ry = 0.0, rz = 0.0, rdx = 0.0, rdy = 0.0, rdz = 0.0, rE = 0.0;
    int rsp = 0, rtag = -1;
    // event-record inputs (per-history kernels): recoil above threshold and
    // not lost, followed, its element (Z, m, Elbind)
    bool lg_above = false, lg_follow = false;
    int lg_rZ = 0;
    double lg_rm = 0.0, lg_lb = 0.0;
    if (active) {
      ++F.ic;
      int mtag;
      const double pos[3] = {F.px, F.py, F.pz};
      const int k = lookup_material<GEO>(G, pos, mtag, F.cc);
      if constexpr (PERHIST && GEO == 1) {
        // layer transport: a recoil's atom leaves the layer its flight starts
        // in (apps/mytrim_dynamic.C:160-164)
        if (F.ic == 1 && F.gen > 0) {
          const ArgPtr A = cold_args();
          if (A->log.mask & kLogTransport)
            atomicAdd(&A->log.transport[k * kTransportZ + H->species[F.sp].Z], ~0ull);
        }
      }
      if (k < 0) {
        vacuum = true;
        ended = true;
      } else {
        {  // v_norm (functions.h:66-70)
          const double inv = fx_rsqrt(F.dx * F.dx + F.dy * F.dy + F.dz * F.dz);
          F.dx *= inv;
          F.dy *= inv;
          F.dz *= inv;
        }
        const uint32_t klo = (uint32_t)F.key, khi = (uint32_t)(F.key >> 32);
        const ArgPtr C0 = cold_args();
        const uint32_t seed = C0->seed;
        const P4 blk = philox10((uint32_t)F.ic, klo, khi, 0u, seed, F.hist);
        const MatConst& mc = T.mat[F.sp * T.n_mat + k];
        double ls, dee, s2, ct, st;
        int it, nn;
        physics<ARITH, POT>(T, F.sp, k, F.E, F.ic, klo, khi, seed, F.hist, blk, C0->cw, P.mlp, nn, ls, dee, s2, ct,
                            st, it);
        const ElemConst& ec = T.el[F.sp * T.n_el + mc.el_begin + nn];
        ruth = (it == 0);
        surr = (it < 0);
        l_newton32 += it > 0 ? (unsigned)it : 0u;
        l_ruth32 += ruth ? 1u : 0u;
        if constexpr (ARITH == 2) l_surr32 += surr ? 1u : 0u;
        if (it >= kMaxNewton) err |= kErrNewton;
        // energy transfer and momenta (trim.C:275-297)
        double den = ec.ec * s2 * F.E;
        if (dee > F.E) dee = F.E;
        F.E -= dee;
        kadd(eel_s, eel_e, dee);
        if (PERHIST) F.eel += dee;
        // The momenta and the deflection run in RD: FP64 on the FP64 path,
        // FP32 on the single-precision paths (positions and energies stay FP64)
        using RD = typename std::conditional<ARITH == 0, double, float>::type;
        const RD m1 = RD(H->species[F.sp].m);
        const RD p1 = m_sqrt(RD(2.0) * m1 * RD(F.E));
        if (den > F.E) den = F.E;
        F.E -= den;
        const RD p2 = m_sqrt(RD(2.0) * m1 * RD(F.E));
        // recoil at the collision site, ion moves on (trim.C:306-318)
        rx = F.px;
        ry = F.py;
        rz = F.pz;
        const ArgPtr C = cold_args();
        const double step = (ls - C->tau) * C->invscale;
        F.px += F.dx * step;
        F.py += F.dy * step;
        F.pz += F.dz * step;
        const RD dx = RD(F.dx), dy = RD(F.dy), dz = RD(F.dz);
        RD rdxr = dx * p1, rdyr = dy * p1, rdzr = dz * p1;
        rE = den - ec.Elbind;
        // azimuth about the flight direction (replaces trim.C:320-333)
        // u: unit vector perpendicular to d in the plane of its two largest
        // components (x smallest: (0, dz, -dy); y: (-dz, 0, dx); z: (dy, -dx, 0)),
        // chosen by selects so that a wave evaluates one reciprocal square root
        RD ux, uy, uz;
        {
          const RD ax = m_abs(dx), ay = m_abs(dy), az = m_abs(dz);
          const bool xs = ax <= ay && ax <= az, ys = !xs && ay <= az;
          const RD c0 = xs ? dy : dx, c1 = (xs || ys) ? dz : dy;
          const RD n = m_rsqrt(c0 * c0 + c1 * c1);
          ux = xs ? RD(0.0) : (ys ? -dz : dy) * n;
          uy = xs ? dz * n : (ys ? RD(0.0) : -dx * n);
          uz = xs ? -dy * n : (ys ? dx * n : RD(0.0));
        }
        const RD vx = dy * uz - dz * uy;
        const RD vy = dz * ux - dx * uz;
        const RD vz = dx * uy - dy * ux;
        RD sphi, cphi;
        m_sincos2pi(u32d(blk.w), sphi, cphi);
        const RD qx = cphi * ux + sphi * vx;
        const RD qy = cphi * uy + sphi * vy;
        const RD qz = cphi * uz + sphi * vz;
        // new direction (trim.C:336-341): psi = atan2(st, ct + my), so
        // cos psi and sin psi are the normalised arguments (st >= 0)
        const RD xpsi = RD(ct) + RD(ec.my);
        const RD ipsi = m_rsqrt(xpsi * xpsi + RD(st) * RD(st));
        const RD cpsi = xpsi * ipsi;
        const RD spsi = RD(st) * ipsi;
        const RD ndx = dx * cpsi + qx * spsi;
        const RD ndy = dy * cpsi + qy * spsi;
        const RD ndz = dz * cpsi + qz * spsi;
        F.dx = ndx;
        F.dy = ndy;
        F.dz = ndz;
        rdxr -= ndx * p2;
        rdyr -= ndy * p2;
        rdzr -= ndz * p2;
        rdx = rdxr;
        rdy = rdyr;
        rdz = rdzr;
        // boundaries (trim.C:344-352)
        bool lost = false;
        if (C->geo.bc[0] == 2 && (F.px > C->geo.w[0] || F.px < 0.0)) lost = true;
        if (C->geo.bc[1] == 2 && (F.py > C->geo.w[1] || F.py < 0.0)) lost = true;
        if (C->geo.bc[2] == 2 && (F.pz > C->geo.w[2] || F.pz < 0.0)) lost = true;
        if (PERHIST) {
          lg_rZ = ec.Z;
          lg_rm = H->species[ec.recoil_species].m;
          lg_lb = ec.Elbind;
          lg_above = !lost && rE > ec.Edisp - ec.Elbind;
          lg_follow = lg_above && (P.follow_max_gen < 0 || F.gen + 1 < P.follow_max_gen);
        }
        // fate of recoil and ion (trim.C:357-411)
        if (lost) {
          state = 4;
          atomicAdd(&cold_args()->energy[eElost], F.E);
          atomicAdd(&cold_args()->energy[eEdrop], den);
        } else if (rE > ec.Edisp - ec.Elbind) {
          const bool follow = C->follow_max_gen < 0 || F.gen + 1 < C->follow_max_gen;
          if (follow) {
            push = true;
            kadd(enuc_s, enuc_e, ec.Elbind);
            const double inv = fx_rsqrt(rdx * rdx + rdy * rdy + rdz * rdz);
            rdx *= inv;
            rdy *= inv;
            rdz *= inv;
            rsp = ec.recoil_species;
            rtag = mtag;
          } else {
            atomicAdd(&cold_args()->energy[eEunf], den);
          }
          // a vacancy or a replacement at the recoil's site: one depth bin,
          // one count in the vacancy or the replacement histogram
          const bool vac = F.E > ec.Edisp;
          {
            const int bn = bin_of_r(rx, C->vac_x0, C->vac_bw, C->vac_ibw, C->vac_nb);
            if (bn < C->lds_hist)
              atomicAdd(&(vac ? lvac : lrepl)[bn], 1u);
            else
              atomicAdd(&(vac ? cold_args()->vac_hist : cold_args()->repl_hist)[bn], 1ull);
          }
          if (vac) {
            // vacancyCreation (trim.C:439-443, TrimVacCount.C:31-42, KP trim.C:445-464)
            int extra = 0;
            if (C->kp && F.gen + 1 == C->follow_max_gen) {
              const double ed = 0.0115 * m_pow(mc.az, -7.0 / 3.0) * rE;
              const double g = 3.4008 * m_pow(ed, 1.0 / 6.0) + 0.40244 * m_pow(ed, 3.0 / 4.0) + ed;
              const double kd = 0.1337 * m_pow(mc.az, 2.0 / 3.0) / m_sqrt(mc.am);
              const double Ev = rE / (1.0 + kd * g);
              extra = (int)(0.8 * Ev / (2.0 * ec.Edisp));
              l_kpvac32 += (unsigned)extra;
            }
            if (PERHIST) atomicAdd(&cold_args()->h_vac[F.hist - (uint32_t)P.hist0], 1 + extra);
            state = -1;  // vacancy marker, ion keeps moving
          } else {
            // replacementCollision (trim.C:388-395)
            state = (H->species[F.sp].Z == ec.Z) ? 1 : 2;
            kadd(enuc_s, enuc_e, F.E);
          }
        } else {
          kadd(enuc_s, enuc_e, rE + ec.Elbind);
          if (F.E < C->Ef) {
            state = 3;
            kadd(enuc_s, enuc_e, F.E);
          }
        }
        ended = state > 0;
        if (F.ic >= P.max_ic && !ended) {
          err |= kErrRunaway;
          ended = true;
        }
        // a NaN or infinite energy (a corrupt table row, an overflow) fails
        // every threshold above and would keep the flight moving until
        // max_ic; a NaN free path leaves the energy finite but the position
        // not.  x - x is 0 exactly for finite x.  The flight ends and the run
        // reports it (the reference's only such check is the warning of
        // trim.C:277-285).
        if (!(F.E - F.E == 0.0) || !(step - step == 0.0)) {
          err |= kErrNonFinite;
          ended = true;
          push = false;
        }
      }
    }
    l_vac32 += (active && state == -1) ? 1u : 0u;

    // ------------------------------------------------------------ event records
    // the lines TrimHistory, TrimPhononOut and TrimDefectLog write for this
    // collision, in trim.C order: followRecoil (history, spawn), vacancy,
    // dissipation, ion at rest; and TrimVacMap's cell (trim.C:479-501)
    if constexpr (PERHIST) {
      const ArgPtr A = cold_args();
      const int lmask = A->log.mask;
      if (lmask) {
        LogRecord* const lrec = A->log.rec;
        const unsigned long long lcap = A->log.cap;
        unsigned long long* lc = &ctrl[kLogCount];
        const bool coll = active && !vacuum;
        LogRecord ion, rec;
        ion.pos[0] = F.px;
        ion.pos[1] = F.py;
        ion.pos[2] = F.pz;
        ion.E = F.E;
        ion.value = 0.0;
        ion.m = active ? H->species[F.sp].m : 0.0;
        ion.id = F.key;
        ion.hist = F.hist;
        ion.Z = active ? H->species[F.sp].Z : 0;
        ion.gen = F.gen;
        ion.tag = F.tag;
        ion.pad = 0;
        rec.pos[0] = rx;
        rec.pos[1] = ry;
        rec.pos[2] = rz;
        rec.E = rE;
        rec.m = lg_rm;
        rec.id = 0ull;
        rec.hist = F.hist;
        rec.Z = lg_rZ;
        rec.gen = F.gen + 1;
        rec.tag = -1;
        rec.pad = 0;
        if (lmask & kLogHistory) {
          ion.kind = 5;
          log_put(lrec, lcap, lc, coll && lg_above, ion);
        }
        if (lmask & kLogPhonons) {
          rec.kind = 6;
          rec.value = lg_lb;
          log_put(lrec, lcap, lc, coll && lg_follow, rec);
        }
        if (lmask & kLogDefects) {
          rec.kind = 1;
          rec.value = 0.0;
          rec.id = lg_follow ? child_key(F.key, (uint32_t)F.ic) : 0ull;
          rec.tag = lg_follow ? rtag : -1;
          log_put(lrec, lcap, lc, coll && state == -1, rec);
          rec.id = 0ull;
          rec.tag = -1;
        }
        if ((lmask & kLogVacmap) && coll && state == -1) {
          int x = (int)((rx * kVacMapN) / G.w[0]);
          int y = (int)((ry * kVacMapN) / G.w[1]);
          x -= (x / kVacMapN) * kVacMapN;
          y -= (y / kVacMapN) * kVacMapN;
          const int c = lg_rZ == A->log.vz[0] ? 0 : lg_rZ == A->log.vz[1] ? 1 : lg_rZ == A->log.vz[2] ? 2 : -1;
          // the reference indexes out of its map for x or y < 0; those are skipped
          if (c >= 0 && x >= 0 && y >= 0) atomicAdd(&A->log.vacmap[(x * kVacMapN + y) * 3 + c], 1);
        }
        if (lmask & kLogPhonons) {
          rec.kind = 7;
          rec.value = rE + lg_lb;
          log_put(lrec, lcap, lc, coll && !lg_above && state != 4, rec);
        }
        const bool rest = coll && state >= 1 && state <= 3;
        if (lmask & kLogDefects) {
          ion.kind = state == 3 ? 2 : state == 2 ? 3 : 4;
          log_put(lrec, lcap, lc, rest, ion);
        }
        if (lmask & kLogPhonons) {
          ion.kind = 8;
          ion.value = F.E;
          log_put(lrec, lcap, lc, rest, ion);
        }
      }
    }

    // ------------------------------------------------------------ push recoils
    const uint64_t push_m = __ballot(push);
    if (push_m) {
      const int n_push = __popcll(push_m);
      // a feeding wave keeps no recoils: they go to the waves that run recoils
      const int room = feeding ? 0 : cap - scount;
      const int r = lane_rank(push_m);
      const bool over = push && r >= room;  // goes to the HBM queue
      const uint32_t meta = (uint32_t)rsp | ((uint32_t)(F.gen + 1) << 8);
      const uint64_t rkey = child_key(F.key, (uint32_t)F.ic);
      int inc_old = 1;
      if (push) {
        const ArgPtr C1 = cold_args();
        int* ho = &C1->h_out[F.hist - (uint32_t)C1->hist0];
        if (over)  // returning: performed before the entry is published (below)
          inc_old = __hip_atomic_fetch_add(ho, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else       // the entry stays with this wave, which ends it after this
          atomicAdd(ho, 1);
      }
      if (push && r < room) {
        const int j = scount + r;
        S.px[j] = rx;
        S.py[j] = ry;
        S.pz[j] = rz;
        S.dx[j] = rdx;
        S.dy[j] = rdy;
        S.dz[j] = rdz;
        S.E[j] = rE;
        S.key[j] = rkey;
        S.meta[j] = meta;
        S.tag[j] = rtag;
        S.hist[j] = F.hist;
      }
      const int n_over = n_push - min(n_push, room);
      scount += min(n_push, room);
      if (n_over > 0) {
        const ArgPtr A = cold_args();
        ++c_qops;
        // the home shard, unless it is over half full: then the least full
        // of it and three others, so that no shard runs out of slots while
        // others have room (a shard's fill is bounded by its own capacity,
        // whatever the global cascade budget)
        const long long half = (long long)(A->q.sub_mask >> 1);
        int qs = home;
        long long qfill = (long long)(ld_relaxed(&ctrl[kQTail(qs)]) - ld_relaxed(&ctrl[kQHead(qs)]));
        for (int v = 1; v < 4 && qfill > half && A->q.nq > 1; ++v) {
          const int s2 = (home + v * (A->q.nq >> 2 ? A->q.nq >> 2 : 1)) & (A->q.nq - 1);
          const long long f2 = (long long)(ld_relaxed(&ctrl[kQTail(s2)]) - ld_relaxed(&ctrl[kQHead(s2)]));
          if (f2 < qfill) {
            qfill = f2;
            qs = s2;
          }
        }
        uint64_t base = 0, head = 0;
        if (lane == 0) {
          base = atomicAdd(&ctrl[kQTail(qs)], (unsigned long long)n_over);
          head = ld_relaxed(&ctrl[kQHead(qs)]);
        }
        base = bcast64(base, 0);
        head = bcast64(head, 0);
        // occupancy (a diagnostic; consumers may already have claimed
        // reserved slots, so head can be past base)
        const long long fill = (long long)(base + (uint64_t)n_over) - (long long)head;
        if (fill > peak_seen) {  // this wave's own maximum gates the shared one
          peak_seen = fill;
          if (lane == 0) atomicMax(&ctrl[kQueuePeak], (unsigned long long)fill);
        }
        // a slot takes lap L only once its lap L - 1 entry has been read out
        // (its ready word is that lap's consumed marker, or zero for lap 1);
        // otherwise the ring is full there: abort, every wave leaves and the
        // caller reports queue-full
        bool full = false;
        uint64_t slot = 0;
        uint32_t lap = 0;
        if (over) {
          const uint64_t pos = base + (uint64_t)(r - room);
          slot = ((uint64_t)qs << A->q.sub_log2) | (pos & A->q.sub_mask);
          lap = (uint32_t)(pos >> A->q.sub_log2) + 1u;
          const uint64_t prev = ld_agent(&A->q.ready[slot]);
          full = prev != (lap == 1u ? 0ull : consumed_word(lap - 1u));
        }
        const uint64_t full_m = __ballot(full);
        if (full_m) {
          err |= kErrQueueFull;
          if (lane == __ffsll((long long)full_m) - 1) atomicOr((unsigned int*)&ctrl[kError], (unsigned)kErrQueueFull);
        }
        if (over && !full) {
          QSum h = sum_init(lap);
          uint64_t w;
          w = (uint64_t)__double_as_longlong(rx); st_agent(&A->q.px[slot], w); h = sum_add(h, w);
          w = (uint64_t)__double_as_longlong(ry); st_agent(&A->q.py[slot], w); h = sum_add(h, w);
          w = (uint64_t)__double_as_longlong(rz); st_agent(&A->q.pz[slot], w); h = sum_add(h, w);
          w = (uint64_t)__double_as_longlong(rdx); st_agent(&A->q.dx[slot], w); h = sum_add(h, w);
          w = (uint64_t)__double_as_longlong(rdy); st_agent(&A->q.dy[slot], w); h = sum_add(h, w);
          w = (uint64_t)__double_as_longlong(rdz); st_agent(&A->q.dz[slot], w); h = sum_add(h, w);
          w = (uint64_t)__double_as_longlong(rE); st_agent(&A->q.E[slot], w); h = sum_add(h, w);
          st_agent(&A->q.key[slot], rkey);
          h = sum_add(h, rkey);
          w = ((uint64_t)meta << 32) | (uint32_t)rtag;
          st_agent(&A->q.mt[slot], w);
          h = sum_add(h, w);
          st_agent(&A->q.hist[slot], F.hist);
          const uint32_t sum = sum_end(sum_add(h, F.hist));
          // A reader on another XCD may end this flight as soon as it sees
          // the ready word, so the history's flight-count increment must be
          // performed first: the ready word depends on the returned count,
          // which makes the store wait for the atomic.  A count below 1
          // cannot happen (the pushing flight is outstanding); it would mean
          // corrupt bookkeeping, and the run is aborted.  The fields need no
          // ordering: the reader checks the checksum.
          const bool sane = inc_old >= 1;
          if (!sane) err |= kErrBadEntry;
          st_agent(&A->q.ready[slot], ((uint64_t)(sane ? lap : 0u) << 32) | sum);
        }
        const uint64_t insane = __ballot((err & kErrBadEntry) != 0u);
        if (insane && lane == __ffsll((long long)insane) - 1)
          atomicOr((unsigned int*)&ctrl[kError], (unsigned)kErrBadEntry);
      }
    }

    // ------------------------------------------------------------ flight ends
    const bool fin = active && (ended || vacuum);
    const uint64_t fin_m = __ballot(fin);
    if (fin_m) {
      const uint64_t nf_m = __ballot((err & kErrNonFinite) != 0u);
      if (nf_m && lane == __ffsll((long long)nf_m) - 1)  // every wave leaves
        atomicOr((unsigned int*)&ctrl[kError], (unsigned)kErrNonFinite);
      // the decrement's old value is read one iteration later (close_pending)
      // so its latency overlaps the next collision
      close_pending();
      if (fin) {
        const ArgPtr C2 = cold_args();
        pend_old = atomicAdd(&C2->h_out[F.hist - (uint32_t)C2->hist0], -1);
        pend = true;
      }
      fe_pending += __popcll(fin_m);
      if (lane == 0 && fe_pending >= 256ull) atomicAdd(&ctrl[kFlightsEnded], fe_pending);
      if (fin) {
        ++l_flight32;
        l_coll32 += (unsigned)(vacuum ? F.ic - 1 : F.ic);
        l_coll0_32 += F.gen == 0 ? (unsigned)(vacuum ? F.ic - 1 : F.ic) : 0u;
        l_vacuum32 += vacuum ? 1u : 0u;
#pragma unroll
        for (int f = 0; f < 4; ++f) l_fate32[f] += (state == f + 1) ? 1u : 0u;
        l_rangen32 += (F.gen == 0 && !vacuum && state != 4) ? 1u : 0u;
        if (vacuum) atomicAdd(&cold_args()->energy[eEvac], F.E);
        const uint32_t hl = F.hist - (uint32_t)P.hist0;
        if (PERHIST) {
          const ArgPtr A = cold_args();
          atomicAdd(&A->h_coll[hl], (unsigned long long)(vacuum ? F.ic - 1 : F.ic));
          atomicAdd(&A->h_flights[hl], 1);
          atomicAdd(&A->h_Eel[hl], F.eel);
          if constexpr (GEO == 1) {
            // layer transport: the atom joins the layer its flight ends in
            // when it stopped at x > 0 (apps/mytrim_dynamic.C:174-178)
            if ((A->log.mask & kLogTransport) && F.px > 0.0) {
              int t;
              const double pos[3] = {F.px, F.py, F.pz};
              const int k = lookup_material<GEO>(G, pos, t);
              atomicAdd(&A->log.transport[k * kTransportZ + H->species[F.sp].Z], 1ull);
            }
          }
        }
        if (F.gen == 0) {
          const ArgPtr A = cold_args();
          if (PERHIST) {
            A->h_x[hl] = F.px;
            A->h_y[hl] = F.py;
            A->h_z[hl] = F.pz;
            A->h_state[hl] = vacuum ? -1 : (state > 0 ? state : 0);
          }
          if (!vacuum && state != 4) {
            atomicAdd(&A->range_hist[bin_of(F.px, A->range_x0, A->range_bw, A->range_nb)], 1ull);
            const double lo = -0.5 * A->lat_nb * A->lat_bw;
            atomicAdd(&A->lat_y_hist[bin_of(F.py - A->lat_y0, lo, A->lat_bw, A->lat_nb)], 1ull);
            atomicAdd(&A->lat_z_hist[bin_of(F.pz - A->lat_z0, lo, A->lat_bw, A->lat_nb)], 1ull);
            atomicAdd(&A->energy[eRangeSum], F.px);
            atomicAdd(&A->energy[eRangeSum2], F.px * F.px);
          }
        }
        active = false;
      }
      if (fe_pending >= 256ull) fe_pending = 0;
    }
  }

  // ------------------------------------------------------------ flush
  const ArgPtr A = cold_args();
  // wave sums of the per-lane counters (each 64-bit; lane sums are 32-bit)
  auto wsum = [&](unsigned v) {
    unsigned long long t = v;
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    return t;
  };
  const unsigned long long n_coll = wsum(l_coll32), n_flight = wsum(l_flight32), n_vac = wsum(l_vac32),
                           n_repl = wsum(l_fate32[0]), n_subst = wsum(l_fate32[1]), n_inter = wsum(l_fate32[2]),
                           n_lost = wsum(l_fate32[3]), n_vacuum = wsum(l_vacuum32), n_ruth = wsum(l_ruth32),
                           n_newton = wsum(l_newton32), n_kpvac = wsum(l_kpvac32), n_rangen = wsum(l_rangen32),
                           n_surr = wsum(l_surr32), n_coll0 = wsum(l_coll0_32);
  for (int off = 32; off > 0; off >>= 1) {
    const double s_a = __shfl_xor(eel_s, off), e_a = __shfl_xor(eel_e, off);
    kadd(eel_s, eel_e, s_a);
    eel_e += e_a;
    const double s_b = __shfl_xor(enuc_s, off), e_b = __shfl_xor(enuc_e, off);
    kadd(enuc_s, enuc_e, s_b);
    enuc_e += e_b;
  }
  if (lane == 0) {
    unsigned long long* C = A->counters;
    atomicAdd(&C[cColl], n_coll);
    atomicAdd(&C[cFlight], n_flight);
    atomicAdd(&C[cVac], n_vac + n_kpvac);
    atomicAdd(&C[cRepl], n_repl);
    atomicAdd(&C[cSubst], n_subst);
    atomicAdd(&C[cInter], n_inter);
    atomicAdd(&C[cLost], n_lost);
    atomicAdd(&C[cVacuum], n_vacuum);
    atomicAdd(&C[cRuth], n_ruth);
    atomicAdd(&C[cNewton], n_newton);
    atomicAdd(&C[cDisp], n_vac + n_repl + n_subst);
    atomicAdd(&C[cRangeN], n_rangen);
    atomicAdd(&C[cSurr], n_surr);
    atomicAdd(&C[cCollPrimary], n_coll0);
    atomicAdd(&ctrl[kQueueOps], c_qops);
    if (fe_pending) atomicAdd(&ctrl[kFlightsEnded], fe_pending);
    atomicAdd(&ctrl[kClkSum], (unsigned long long)clock64() - clk0);
    atomicAdd(&ctrl[kWallSum], (unsigned long long)wall_clock64() - wall0);
    const int gw = blockIdx.x * nwb + wid;
    double* we = A->wave_energy + (size_t)gw * (2 * kNumEnergy);
    we[2 * eEel] = eel_s;
    we[2 * eEel + 1] = eel_e;
    we[2 * eEnuc] = enuc_s;
    we[2 * eEnuc + 1] = enuc_e;
  }
  if (err) atomicOr((unsigned int*)&ctrl[kError], err);
  if (P.lds_hist) {
    __syncthreads();
    for (int i = threadIdx.x; i < P.lds_hist; i += blockDim.x) {
      if (lvac[i]) atomicAdd(&A->vac_hist[i], (unsigned long long)lvac[i]);
      if (lrepl[i]) atomicAdd(&A->repl_hist[i], (unsigned long long)lrepl[i]);
    }
  }
}

int cascade_stack_bytes(int cap) { return stack_bytes(cap); }

// ------------------------------------------------------------ reset and flush
// Zero the spans of a run (queue ready words, tallies, control words): a
// kernel rather than memset nodes, so the whole sequence can be replayed from
// a graph.  Spans are hipMalloc'd (256-byte aligned) with sizes in bytes that
// are multiples of 4.
__global__ void __launch_bounds__(256) zero_spans(ZeroSpans z) {
  const size_t gid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (int s = 0; s < z.n; ++s) {
    const size_t words = z.bytes[s] >> 2;
    uint4* p4 = (uint4*)z.p[s];
    const size_t n4 = words >> 2;
    for (size_t i = gid; i < n4; i += stride) p4[i] = make_uint4(0u, 0u, 0u, 0u);
    uint32_t* p1 = (uint32_t*)z.p[s];
    for (size_t i = 4 * n4 + gid; i < words; i += stride) p1[i] = 0u;
  }
}

// Compensated (Neumaier) sum of the per-wave energy partials: lane l sums
// waves l, l + 64, ...; lane 0 then folds the 64 partial pairs in lane order.
// out: Eel sum, Eel compensation, Enuc sum, Enuc compensation.
__device__ __forceinline__ void neumaier(double& s, double& c, double x) {
  const double t = s + x;
  c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
  s = t;
}

__global__ void __launch_bounds__(64) flush_energy(const double* we, int waves, double* out) {
  __shared__ double part[4][64];
  const int l = threadIdx.x;
  double s0 = 0.0, c0 = 0.0, s1 = 0.0, c1 = 0.0;
  for (int w = l; w < waves; w += 64) {
    const double* p = we + (size_t)w * 2 * kNumEnergy;
    neumaier(s0, c0, p[0]);
    neumaier(s0, c0, p[1]);
    neumaier(s1, c1, p[2]);
    neumaier(s1, c1, p[3]);
  }
  part[0][l] = s0;
  part[1][l] = c0;
  part[2][l] = s1;
  part[3][l] = c1;
  __syncthreads();
  if (l == 0) {
    double a0 = 0.0, e0 = 0.0, a1 = 0.0, e1 = 0.0;
    for (int i = 0; i < 64; ++i) {
      neumaier(a0, e0, part[0][i]);
      e0 += part[1][i];
      neumaier(a1, e1, part[2][i]);
      e1 += part[3][i];
    }
    out[0] = a0;
    out[1] = e0;
    out[2] = a1;
    out[3] = e1;
  }
}

}  // namespace mt

// ---------------------------------------------------------------- launchers
namespace mt {

typedef void (*CascadeKernel)(Params);

template <int ARITH, int POT, int GEO>
static CascadeKernel pick3(bool perhist) {
  return perhist ? cascade_kernel<ARITH, POT, GEO, true> : cascade_kernel<ARITH, POT, GEO, false>;
}

template <int ARITH, int POT>
static CascadeKernel pick2(int geo, bool perhist) {
  if (geo == 1) return pick3<ARITH, POT, 1>(perhist);
  if (geo == 2) return pick3<ARITH, POT, 2>(perhist);
  return pick3<ARITH, POT, 0>(perhist);
}

template <int ARITH>
static CascadeKernel pick1(int potential, int geo, bool perhist) {
  if (potential == 1) return pick2<ARITH, 1>(geo, perhist);
  if (potential == 2) return pick2<ARITH, 2>(geo, perhist);
  return pick2<ARITH, 0>(geo, perhist);
}

// the surrogate is trained on the universal potential only
static CascadeKernel pick(int arith, int potential, int geo, bool perhist) {
  if (arith == 2) return pick2<2, 0>(geo, perhist);
  return arith == 1 ? pick1<1>(potential, geo, perhist) : pick1<0>(potential, geo, perhist);
}

hipError_t cascade_occupancy(int arith, int potential, int geo, bool perhist, int block, size_t lds,
                             int* blocks_per_cu) {
  CascadeKernel k = pick(arith, potential, geo, perhist);
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)k, block, lds);
}

hipError_t cascade_launch(const Params& P, int arith, int potential, bool perhist, int grid, int block, size_t lds,
                          hipStream_t stream) {
  CascadeKernel k = pick(arith, potential, P.geo.kind, perhist);
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds, stream, P);
  return hipGetLastError();
}

const void* cascade_function(int arith, int potential, int geo, bool perhist) {
  return (const void*)pick(arith, potential, geo, perhist);
}

int zero_spans_grid(const ZeroSpans& z) {
  size_t most = 0;
  for (int s = 0; s < z.n; ++s) most = z.bytes[s] > most ? z.bytes[s] : most;
  const size_t blocks = (most / 16 + 255) / 256;
  return (int)(blocks < 1 ? 1 : blocks > 4096 ? 4096 : blocks);
}

const void* zero_spans_function() { return (const void*)zero_spans; }
const void* flush_energy_function() { return (const void*)flush_energy; }

// Species index of each primary: the first entry of the run's species list
// with its (Z, m); -1 where none matches (the cascade kernel rejects such a
// flight where it starts).
__global__ void __launch_bounds__(256) primary_species(SpeciesList L, const int* Z, const double* m, long long n,
                                                       int* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    int s = -1;
    for (int k = L.n - 1; k >= 0; --k)
      if (L.Z[k] == Z[i] && L.m[k] == m[i]) s = k;
    out[i] = s;
  }
}

hipError_t primary_species_launch(const SpeciesList& L, const int* Z, const double* m, long long n, int* out,
                                  hipStream_t st) {
  const long long g = (n + 255) / 256;
  hipLaunchKernelGGL(primary_species, dim3((unsigned)(g < 4096 ? (g > 0 ? g : 1) : 4096)), dim3(256), 0, st, L, Z, m,
                     n, out);
  return hipGetLastError();
}

hipError_t zero_spans_launch(const ZeroSpans& z, hipStream_t st) {
  hipLaunchKernelGGL(zero_spans, dim3(zero_spans_grid(z)), dim3(256), 0, st, z);
  return hipGetLastError();
}

hipError_t flush_energy_launch(const double* we, int waves, double* out, hipStream_t st) {
  hipLaunchKernelGGL(flush_energy, dim3(1), dim3(64), 0, st, we, waves, out);
  return hipGetLastError();
}

}  // namespace mt
