/*
 * mytrim_gpu.h — C ABI of the MI355X (gfx950) binary-collision cascade engine.
 *
 * Plain pointers, sizes and POD structs only.  Every function returns 0 on
 * success or a negative MT_E* code; mt_last_error() then holds a message.
 * Host arrays are read during the call and may be freed afterwards.
 *
 * What each entry point replaces in the reference (MyTRIM):
 *   mt_ctx_create / mt_ctx_set_tables   SimconfType construction and
 *   mt_default_tables                   readDataFiles (simconf.C:36-137)
 *   mt_ctx_set_target                   MaterialBase + Element set-up and
 *                                       prepare() (material.C:37-74), the
 *                                       Sample* geometry (sample*.h) and the
 *                                       SimconfType knobs (simconf.h:66)
 *   mt_run_cascades                     the drain loop of apps/runmytrim.C:71-93
 *                                       (averages + TrimBase::trim per ion,
 *                                       trim.C:35-425) for a batch of primaries,
 *                                       with the tally hooks of trim.C:439-527
 *                                       and apps/src/TrimVacCount.C:31-53
 *   mt_trim_flight                      one TrimBase::trim(pka, recoils) call
 *                                       (trim.C:35-425) driven by the caller's
 *                                       SimconfType::drand() stream, with a
 *                                       per-collision event log so the caller
 *                                       can replay the virtual hooks
 *   mt_stopping_batch                   MaterialBase::getrstop (material.C:114-122)
 *   mt_scatter_batch                    the scattering block of trim.C:172-272
 *   mt_stopping_batch_arith,            the same two in the arithmetic of a
 *   mt_scatter_batch_arith              cascade precision mode
 */
#ifndef MYTRIM_GPU_H
#define MYTRIM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MT_OK = 0,
  MT_E_ARG = -1,      /* invalid argument */
  MT_E_HIP = -2,      /* HIP runtime error */
  MT_E_STATE = -3,    /* call out of order (e.g. no target set) */
  MT_E_DEVICE = -4,   /* device-side failure flag (queue overflow, runaway loop, NaN) */
  MT_E_NOMEM = -5,
  MT_E_LIMIT = -6,    /* input exceeds a documented limit */
  MT_E_MORE = -7      /* mt_trim_flight: the uniform stream or the event
                         buffer ran out; call again with more */
};

/* boundary conditions (SampleBase::sampleBoundary, sample.h:49-54) */
enum { MT_PBC = 0, MT_INF = 1, MT_CUT = 2 };
/* geometry kinds: SampleSolid (sample_solid.C:25-29), SampleLayers
 * (sample_layers.C:26-49), sampleClusters (sample_clusters.C:43-133),
 * SampleWire (a cylinder along z inscribed in the w[0] x w[1] box, material 0
 * inside, vacuum outside; sample_wire.C:37-46) and SampleBurriedWire (the
 * wire under a 250 A cover of material 1 at -250 <= z < 0, material 1
 * around the wire, vacuum above z = w[2] and below z = -250;
 * sample_burried_wire.C:34-55) */
enum { MT_GEO_SOLID = 0, MT_GEO_LAYERS = 1, MT_GEO_CLUSTERS = 2, MT_GEO_WIRE = 3, MT_GEO_BURRIED_WIRE = 4 };
/* interatomic potentials (TrimBase::Potential, trim.h:63-68) */
enum { MT_UNIVERSAL = 0, MT_MOLIERE = 1, MT_CKR = 2 };
/* arithmetic of the cascade kernel */
enum { MT_FP64 = 0, MT_FP32 = 1, MT_FP32_SURROGATE = 2 };

typedef struct mt_ctx mt_ctx;

/* Target: materials, geometry and the SimconfType knobs. */
typedef struct {
  int n_mat;
  const double* mat_rho;  /* g/cm^3 */
  const int* mat_el_off;  /* n_mat + 1 offsets into the element arrays */
  const int* el_Z;
  const double* el_m;
  const double* el_t;     /* stoichiometry (normalised by prepare()) */
  const double* el_Edisp;
  const double* el_Elbind;
  int geo;                /* MT_GEO_* */
  double w[3];            /* sample size [A] */
  int bc[3];              /* MT_PBC / MT_INF / MT_CUT */
  int n_layers;           /* MT_GEO_LAYERS: thickness of layer i is layer_thick[i] */
  const double* layer_thick;
  int kn[3];              /* MT_GEO_CLUSTERS: spatial hash cells per axis */
  int n_cl;               /* clusters (sphere centre, radius); material 1 inside */
  const double* cl_x;
  const double* cl_y;
  const double* cl_z;
  const double* cl_r;
  double tmin, tau, cw, scale; /* SimconfType (simconf.C:44-49), length scale */
  int potential;               /* MT_UNIVERSAL / MT_MOLIERE / MT_CKR */
} mt_target;

/* A batch of primaries (generation 0).  With pka_uniform != 0 the arrays hold
 * one ion used for every history, else one entry per history.  Runs refuse
 * (MT_E_ARG) an energy that is not positive and finite, a non-finite position
 * and a direction that is zero or not finite (it need not be normalised). */
typedef struct {
  long long n_hist;
  long long hist0;   /* index of the first history (selects its random streams) */
  int pka_uniform;
  const int* Z;
  const double* m;
  const double* E;   /* eV */
  const double* pos; /* 3 per entry */
  const double* dir; /* 3 per entry */
  double Ef;         /* IonBase::_Ef, final energy (ion.C:26) */
  uint32_t seed;
} mt_primaries;

/* Which recoils are followed and which tallies are binned. */
typedef struct {
  int follow_max_gen; /* < 0: every recoil (TrimBase); k: gen < k (TrimPrimaries/TrimRecoils) */
  int kp;             /* modified Kinchin-Pease vacancies at gen == follow_max_gen */
  int arith;          /* MT_FP64 / MT_FP32 / MT_FP32_SURROGATE */
  double range_x0, range_bw; int range_nb;           /* final x of resting gen-0 ions */
  double lat_y0, lat_z0, lat_bw; int lat_nb;         /* final y - y0, z - z0 of the same */
  double vac_x0, vac_bw; int vac_nb;                 /* recoil x at vacancy / replacement */
  int per_history;    /* fill the per-history arrays of mt_tallies */
} mt_run_opts;

/* Tallies.  Histogram arrays have nb + 2 bins: [0] underflow, [1..nb], [nb+1]
 * overflow; they and the per-history arrays are caller-allocated (may be NULL). */
typedef struct {
  long long n_coll, n_flight, vac, repl, subst, inter, lost, vacuum, rutherford, newton, disp;
  double E0, Eel, Enuc, Eunf, Elost, Edrop, Evac;
  long long range_n;
  double range_sum, range_sum2;
  long long* range_hist;
  long long* lat_y_hist;
  long long* lat_z_hist;
  long long* vac_hist;
  long long* repl_hist;
  int* h_vac;            /* per history, index h - hist0 */
  long long* h_coll;
  int* h_flights;
  double* h_x;           /* final position and state of the gen-0 ion */
  double* h_y;
  double* h_z;
  int* h_state;          /* IonBase::StateType, -1 = ended in vacuum */
  double* h_Eel;
  /* launch record */
  double kernel_ms;      /* cascade kernel time (HIP events on its stream) */
  int grid, block, waves;
  long long queue_peak;  /* most recoils held in the HBM queue */
  long long queue_ops;   /* wave-level HBM queue transfers */
  long long surrogate;   /* collisions whose angle came from the surrogate */
  long long coll_primary;/* collisions of gen-0 flights (the primaries' share of the work) */
  double sclk_mhz;       /* mean shader clock of the launch's waves, start to exit (their cycle
                            counters over the constant-rate wall clock) */
  double half_ms;        /* first wave's start to the admission of the middle primary (the
                            ramp-up is over); 0 if not recorded */
  long long flights_at_half; /* flights ended by then (each wave adds its own in batches) */
  double admit_ms;       /* first wave's start to the admission of the last primary (the drain
                            follows); 0 if not recorded */
  long long flights_at_admit; /* flights ended by then */
} mt_tallies;

/* Tuning knobs of the persistent kernel (0 = default). */
typedef struct {
  int block;          /* threads per workgroup (multiple of 64, at most 256) */
  int blocks_per_cu;  /* resident workgroups per CU (0: two waves per SIMD) */
  int stack_cap;      /* per-wave LDS recoil stack entries (0: 128) */
  long long queue_cap;/* HBM queue entries (power of two) */
  int lds_hist;       /* vacancy/replacement depth bins counted in LDS (the
                         leading ones; the rest in HBM): 0 as many as fit
                         without costing a resident workgroup, < 0 none,
                         n > 0 at most n */
  int graph;          /* >= 0: replay the run sequence (reset, kernel, flush,
                         copies) from a cached hipGraph; < 0: launch directly */
  int primary_waves64;/* 0: every wave admits primaries and keeps its recoils;
                         k in 1..63: k waves of every 64 admit the primaries and
                         hand every recoil to the HBM queue while primaries
                         remain, the others run recoils only (waves of one ion
                         class take one stopping branch) */
} mt_tuning;

/* One flight (TrimBase::trim) for the drop-in single-call path. */
typedef struct {
  int Z;
  double m, E;
  double pos[3], dir[3];
  int gen, tag, state;
  double Ef;
} mt_ion;

/* Per-collision record of mt_trim_flight, in trim.C order. */
typedef struct {
  double pka_E, pka_pos[3], pka_dir[3];  /* after the collision */
  double ls, dee, den;
  double rec_E, rec_pos[3], rec_dir[3];  /* recoil: E = den - Elbind, dir before normalisation */
  int material, element;                 /* indices into the target */
  int pka_state;                         /* IonBase::StateType after the collision */
  int fate;                              /* 0 dissipated, 1 above threshold, 2 lost (dropped) */
  int vacancy;                           /* 1: vacancyCreation(), 0: replacementCollision() (fate 1) */
  int mat_tag;                           /* material tag at this collision (cluster index) */
} mt_event;

/* Event streams of the batched path: one record per line the reference's
 * stream-writing TrimBase subclasses print, and TrimVacMap's map.
 *   kind 1      TrimDefectLog "V" + the recoil at vacancyCreation (trim.C:466-469)
 *   kind 2/3/4  TrimDefectLog "I" / "S" / "R" + the ion at rest (trim.C:471-477)
 *   kind 5      TrimHistory: the ion when a recoil is offered to followRecoil (trim.h:139-155)
 *   kind 6      TrimPhononOut::followRecoil: value Elbind + the recoil (trim.C:520-526)
 *   kind 7      TrimPhononOut::dissipateRecoilEnergy: value recoil E + Elbind + the recoil (trim.C:512-518)
 *   kind 8      TrimPhononOut::checkPKAState: value the ion's E + the ion at rest (trim.C:503-510)
 * The ion fields are those IonBase's operator<< prints (ion.C:74-79); id is
 * the flight key (unique per flight; the reference numbers ions in creation
 * order from a global counter).  Records of one history are in trim.C order;
 * histories interleave.  Recoils in kinds 6 and 7 carry id 0 and tag -1, as
 * they have not been queued when the reference prints them.
 * MT_LOG_TRANSPORT (layers geometry only) counts the atom moves a
 * SampleDynamic drain loop applies with addAtomsToLayer
 * (apps/mytrim_dynamic.C:160-178), the layers held fixed during the run: per
 * layer k and element Z, +1 for every flight that ends at x > 0, in the layer
 * it ends in, and -1 for every recoil flight (gen > 0), in the layer it starts
 * in. */
enum { MT_LOG_DEFECTS = 1, MT_LOG_HISTORY = 2, MT_LOG_PHONONS = 4, MT_LOG_VACMAP = 8, MT_LOG_TRANSPORT = 16 };
typedef struct {
  double pos[3];
  double E, value, m;
  unsigned long long id;
  uint32_t hist;
  int Z, gen, tag, kind, pad;
} mt_record;
typedef struct {
  int mask;                 /* MT_LOG_* */
  mt_record* records;       /* caller-allocated, max_records entries (may be NULL) */
  long long max_records;
  long long n_records;      /* out: records produced; those past max_records are dropped */
  int vacmap_Z[3];          /* TrimVacMap z1 z2 z3 (trim.h:176-201) */
  int* vacmap;              /* caller-allocated int[20][20][3] (x cell, y cell, constituent):
                               the run's counts are added (MT_LOG_VACMAP) */
  long long* transport;     /* caller-allocated long long[n_mat][93], indexed [layer][Z]:
                               the run's net moves are added (MT_LOG_TRANSPORT) */
} mt_log;

/* Scattering surrogate (no counterpart in the reference): an MLP
 * 3 -> 64 -> 64 -> 1 (tanh) from (log10 eps, log10 b, b) to ln sin^2(theta/2),
 * trained on the device on the engine's own universal-potential scattering
 * over the (eps, b) pairs cascades in the context's target draw; used by
 * MT_FP32_SURROGATE runs inside its domain, the MAGIC solve outside it. */
#define MT_SURROGATE_PARAMS 4481
#define MT_SURROGATE_NORM 12
typedef struct {
  int n_ions;              /* species the cascades start with; target atoms are added */
  const int* ion_Z;
  const double* ion_m;
  double E_min, E_max;     /* flight energies [eV] the training pairs cover */
  double eps_max;          /* largest reduced energy handed to the surrogate */
  long long n_train;       /* training pairs generated on the device */
  int epochs, batch;
  double lr;               /* Adam step size, cosine-annealed to lr / 100 */
  uint32_t seed;
  int bf16;                /* inference with BF16 dot products (else FP32 FMA) */
  int inputs;              /* 0 or 3: (log eps, log b, b); 2: (log eps, b), the log b
                              input held at zero (a 2 -> 64 -> 64 -> 1 network) */
  int sampling;            /* 0: pairs drawn as the target's cascades draw them;
                              1: eps log-uniform on [eps_min, eps_max], b uniform on
                              [0, b_max], b_max the widest reduced impact parameter
                              the cascades draw */
  double eps_min;          /* sampling 1: lower end of eps */
} mt_surrogate_opts;

const char* mt_last_error(void);
const char* mt_version(void);

int mt_ctx_create(int device, mt_ctx** out);
int mt_ctx_destroy(mt_ctx* ctx);
/* 92 rows x 16 doubles: mm1 m1 mnat rho atrho vfermi heat pcoef[8] lfctr
 * (SCOEF.95A + SLFCTR.dat columns).  Defaults to the tables built into the
 * library. */
int mt_ctx_set_tables(mt_ctx* ctx, const double* rows, int n_rows, int n_cols);
/* copy the tables built into the library (same 92 x 16 layout) */
int mt_default_tables(double* rows, int n_rows, int n_cols);
int mt_ctx_set_target(mt_ctx* ctx, const mt_target* target);
int mt_ctx_set_tuning(mt_ctx* ctx, const mt_tuning* tuning);

int mt_run_cascades(mt_ctx* ctx, const mt_primaries* pkas, const mt_run_opts* opts,
                    mt_tallies* out);

/* mt_run_cascades without allocating: no counterpart in the reference.  It
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * device buffer of the run (tables, primaries, queue, control words,
 * tallies) comes from one static array of 205 MiB that each device running
 * the library holds until the process ends; the results are copied straight
 * into the caller's mt_tallies arrays.  Same inputs and results as
 * mt_run_cascades (the hot kernel only: per_history must be 0).  Limits: at
 * most 2^22 histories, 2^18 per-history primaries, about 130 000 histogram
 * bins in all, 16 384 resident waves, and 2^21 HBM queue slots (a run that
 * needs more stops with "queue-full"); larger inputs return MT_E_LIMIT.
 * Calls on one device share the array: calls that overlap on one GPU race. */
int mt_run_cascades_noalloc(mt_ctx* ctx, const mt_primaries* pkas, const mt_run_opts* opts,
                            mt_tallies* out);

/* mt_run_cascades with event streams (TrimDefectLog, TrimHistory,
 * TrimPhononOut, TrimVacMap; trim.C:466-527, trim.h:139-227) written to
 * `log`.  Runs the per-history kernel. */
int mt_run_cascades_log(mt_ctx* ctx, const mt_primaries* pkas, const mt_run_opts* opts, mt_log* log,
                        mt_tallies* out);

/* Autotuner (no counterpart in the reference: the reference's tuning knob is
 * its thread count, apps/runmytrim.C:59-68).  Runs the first `sample`
 * histories of `pkas` once at the default shape (for the primaries' share of
 * the collisions), then under each of 17 candidate launch shapes (block size,
 * LDS stack depth, resident blocks per CU) without wave roles and with feeder
 * waves for 2.5 times that share, each with as many depth bins in LDS as fit
 * and with every bin in HBM (unless the context's tuning sets lds_hist), then
 * with three other feeder counts at the fastest roles shape of each
 * occupancy (waves per CU), `reps` times each.  A shape's rate is projected for a launch of all of `pkas`:
 * the sample launch's steady rate (collisions from the start to the
 * admission of the last primary, over that time) for the collisions of that
 * many histories, plus the sample launch's drain (the time after the last
 * admission).  When `pkas` holds at least twice `sample` histories, the
 * finalists (the fastest shape of each occupancy within 10 % of the fastest,
 * without and with roles, and without roles the runner-up of another shape)
 * are timed again on up to four times `sample`
 * histories, each with both depth-bin splits, and those timings decide; otherwise the two fastest (without and with roles) are
 * timed once more on the sample.  Roles are kept only when the primaries
 * make at least 2 % of the collisions and they run 3 % faster than the
 * fastest shape without them.  The pick stays in the context and is returned
 * in *best, its projected rate in *best_rate (collisions per second). */
int mt_autotune(mt_ctx* ctx, const mt_primaries* pkas, const mt_run_opts* opts, long long sample,
                int reps, mt_tuning* best, double* best_rate);

/* Single flight with the caller's uniform stream u[0..n_u): consumes draws in
 * the reference's order (trim.C:71, 143, 147, 322-333).  Returns MT_E_MORE
 * when the stream ran out (*n_used = n_u) or the event buffer filled
 * (*n_events = max_events): call again with more.  MT_E_DEVICE for a runaway
 * flight or a non-finite energy or position.  The ion is updated in place;
 * events[0..*n_events) hold every collision. */
int mt_trim_flight(mt_ctx* ctx, mt_ion* ion, const double* u, long long n_u, long long* n_used,
                   mt_event* events, int max_events, int* n_events);

/* Input / output scaling and domain the options give in the context's target:
 * e_mid e_inv_half lb_mid lb_inv_half b_mid b_inv_half b_floor y_off y_scale
 * eps_lo eps_hi b_hi (x0 = (log10 eps - e_mid) e_inv_half, x1 = (log10
 * max(b, b_floor) - lb_mid) lb_inv_half, x2 = (b - b_mid) b_inv_half,
 * ln s2 = y y_scale + y_off). */
int mt_surrogate_norm(mt_ctx* ctx, const mt_surrogate_opts* opts, float* norm);
/* Train the surrogate from `params` (initial values in, trained values out);
 * epoch_loss[e] = mean squared error (normalised units) of epoch e. */
int mt_surrogate_train(mt_ctx* ctx, const mt_surrogate_opts* opts, float* params, double* epoch_loss);
/* Use `params` for MT_FP32_SURROGATE runs on this context. */
int mt_surrogate_install(mt_ctx* ctx, const mt_surrogate_opts* opts, const float* params);
/* Installed surrogate at (eps, b): s2 = sin^2(theta/2), -1 outside its domain. */
int mt_surrogate_eval(mt_ctx* ctx, const double* eps, const double* b, long long n, double* s2);
/* Inference ablation of the learned scattering (no counterpart in the
 * reference: it measures whether the network can replace trim.C:181-272).
 * Evaluates n (eps, b) pairs `reps` times after one untimed pass, in one of
 * four ways: mode 0 the FP32 MAGIC solve of the universal potential (no
 * network needed), 1 the installed network with layer 2 in FP32 FMAs, 2 in
 * BF16 dot products (v_dot2_f32_bf16), 3 on the matrix cores
 * (v_mfma_f32_32x32x16_bf16, one 64 x 64 x 64 product per wave).  s2 receives
 * sin^2(theta/2) (the network's without its domain check); *kernel_ms the mean
 * time of one pass. */
int mt_surrogate_ablate(mt_ctx* ctx, int mode, const double* eps, const double* b, long long n, int reps,
                        double* s2, double* kernel_ms);
/* Loss and gradient of one batch (targets ln s2), FP32 back-propagation. */
int mt_surrogate_grad(mt_ctx* ctx, const mt_surrogate_opts* opts, const float* params, const double* eps,
                      const double* b, const double* ln_s2, int n, float* grad, double* loss);

/* getrstop for ion (Z1, m1) at energies E[0..n) in material `mat` [eV/A]. */
int mt_stopping_batch(mt_ctx* ctx, int Z1, double m1, int mat, const double* E, long long n,
                      double* se_out);
/* scattering for (eps, b) pairs: s2 = sin^2(theta/2), ct = cos(theta),
 * st = sin(theta), iters = Newton iterations (0: Rutherford branch). */
int mt_scatter_batch(mt_ctx* ctx, int potential, const double* eps, const double* b, long long n,
                     double* s2, double* ct, double* st, int* iters);

/* The two batches above in the arithmetic of a cascade precision mode
 * (MT_FP64, MT_FP32, MT_FP32_SURROGATE): each runs the device function the
 * cascade kernel of that mode evaluates (getrstop of material.C:114-122 and
 * the scattering of trim.C:172-272 in single precision; with
 * MT_FP32_SURROGATE the installed network answers inside its domain for
 * eps <= 10, with iters = -1 there), for per-function tolerance checks.
 * Inputs are rounded to single precision first in the FP32 modes. */
int mt_stopping_batch_arith(mt_ctx* ctx, int arith, int Z1, double m1, int mat, const double* E, long long n,
                            double* se_out);
int mt_scatter_batch_arith(mt_ctx* ctx, int arith, int potential, const double* eps, const double* b,
                           long long n, double* s2, double* ct, double* st, int* iters);

#ifdef __cplusplus
}
#endif
#endif /* MYTRIM_GPU_H */
