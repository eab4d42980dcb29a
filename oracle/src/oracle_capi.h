/* ORACLE — test infrastructure only: the C interface tests/ and bench.py's
 * cpu_baseline leg load through ctypes (oracle/_build/liboracle.so). */
#ifndef MYTRIM_ORACLE_CAPI_H
#define MYTRIM_ORACLE_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

/* A cascade run: target, geometry, primaries, draw scheme and tally bins. */
typedef struct {
  /* materials: element e of material k is e in [mat_el_off[k], mat_el_off[k+1]) */
  int n_mat;
  const double* mat_rho;
  const int* mat_el_off;
  const int* el_Z;
  const double* el_m;
  const double* el_t;
  const double* el_Edisp;
  const double* el_Elbind;
  /* geometry: 0 SampleSolid, 1 SampleLayers, 2 sampleClusters, 3 SampleWire,
     4 SampleBurriedWire */
  int geo;
  double w[3];
  int bc[3]; /* 0 PBC, 1 INF, 2 CUT */
  int n_layers;
  const double* layer_thick;
  int kn[3];
  int n_cl;
  const double* cl_x;
  const double* cl_y;
  const double* cl_z;
  const double* cl_r;
  /* SimconfType fields */
  double tmin, tau, cw, scale, Ef;
  int potential;
  /* primaries: histories hist0 .. hist0+n_hist-1; pka arrays hold one entry
     (pka_uniform != 0) or one per history */
  long long n_hist, hist0;
  int pka_uniform;
  const int* pka_Z;
  const double* pka_m;
  const double* pka_E;
  const double* pka_pos; /* 3 per entry */
  const double* pka_dir; /* 3 per entry */
  unsigned seed;
  int rng; /* 0: std::mt19937 reseeded per history, 1: device Philox scheme */
  int follow_max_gen; /* < 0: follow every recoil (TrimBase); else gen < max (TrimPrimaries) */
  int kp;             /* modified Kinchin-Pease at gen == follow_max_gen */
  /* histogram bins: bin 0 underflow, 1..nb, nb+1 overflow */
  double range_x0, range_bw;
  int range_nb;
  double lat_y0, lat_z0, lat_bw;
  int lat_nb;
  double vac_x0, vac_bw;
  int vac_nb;
  int nthreads;
} orc_run_spec;

typedef struct {
  long long n_coll, n_flight, vac, repl, subst, inter, lost, vacuum, rutherford, newton, disp;
  double E0, Eel, Enuc, Eunf, Elost, Edrop, Evac;
  long long range_n;
  double range_sum, range_sum2;
  long long* range_hist; /* range_nb + 2, caller-allocated */
  long long* lat_y_hist; /* lat_nb + 2 */
  long long* lat_z_hist;
  long long* vac_hist; /* vac_nb + 2 */
  long long* repl_hist;
  /* per history, indexed h - hist0 (each may be NULL) */
  int* h_vac;
  long long* h_coll;
  int* h_flights;
  double* h_x;
  double* h_y;
  double* h_z;
  int* h_state;
  double* h_Eel;
  double seconds;
} orc_run_out;

/* One line of the reference's stream-writing TrimBase subclasses (kinds in
 * runs.cpp, LogHooks); 80 bytes, the layout of the device's mt_record. */
typedef struct {
  double pos[3];
  double E, value, m;
  unsigned long long id;
  unsigned hist;
  int Z, gen, tag, kind, pad;
} orc_record;

int orc_load_tables(const char* path);
// Plants a deliberate error in the device draw scheme (see oracle.h); 0 clears it.
void orc_plant_fault(int f);
double orc_rpstop(int z2, double e);
double orc_rstop(int z1, double m1, double E, int z2);
double orc_getrstop(int Z1, double m1, double E, double rho, int n_el, const int* Z2,
                    const double* m2, const double* t);
/* out: am az arho mu a f epsdg fd kd, then my ec ai fi per element */
void orc_average(int Z1, double m1, double tmin, double rho, int n_el, const int* Z2,
                 const double* m2, const double* t, double* out);
int orc_scatter(double eps, double b, int potential, double* out3);
void orc_philox(const unsigned* ctr4, const unsigned* key2, unsigned* out4);
unsigned long long orc_flight_key_child(unsigned long long parent, unsigned ic);
/* 64 bits of a flight's first draw (r1) */
unsigned long long orc_first_draw(unsigned long long key, unsigned seed, unsigned hist);
void orc_mt_stream(unsigned seed, long long skip, long long n, double* out);
unsigned orc_mt_history_seed(unsigned master, long long h);
int orc_uo2(int seed, double r, double cbf, int nev, char* erec, int erec_len, char* clcoor,
            int clcoor_len, char* dist, int dist_len);
int orc_run(const orc_run_spec* spec, orc_run_out* out);
/* algorithmic work per collision: 16 doubles (see runs.cpp) */
int orc_workcount(const orc_run_spec* spec, double* out);
/* (eps, b, E, target Z) of every collision of the spec's histories */
long long orc_scatter_inputs(const orc_run_spec* spec, double* rows, long long max_rows);
/* event records (mask: 1 defects, 2 history, 4 phonons) and the vacancy map
 * (int[20][20][3], accumulated; NULL: none) of the spec's histories, run in
 * order on one thread; mask 16: the layer transport (long long[n_mat][93],
 * accumulated); returns the number of records (those past max_records are
 * not stored) */
long long orc_run_log(const orc_run_spec* spec, int mask, const int* vacmap_Z, orc_record* records,
                      long long max_records, int* vacmap, long long* transport);
/* per-collision trace of history `h` (rows of 21 doubles, see runs.cpp) */
int orc_trace(const orc_run_spec* spec, long long h, double* rows, int max_rows);

#ifdef __cplusplus
}
#endif
#endif
