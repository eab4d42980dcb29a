// C ABI of the cascade engine (include/mytrim_gpu.h): context, target
// upload, batched cascades, per-function batches.  Host code only; the
// kernels live in mt_cascade.hip, mt_batch.hip and mt_flight.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <cstring>
#include <string>
#include <vector>

#include "mt_cascade.h"
#include "mt_flight.h"
#include "mt_tune.h"
#include "mt_host.h"
#include "mytrim_gpu.h"

namespace mt {
hipError_t cascade_occupancy(int arith, int potential, int geo, bool perhist, int block, size_t lds,
                             int* blocks_per_cu);
hipError_t cascade_launch(const Params& P, int arith, int potential, bool perhist, int grid, int block, size_t lds,
                          hipStream_t stream);
int cascade_stack_bytes(int cap);
const void* cascade_function(int arith, int potential, int geo, bool perhist);
int zero_spans_grid(const ZeroSpans& z);
const void* zero_spans_function();
const void* flush_energy_function();
hipError_t zero_spans_launch(const ZeroSpans& z, hipStream_t st);
hipError_t flush_energy_launch(const double* we, int waves, double* out, hipStream_t st);
hipError_t primary_species_launch(const SpeciesList& L, const int* Z, const double* m, long long n, int* out,
                                  hipStream_t st);
hipError_t stopping_launch(int arith, const unsigned char* tables, int species, int mat, const double* E, long long n,
                           double* out, hipStream_t stream);
hipError_t scatter_launch(int arith, int potential, const MlpDev& mlp, const double* eps, const double* b, long long n,
                          double* s2, double* ct, double* st, int* iters, hipStream_t stream);
hipError_t ablate_launch(int mode, const MlpDev& S, const double* eps, const double* b, long long n, double* s2,
                         hipStream_t stream);
hipError_t mlp_make_data(uint32_t seed, long long n, const MlpTriple* tri, int n_tri, double E_lo, int rect,
                         const MlpNorm& nm, float* X, float* T, hipStream_t st);
hipError_t mlp_train(float* p, const float* X, const float* T, long long n, int B, int epochs, float lr0,
                     float* rec, float* part, float* g, float* m, float* v, int* t, float* batch_loss,
                     double* losses, hipStream_t st);
hipError_t mlp_batch_grad(const float* p, const float* X, const float* T, int B, float* rec, float* part,
                          float* g, float* loss, hipStream_t st);
hipError_t mlp_pack(const float* p, uint32_t* w2bf, hipStream_t st);
hipError_t mlp_infer_launch(const MlpDev& S, const double* eps, const double* b, long long n, double* s2,
                            hipStream_t st);
}  // namespace mt

using namespace mt;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHECK(x)                                                                   \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return fail(MT_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, std::max<size_t>(n, 16));
    if (e == hipSuccess) bytes = std::max<size_t>(n, 16);
    return e;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// pinned host staging (hipHostMalloc) for the run's results
struct HBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&p, std::max<size_t>(n, 16), hipHostMallocDefault);
    if (e == hipSuccess) bytes = std::max<size_t>(n, 16);
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// pinned host memory that kernels write directly: mapped and coherent, at the
// device address hipHostGetDevicePointer gives
struct MBuf {
  void* p = nullptr;
  void* dev = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    const size_t want = std::max<size_t>(n, 256);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&dev, p, 0);
    if (e == hipSuccess)
      bytes = want;
    else
      release();
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    dev = nullptr;
    bytes = 0;
  }
};

// result block: control words, counters, energies, flushed energy sums
constexpr int kResCtrl = 0;
constexpr int kResCounters = kResCtrl + kNumCtrl;
constexpr int kResEnergy = kResCounters + kNumCounters;
constexpr int kResReduced = kResEnergy + kNumEnergy;
constexpr int kResWords = kResReduced + 4;

// what a cached run graph was built for (its memcpy nodes and kernel shape)
struct RunGraphKey {
  const void* fn = nullptr;
  int grid = 0, block = 0;
  size_t lds = 0;
  void *res = nullptr, *hist = nullptr;
  size_t hist_bytes = 0;
  void *stage_res = nullptr, *stage_hist = nullptr;
  bool operator==(const RunGraphKey& o) const {
    return fn == o.fn && grid == o.grid && block == o.block && lds == o.lds && res == o.res && hist == o.hist &&
           hist_bytes == o.hist_bytes && stage_res == o.stage_res && stage_hist == o.stage_hist;
  }
};

struct RunGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraphNode_t n_zero = nullptr, n_kern = nullptr, n_flush = nullptr;
  RunGraphKey key;
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key = RunGraphKey{};
  }
};

uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// Device memory of mt_run_cascades_noalloc, carved from one static array per
// device (`scratch` below); limits of one call:
constexpr size_t na_align(size_t x) { return (x + 255) & ~size_t(255); }
constexpr long long kNaMaxEntries = 1ll << 18;      // per-history primaries
constexpr long long kNaMaxHist = 1ll << 22;         // histories
constexpr size_t kNaResHistBytes = size_t(1) << 20;  // result block + the five histograms
constexpr int kNaMaxWaves = 16384;                  // resident waves of the launch
constexpr uint64_t kNaQueueSlots = uint64_t(1) << 21;
constexpr size_t kNaSlotBytes = 9 * 8 + 4 + 8;      // one entry of each queue array
// layout: tables | primaries (species, Z, m, E, pos, dir) | per-history
// flight counts | result block + histograms | per-wave energies | queue
constexpr size_t kNaOffTables = 0;
constexpr size_t kNaOffPka = na_align(kNaOffTables + kMaxTableBytes);
constexpr size_t kNaOffHout = na_align(kNaOffPka + size_t(kNaMaxEntries) * (2 * sizeof(int) + 8 * sizeof(double)));
constexpr size_t kNaOffRes = na_align(kNaOffHout + size_t(kNaMaxHist) * sizeof(int));
constexpr size_t kNaOffWave = na_align(kNaOffRes + kNaResHistBytes);
constexpr size_t kNaOffQueue = na_align(kNaOffWave + size_t(kNaMaxWaves) * 2 * kNumEnergy * sizeof(double));
constexpr size_t kNaBytes = na_align(kNaOffQueue + size_t(kNaQueueSlots) * kNaSlotBytes);
static_assert(kNaBytes <= (size_t(224) << 20), "static scratch above its stated 224 MiB");

// the array's device address on each device, resolved once per device at
// context creation (hipGetSymbolAddress with that device current)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

}  // namespace

// Scratch of mt_run_cascades_noalloc: each device that runs the library holds
// one copy until the process ends.  Host code never uses this name as an
// address; it uses the pointer resolved for the device (g_scratch_dev).
static __device__ __align__(256) unsigned char scratch[kNaBytes];

struct mt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int cus = 0;
  int lds_max = 64 * 1024;  // LDS bytes one workgroup may use
  double clock_khz = 100000.0;
  std::vector<CoefRow> coef;
  HostTarget tgt;
  bool has_target = false;
  mt_tuning tune{};
  // the last cascade launch's wall-clock stamps (kT0, kTAdmit) and the
  // flights ended at its last admission (kFlightsAtAdmit)
  unsigned long long last_t0 = 0, last_tadm = 0, last_fadm = 0, last_thalf = 0, last_fhalf = 0;
  DBuf tables, sh, cl, cx, cy, cz, cr, occ, occ_rank, occ_off, occ_idx;
  DBuf pka_sp, pka_E, pka_pos, pka_dir;
  DBuf res, wave_energy, hist_all, h_out;
  HBuf stage_res, stage_hist;
  RunGraph graph;
  DBuf per[8];
  DBuf q[kQueueArrays];
  uint64_t qcap = 0;
  // single-flight path: device tables per ion species (Z, m), event buffers
  std::map<std::pair<int, double>, std::unique_ptr<DBuf>> flight_tables;
  std::map<std::pair<int, double>, int> flight_table_bytes;
  DBuf f_in;     // single flight: the ion and the caller's uniforms
  HBuf f_stage;  // their pinned staging
  MBuf f_out;    // what the flight kernel writes: status words, the ion, events
  // scattering surrogate
  DBuf mlp_p, mlp_w2bf, mlp_X, mlp_T, mlp_rec, mlp_part, mlp_g, mlp_m, mlp_v, mlp_t, mlp_bl, mlp_tri;
  MlpNorm mlp_norm{};
  bool mlp_ready = false;
  int mlp_bf16 = 0;
  DBuf scratch_a, scratch_b, scratch_c, scratch_d, scratch_e;
  DBuf log_rec, log_map, log_tr;  // event records, TrimVacMap map, layer transport
  // mt_run_cascades_noalloc: the device's static scratch and a host buffer
  // for the table blob, sized once here
  void* na_scratch = nullptr;
  std::vector<unsigned char> na_blob;

  ~mt_ctx() {
    graph.release();
    DBuf* all[] = {&tables, &sh, &cl, &cx, &cy, &cz, &cr, &occ, &occ_rank, &occ_off, &occ_idx, &pka_sp, &pka_E, &pka_pos, &pka_dir, &res,
                   &wave_energy, &hist_all, &h_out, &scratch_a, &scratch_b, &scratch_c, &scratch_d, &scratch_e,
                   &log_rec, &log_map, &log_tr};
    for (DBuf* b : all) b->release();
    stage_res.release();
    stage_hist.release();
    for (auto& b : per) b.release();
    for (auto& b : q) b.release();
    for (auto& kv : flight_tables) kv.second->release();
    f_in.release();
    f_stage.release();
    f_out.release();
    DBuf* ml[] = {&mlp_p,     &mlp_w2bf, &mlp_X, &mlp_T, &mlp_rec, &mlp_part,
                  &mlp_g,     &mlp_m,    &mlp_v, &mlp_t, &mlp_bl,  &mlp_tri};
    for (DBuf* b : ml) b->release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

extern "C" {

const char* mt_last_error(void) { return g_err.c_str(); }
const char* mt_version(void) { return "mytrim_amd 0.1 (gfx950)"; }

int mt_ctx_create(int device, mt_ctx** out) {
  if (!out) return fail(MT_E_ARG, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  HIPCHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(MT_E_ARG, "no such device");
  HIPCHECK(hipSetDevice(device));
  mt_ctx* c = new mt_ctx;
  c->device = device;
  c->coef = default_coefs();
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device);
  int khz = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device);
  int lds = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  if (e == hipSuccess && lds > 0) c->lds_max = lds;
  if (e != hipSuccess) {
    delete c;
    return fail(MT_E_HIP, std::string("context set-up: ") + hipGetErrorString(e));
  }
  if (khz > 0) c->clock_khz = khz;
  // the static scratch's address on this device (resolved once per device)
  if (device < kMaxDevices) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
      void* p = nullptr;
      const hipError_t se = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (se != hipSuccess) {
        delete c;
        return fail(MT_E_HIP, std::string("static scratch address: ") + hipGetErrorString(se));
      }
      g_scratch_dev[device] = p;
    }
    c->na_scratch = g_scratch_dev[device];
  }
  c->na_blob.assign(kMaxTableBytes, 0);
  *out = c;
  return MT_OK;
}

int mt_ctx_destroy(mt_ctx* ctx) {
  if (!ctx) return fail(MT_E_ARG, "ctx is NULL");
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return MT_OK;
}

int mt_ctx_set_tables(mt_ctx* ctx, const double* rows, int n_rows, int n_cols) {
  if (!ctx || !rows) return fail(MT_E_ARG, "NULL argument");
  if (n_rows != 92 || n_cols != 16) return fail(MT_E_ARG, "tables must be 92 x 16");
  for (int i = 0; i < 92; ++i) {
    const double* v = rows + 16 * i;
    CoefRow& c = ctx->coef[i];
    c.mm1 = v[0];
    c.m1 = v[1];
    c.mnat = v[2];
    c.rho = v[3];
    c.atrho = v[4];
    c.vfermi = v[5];
    c.heat = v[6];
    for (int j = 0; j < 8; ++j) c.pcoef[j] = v[7 + j];
    c.lfctr = v[15];
  }
  // single-flight tables are cached per ion; the next flight rebuilds them
  for (auto& kv : ctx->flight_tables) kv.second->release();
  ctx->flight_tables.clear();
  ctx->flight_table_bytes.clear();
  return MT_OK;
}

int mt_default_tables(double* rows, int n_rows, int n_cols) {
  if (!rows) return fail(MT_E_ARG, "NULL argument");
  if (n_rows != 92 || n_cols != 16) return fail(MT_E_ARG, "tables are 92 x 16");
  const std::vector<CoefRow>& c = default_coefs();
  for (int i = 0; i < 92; ++i) {
    double* v = rows + 16 * i;
    v[0] = c[i].mm1;
    v[1] = c[i].m1;
    v[2] = c[i].mnat;
    v[3] = c[i].rho;
    v[4] = c[i].atrho;
    v[5] = c[i].vfermi;
    v[6] = c[i].heat;
    for (int j = 0; j < 8; ++j) v[7 + j] = c[i].pcoef[j];
    v[15] = c[i].lfctr;
  }
  return MT_OK;
}

int mt_ctx_set_tuning(mt_ctx* ctx, const mt_tuning* t) {
  if (!ctx || !t) return fail(MT_E_ARG, "NULL argument");
  if (t->block && (t->block % 64 || t->block > kCascadeMaxBlock))
    return fail(MT_E_ARG, "block must be a multiple of 64 and at most 256 (the kernel's launch bound)");
  if (t->queue_cap && (t->queue_cap & (t->queue_cap - 1))) return fail(MT_E_ARG, "queue_cap must be a power of two");
  if (t->stack_cap < 0 || t->stack_cap > 1024) return fail(MT_E_ARG, "stack_cap out of range");
  if (t->primary_waves64 < 0 || t->primary_waves64 > 63) return fail(MT_E_ARG, "primary_waves64 out of range 0..63");
  ctx->tune = *t;
  return MT_OK;
}

int mt_ctx_set_target(mt_ctx* ctx, const mt_target* t) {
  if (!ctx || !t) return fail(MT_E_ARG, "NULL argument");
  if (t->n_mat < 1 || t->n_mat > kMaxMaterials) return fail(MT_E_ARG, "n_mat out of range");
  HostTarget tg;
  int n_el = 0;
  for (int k = 0; k < t->n_mat; ++k) {
    HostMaterial m;
    m.rho = t->mat_rho[k];
    if (!(m.rho > 0.0)) return fail(MT_E_ARG, "material density must be positive");
    const int b = t->mat_el_off[k], e = t->mat_el_off[k + 1];
    if (e <= b) return fail(MT_E_ARG, "material without elements");
    for (int i = b; i < e; ++i) {
      if (t->el_Z[i] < 1 || t->el_Z[i] > 92) return fail(MT_E_ARG, "element Z out of range 1..92");
      if (!(t->el_m[i] > 0.0)) return fail(MT_E_ARG, "element mass must be positive");
      m.el.push_back(HostElement{t->el_Z[i], t->el_m[i], t->el_t[i], t->el_Edisp[i], t->el_Elbind[i]});
    }
    n_el += e - b;
    m.prepare();
    if (!(m.am > 0.0) || !(m.arho > 0.0)) return fail(MT_E_ARG, "stoichiometry invalid");
    tg.mat.push_back(m);
  }
  if (n_el > kMaxElements) return fail(MT_E_LIMIT, "too many elements");
  Geometry& g = tg.geo;
  std::memset(&g, 0, sizeof(g));
  g.kind = t->geo;
  g.n_mat = t->n_mat;
  for (int i = 0; i < 3; ++i) {
    g.w[i] = t->w[i];
    g.bc[i] = t->bc[i];
    if (!(g.w[i] > 0.0)) return fail(MT_E_ARG, "sample size must be positive");
    if (g.bc[i] < 0 || g.bc[i] > 2) return fail(MT_E_ARG, "bad boundary condition");
  }
  if (g.kind == MT_GEO_LAYERS) {
    if (t->n_layers < 1 || t->n_layers > kMaxLayers) return fail(MT_E_ARG, "n_layers out of range");
    g.n_layers = t->n_layers;
    double d = 0.0;
    for (int i = 0; i < t->n_layers; ++i) {  // sample_layers.C:32-37
      d += t->layer_thick[i];
      g.layer_end[i] = d;
      tg.layers.push_back(t->layer_thick[i]);
    }
  } else if (g.kind == MT_GEO_CLUSTERS) {
    if (t->n_mat < 2) return fail(MT_E_ARG, "clusters need two materials (matrix, inside)");
    for (int i = 0; i < 3; ++i) {
      if (t->kn[i] < 1) return fail(MT_E_ARG, "kn must be >= 1");
      g.kn[i] = t->kn[i];
      g.kd[i] = g.w[i] / double(g.kn[i]);
    }
    const size_t ncell = size_t(g.kn[0]) * g.kn[1] * g.kn[2];
    if (ncell > (size_t(1) << 28)) return fail(MT_E_LIMIT, "spatial hash too large");
    tg.sh.assign(ncell, -1);
    tg.cl.assign(size_t(std::max(t->n_cl, 0)), -1);
    g.cmr = 0.0;
    for (int cn = 0; cn < t->n_cl; ++cn) {  // sampleClusters::addCluster (sample_clusters.C:187-224)
      const double c[3] = {t->cl_x[cn], t->cl_y[cn], t->cl_z[cn]};
      int k[3];
      for (int i = 0; i < 3; ++i) {
        k[i] = int(std::floor((c[i] * g.kn[i]) / g.w[i])) % g.kn[i];
        k[i] = (k[i] < 0) ? k[i] + g.kn[i] : k[i];
      }
      int l = k[0] + g.kn[0] * (k[1] + g.kn[1] * k[2]);
      if (tg.sh[l] < 0) {
        tg.sh[l] = cn;
      } else {
        l = tg.sh[l];
        while (tg.cl[l] >= 0) l = tg.cl[l];
        tg.cl[l] = cn;
      }
      if (t->cl_r[cn] > g.cmr) g.cmr = t->cl_r[cn];
      tg.clusters.push_back(HostCluster{c[0], c[1], c[2], t->cl_r[cn]});
    }
    HIPCHECK(hipSetDevice(ctx->device));
    const size_t n = std::max<size_t>(tg.clusters.size(), 1);
    HIPCHECK(ctx->sh.ensure(ncell * sizeof(int)));
    HIPCHECK(ctx->cl.ensure(n * sizeof(int)));
    HIPCHECK(ctx->cx.ensure(n * sizeof(double)));
    HIPCHECK(ctx->cy.ensure(n * sizeof(double)));
    HIPCHECK(ctx->cz.ensure(n * sizeof(double)));
    HIPCHECK(ctx->cr.ensure(n * sizeof(double)));
    std::vector<double> v[4];
    for (auto& c : tg.clusters) {
      v[0].push_back(c.x);
      v[1].push_back(c.y);
      v[2].push_back(c.z);
      v[3].push_back(c.r);
    }
    HIPCHECK(hipMemcpy(ctx->sh.p, tg.sh.data(), ncell * sizeof(int), hipMemcpyHostToDevice));
    if (!tg.clusters.empty()) {
      HIPCHECK(hipMemcpy(ctx->cl.p, tg.cl.data(), tg.cl.size() * sizeof(int), hipMemcpyHostToDevice));
      DBuf* dst[4] = {&ctx->cx, &ctx->cy, &ctx->cz, &ctx->cr};
      for (int i = 0; i < 4; ++i)
        HIPCHECK(hipMemcpy(dst[i]->p, v[i].data(), v[i].size() * sizeof(double), hipMemcpyHostToDevice));
    }
    // occupancy grid for the lookup's early exit: cells of about half the
    // largest radius (at most 1024 per axis, 2^30 in all); a cell's bit is set
    // when a sphere, radius padded by 1e-9 of the sample size, meets the cell
    // (at its periodic images along PBC axes)
    {
      const double cs = std::max(0.5 * g.cmr, 1e-9 * std::max(g.w[0], std::max(g.w[1], g.w[2])));
      long long on[3];
      for (int i = 0; i < 3; ++i) on[i] = std::min(1024ll, std::max(1ll, (long long)std::ceil(g.w[i] / cs)));
      while (on[0] * on[1] * on[2] > (1ll << 30)) {
        int m = 0;
        for (int i = 1; i < 3; ++i)
          if (on[i] > on[m]) m = i;
        on[m] = (on[m] + 1) / 2;
      }
      for (int i = 0; i < 3; ++i) {
        g.on[i] = int(on[i]);
        g.on_w[i] = double(on[i]) / g.w[i];
        g.ks[i] = int(g.cmr / g.kd[i]) + 1;
        g.inv_w[i] = 1.0 / g.w[i];
      }
      const long long nbits = on[0] * on[1] * on[2];
      std::vector<uint32_t> bits(size_t((nbits + 31) / 32), 0u);
      const double pad = 1e-9 * std::max(g.w[0], std::max(g.w[1], g.w[2]));
      std::vector<std::pair<uint64_t, int>> meets;  // (cell, cluster) for the candidate lists
      for (int ci = 0; ci < int(tg.clusters.size()); ++ci) {
        const HostCluster& c = tg.clusters[size_t(ci)];
        const double cc[3] = {c.x, c.y, c.z}, rr = c.r + pad;
        long long lo[3], hi[3];
        for (int i = 0; i < 3; ++i) {
          lo[i] = (long long)std::floor((cc[i] - rr) * double(on[i]) / g.w[i]);
          hi[i] = (long long)std::floor((cc[i] + rr) * double(on[i]) / g.w[i]);
          if (g.bc[i] != MT_PBC) {
            lo[i] = std::max(lo[i], 0ll);
            hi[i] = std::min(hi[i], on[i] - 1);
          } else if (hi[i] - lo[i] + 1 > on[i]) {  // spans the axis
            lo[i] = 0;
            hi[i] = on[i] - 1;
          }
        }
        for (long long j0 = lo[0]; j0 <= hi[0]; ++j0)
          for (long long j1 = lo[1]; j1 <= hi[1]; ++j1)
            for (long long j2 = lo[2]; j2 <= hi[2]; ++j2) {
              const long long j[3] = {j0, j1, j2};
              double d2 = 0.0;
              long long b = 0, mul = 1;
              for (int i = 0; i < 3; ++i) {
                const double a = double(j[i]) * g.w[i] / double(on[i]);
                const double e = double(j[i] + 1) * g.w[i] / double(on[i]);
                const double d = std::max(std::max(a - cc[i], cc[i] - e), 0.0);
                d2 += d * d;
                b += (((j[i] % on[i]) + on[i]) % on[i]) * mul;
                mul *= on[i];
              }
              if (d2 <= rr * rr) {
                bits[size_t(b >> 5)] |= 1u << (b & 31);
                meets.emplace_back(uint64_t(b), ci);
              }
            }
      }
      HIPCHECK(ctx->occ.ensure(bits.size() * sizeof(uint32_t)));
      HIPCHECK(hipMemcpy(ctx->occ.p, bits.data(), bits.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      g.occ = ctx->occ.as<uint32_t>();
      // candidate lists (the same (cell, cluster) pairs as the bits)
      std::vector<uint32_t> rank(bits.size() + 1, 0u);
      for (size_t w = 0; w < bits.size(); ++w) rank[w + 1] = rank[w] + uint32_t(__builtin_popcount(bits[w]));
      const size_t nset = rank[bits.size()];
      auto rank_of = [&](uint64_t b) {
        const uint32_t word = bits[size_t(b >> 5)];
        return rank[size_t(b >> 5)] + uint32_t(__builtin_popcount(word & ((1u << (b & 31)) - 1u)));
      };
      std::sort(meets.begin(), meets.end());
      std::vector<int> off(nset + 1, 0), idx;
      idx.reserve(meets.size());
      auto overlap = [&](int a, int b) {  // the lookup's minimum-image distance
        const HostCluster &p = tg.clusters[size_t(a)], &q = tg.clusters[size_t(b)];
        double d[3] = {p.x - q.x, p.y - q.y, p.z - q.z}, r2 = 0.0;
        for (int i = 0; i < 3; ++i) {
          if (g.bc[i] == MT_PBC) d[i] -= std::round(d[i] * g.inv_w[i]) * g.w[i];
          r2 += d[i] * d[i];
        }
        // touching within rounding counts as overlapping
        return r2 < (p.r + q.r + pad) * (p.r + q.r + pad);
      };
      for (size_t i = 0; i < meets.size();) {
        size_t e = i;
        while (e < meets.size() && meets[e].first == meets[i].first) ++e;
        bool clash = false;
        for (size_t a = i; a < e && !clash; ++a)
          for (size_t b = a + 1; b < e && !clash; ++b) clash = overlap(meets[a].second, meets[b].second);
        const uint32_t k = rank_of(meets[i].first);
        off[k] = int(idx.size());
        if (clash) {
          idx.push_back(-1);
        } else {
          for (size_t a = i; a < e; ++a) idx.push_back(meets[a].second);
        }
        i = e;
      }
      // every set bit has at least one pair: off is non-decreasing; close it
      off[nset] = int(idx.size());
      for (size_t k = nset; k-- > 0;)
        if (off[k] > off[k + 1]) off[k] = off[k + 1];
      if (idx.empty()) idx.push_back(-1);
      HIPCHECK(ctx->occ_rank.ensure(rank.size() * sizeof(uint32_t)));
      HIPCHECK(ctx->occ_off.ensure(off.size() * sizeof(int)));
      HIPCHECK(ctx->occ_idx.ensure(idx.size() * sizeof(int)));
      HIPCHECK(hipMemcpy(ctx->occ_rank.p, rank.data(), rank.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(ctx->occ_off.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(ctx->occ_idx.p, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
      g.occ_rank = ctx->occ_rank.as<uint32_t>();
      g.occ_off = ctx->occ_off.as<int>();
      g.occ_idx = ctx->occ_idx.as<int>();
    }
    g.sh = ctx->sh.as<int>();
    g.cl = ctx->cl.as<int>();
    g.cx = ctx->cx.as<double>();
    g.cy = ctx->cy.as<double>();
    g.cz = ctx->cz.as<double>();
    g.cr = ctx->cr.as<double>();
  } else if (g.kind == MT_GEO_BURRIED_WIRE) {
    if (t->n_mat < 2) return fail(MT_E_ARG, "a buried wire needs two materials (wire, cover)");
  } else if (g.kind != MT_GEO_SOLID && g.kind != MT_GEO_WIRE) {
    return fail(MT_E_ARG, "unknown geometry kind");
  }
  if (t->potential < 0 || t->potential > 2) return fail(MT_E_ARG, "unknown potential");
  if (!(t->scale > 0.0)) return fail(MT_E_ARG, "length scale must be positive");
  tg.tmin = t->tmin;
  tg.tau = t->tau;
  tg.cw = t->cw;
  tg.scale = t->scale;
  tg.potential = t->potential;
  ctx->tgt = tg;
  ctx->has_target = true;
  for (auto& kv : ctx->flight_tables) kv.second->release();
  ctx->flight_tables.clear();
  ctx->flight_table_bytes.clear();
  return MT_OK;
}

static int run_cascades(mt_ctx* ctx, const mt_primaries* pk, const mt_run_opts* o, mt_log* lg, mt_tallies* out);

int mt_run_cascades(mt_ctx* ctx, const mt_primaries* pk, const mt_run_opts* o, mt_tallies* out) {
  return run_cascades(ctx, pk, o, nullptr, out);
}

int mt_run_cascades_log(mt_ctx* ctx, const mt_primaries* pk, const mt_run_opts* o, mt_log* lg, mt_tallies* out) {
  if (!lg) return fail(MT_E_ARG, "NULL log");
  if (lg->mask & ~(MT_LOG_DEFECTS | MT_LOG_HISTORY | MT_LOG_PHONONS | MT_LOG_VACMAP | MT_LOG_TRANSPORT))
    return fail(MT_E_ARG, "unknown log bits");
  if (lg->mask & MT_LOG_TRANSPORT) {
    if (!lg->transport) return fail(MT_E_ARG, "transport buffer missing");
    if (ctx && ctx->has_target && ctx->tgt.geo.kind != MT_GEO_LAYERS)
      return fail(MT_E_ARG, "layer transport needs a layers geometry");
  }
  if (lg->max_records < 0 || (lg->max_records > 0 && !lg->records))
    return fail(MT_E_ARG, "records buffer missing");
  if ((lg->mask & MT_LOG_VACMAP) && !lg->vacmap) return fail(MT_E_ARG, "vacmap buffer missing");
  return run_cascades(ctx, pk, o, lg, out);
}

static_assert(sizeof(LogRecord) == sizeof(mt_record) && sizeof(mt_record) == 80, "record layout");

// ------------------------------------------------------------ run pieces
// (shared by mt_run_cascades, mt_run_cascades_log and mt_run_cascades_noalloc)

// bytes per slot of each queue array (px py pz dx dy dz E key mt: 8, hist: 4, ready: 8)
constexpr size_t kQueueFieldBytes[kQueueArrays] = {8, 8, 8, 8, 8, 8, 8, 8, 8, 4, 8};

static int validate_run(const mt_ctx* ctx, const mt_primaries* pk, const mt_run_opts* o, const mt_tallies* out) {
  if (!ctx || !pk || !o || !out) return fail(MT_E_ARG, "NULL argument");
  if (!ctx->has_target) return fail(MT_E_STATE, "no target set");
  if (pk->n_hist < 0 || pk->n_hist > (1ll << 31)) return fail(MT_E_LIMIT, "n_hist out of range");
  if (pk->hist0 < 0 || pk->hist0 + pk->n_hist > (1ll << 32)) return fail(MT_E_LIMIT, "history index out of range");
  if (!pk->Z || !pk->m || !pk->E || !pk->pos || !pk->dir) return fail(MT_E_ARG, "NULL primaries array");
  if (o->arith < MT_FP64 || o->arith > MT_FP32_SURROGATE) return fail(MT_E_ARG, "unknown arithmetic mode");
  if (o->arith == MT_FP32_SURROGATE && !ctx->mlp_ready) return fail(MT_E_STATE, "no surrogate installed");
  if (o->arith == MT_FP32_SURROGATE && ctx->tgt.potential != MT_UNIVERSAL)
    return fail(MT_E_ARG, "the surrogate is trained on the universal potential");
  if (o->range_nb < 1 || o->lat_nb < 1 || o->vac_nb < 1) return fail(MT_E_ARG, "histograms need >= 1 bin");
  if (!(o->range_bw > 0.0) || !(o->lat_bw > 0.0) || !(o->vac_bw > 0.0)) return fail(MT_E_ARG, "bin widths must be positive");
  return MT_OK;
}

// finite positive energies and positions, non-zero directions; the batch's
// total and largest primary energy
static int check_primaries(const mt_primaries* pk, double& E0, double& e_max) {
  const long long n_entries = pk->pka_uniform ? 1 : pk->n_hist;
  E0 = 0.0;
  e_max = 0.0;
  for (long long i = 0; i < n_entries; ++i) {
    if (!(pk->E[i] > 0.0) || !std::isfinite(pk->E[i])) return fail(MT_E_ARG, "primary energy must be positive and finite");
    const double* q = pk->pos + 3 * i;
    const double* d = pk->dir + 3 * i;
    if (!std::isfinite(q[0]) || !std::isfinite(q[1]) || !std::isfinite(q[2]))
      return fail(MT_E_ARG, "primary position must be finite");
    const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (!(dd > 0.0) || !std::isfinite(dd)) return fail(MT_E_ARG, "primary direction must be finite and non-zero");
    E0 += pk->E[i];
    e_max = std::max(e_max, pk->E[i]);
  }
  if (pk->pka_uniform) E0 = pk->E[0] * double(pk->n_hist);
  return MT_OK;
}

static bool perhist_of(const mt_run_opts* o, const mt_log* lg) { return o->per_history != 0 || (lg && lg->mask != 0); }

struct RunShape {
  int block = 0, cap = 0, grid = 0, waves = 0, lds_hist = 0;
  size_t lds = 0;
};

// launch shape of the persistent kernel: the context's tuning, else the shape
// measured best for Cu -> Cu cascades (two waves per SIMD, 128-entry LDS
// stacks).  The leading vacancy/replacement depth bins are counted in LDS, as
// many as fit without costing a resident block: the requested blocks per CU
// (or what the registers, tables and stacks allow, if fewer) decide, and bins
// beyond those in LDS are counted in HBM.
static int run_shape(const mt_ctx* ctx, const mt_run_opts* o, size_t table_bytes, bool perhist, RunShape& s) {
  s.block = ctx->tune.block ? ctx->tune.block : 256;
  s.cap = ctx->tune.stack_cap ? ctx->tune.stack_cap : 128;
  const int nwb = s.block / 64;
  const size_t base = table_bytes + size_t(nwb) * cascade_stack_bytes(s.cap);
  const size_t lds_max = size_t(ctx->lds_max);
  if (base > lds_max) return fail(MT_E_LIMIT, "tables and stacks exceed the LDS of a workgroup");
  const int arith = o->arith, pot = ctx->tgt.potential, geo = ctx->tgt.geo.kind;
  const int want = ctx->tune.blocks_per_cu > 0 ? ctx->tune.blocks_per_cu : kDefaultWavesPerCu * 64 / s.block;
  // LDS bytes of nl bins of each histogram (16-byte multiples)
  auto hist_bytes = [](int nl) { return (2 * size_t(nl) * sizeof(uint32_t) + 15) & ~size_t(15); };
  int nl = ctx->tune.lds_hist < 0 ? 0 : o->vac_nb + 2;
  if (ctx->tune.lds_hist > 0) nl = std::min(nl, ctx->tune.lds_hist);
  while (nl > 0 && base + hist_bytes(nl) > lds_max) nl /= 2;
  if (nl > 0) {
    int top = 0, with = 0;
    HIPCHECK(cascade_occupancy(arith, pot, geo, perhist, s.block, base, &top));
    const int target = std::min(top, want);
    HIPCHECK(cascade_occupancy(arith, pot, geo, perhist, s.block, base + hist_bytes(nl), &with));
    if (with < target) {
      // the most bins that keep `target` blocks: occupancy falls as LDS grows
      int lo = 0, hi = nl;
      while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        HIPCHECK(cascade_occupancy(arith, pot, geo, perhist, s.block, base + hist_bytes(mid), &with));
        if (with >= target)
          lo = mid;
        else
          hi = mid;
      }
      nl = lo;
    }
  }
  s.lds_hist = nl;
  s.lds = base + (nl > 0 ? hist_bytes(nl) : 0);
  int bpc = 0;
  HIPCHECK(cascade_occupancy(arith, pot, geo, perhist, s.block, s.lds, &bpc));
  bpc = std::min(bpc, want);
  if (bpc < 1) return fail(MT_E_LIMIT, "cascade kernel does not fit on a CU");
  s.grid = bpc * ctx->cus;
  s.waves = s.grid * nwb;
  if (s.waves > 0xffff) return fail(MT_E_LIMIT, "too many waves for the idle counter");
  return MT_OK;
}

// flights per cascade before any has completed: the Kinchin-Pease count
// E / (2 Ed) of the most energetic primary against the lowest threshold
static double flights_per_cascade0(const mt_ctx* ctx, double e_max) {
  double ed_min = 1e300;
  for (const HostMaterial& m : ctx->tgt.mat)
    for (const HostElement& e : m.el) ed_min = std::min(ed_min, e.Edisp);
  return 1.0 + e_max / (2.0 * std::max(ed_min, 1.0));
}

// device memory of one run
struct RunMem {
  const unsigned char* tables = nullptr;
  int table_bytes = 0;
  const int* pka_sp = nullptr;
  const double *pka_E = nullptr, *pka_pos = nullptr, *pka_dir = nullptr;
  void* q[kQueueArrays] = {};
  uint64_t qcap = 0;
  unsigned long long* res = nullptr;   // kResWords: control words, counters, energies
  unsigned long long* hist = nullptr;  // the five histograms, at hoff[0..4]
  double* wave_energy = nullptr;
  int* h_out = nullptr;
};

// the kernel parameters every run shares (per-history arrays, event streams
// and the surrogate are set by the caller)
static void fill_params(Params& P, const mt_ctx* ctx, const mt_primaries* pk, const mt_run_opts* o,
                        const RunMem& M, const RunShape& sh, double f_init, const size_t* hoff) {
  std::memset(&P, 0, sizeof(P));
  P.tables = M.tables;
  P.table_bytes = M.table_bytes;
  P.stack_cap = sh.cap;
  P.geo = ctx->tgt.geo;
  P.tau = ctx->tgt.tau;
  P.cw = ctx->tgt.cw;
  P.invscale = 1.0 / ctx->tgt.scale;
  P.Ef = pk->Ef;
  P.follow_max_gen = o->follow_max_gen;
  P.kp = o->kp;
  P.n_hist = pk->n_hist;
  P.hist0 = pk->hist0;
  P.pka_uniform = pk->pka_uniform ? 1 : 0;
  P.pka_species = M.pka_sp;
  P.pka_E = M.pka_E;
  P.pka_pos = M.pka_pos;
  P.pka_dir = M.pka_dir;
  P.seed = pk->seed;
  P.ctrl = M.res + kResCtrl;
  Queue& Q = P.q;
  Q.px = static_cast<uint64_t*>(M.q[0]);
  Q.py = static_cast<uint64_t*>(M.q[1]);
  Q.pz = static_cast<uint64_t*>(M.q[2]);
  Q.dx = static_cast<uint64_t*>(M.q[3]);
  Q.dy = static_cast<uint64_t*>(M.q[4]);
  Q.dz = static_cast<uint64_t*>(M.q[5]);
  Q.E = static_cast<uint64_t*>(M.q[6]);
  Q.key = static_cast<uint64_t*>(M.q[7]);
  Q.mt = static_cast<uint64_t*>(M.q[8]);
  Q.hist = static_cast<uint32_t*>(M.q[9]);
  Q.ready = static_cast<uint64_t*>(M.q[kQueueReady]);
  // shards of at least 2^12 slots, up to 16
  int log2cap = 0;
  while ((uint64_t(1) << log2cap) < M.qcap) ++log2cap;
  int nq = kMaxShards;
  while (nq > 1 && (M.qcap / uint64_t(nq)) < 4096) nq >>= 1;
  int lq = 0;
  while ((1 << lq) < nq) ++lq;
  Q.nq = nq;
  Q.sub_log2 = log2cap - lq;
  Q.sub_mask = (uint64_t(1) << Q.sub_log2) - 1;
  P.admit_limit = (long long)(M.qcap / 4);
  P.budget = double(M.qcap) / 2.0;
  P.f_init = f_init;
  P.c_floor = sh.waves;
  P.h_out = M.h_out;
  P.max_ic = 10000000;
  P.max_wait_ticks = (unsigned long long)(60.0 * ctx->clock_khz * 1000.0);
  P.counters = M.res + kResCounters;
  P.wave_energy = M.wave_energy;
  P.energy = reinterpret_cast<double*>(M.res + kResEnergy);
  P.range_hist = M.hist + hoff[0];
  P.lat_y_hist = M.hist + hoff[1];
  P.lat_z_hist = M.hist + hoff[2];
  P.vac_hist = M.hist + hoff[3];
  P.repl_hist = M.hist + hoff[4];
  P.range_x0 = o->range_x0;
  P.range_bw = o->range_bw;
  P.range_nb = o->range_nb;
  P.lat_y0 = o->lat_y0;
  P.lat_z0 = o->lat_z0;
  P.lat_bw = o->lat_bw;
  P.lat_nb = o->lat_nb;
  P.vac_x0 = o->vac_x0;
  P.vac_bw = o->vac_bw;
  P.vac_ibw = 1.0 / o->vac_bw;
  P.vac_nb = o->vac_nb;
  P.lds_hist = sh.lds_hist;
  P.primary_waves64 = ctx->tune.primary_waves64;
  if (o->arith == MT_FP32_SURROGATE) {
    P.mlp.p = ctx->mlp_p.as<float>();
    P.mlp.w2bf = ctx->mlp_w2bf.as<uint32_t>();
    P.mlp.n = ctx->mlp_norm;
    P.mlp.bf16 = ctx->mlp_bf16;
  }
}

// counters, energies, queue figures and the mean shader clock of a finished
// run's result block (clock_khz: the wall clock's rate)
static void read_results(const unsigned long long* sr, double E0, double clock_khz, mt_tallies* out) {
  const unsigned long long* ctrl = sr + kResCtrl;
  const unsigned long long* cnt = sr + kResCounters;
  const double* en = reinterpret_cast<const double*>(sr + kResEnergy);
  const double* rsum = reinterpret_cast<const double*>(sr + kResReduced);
  out->n_coll = (long long)cnt[cColl];
  out->n_flight = (long long)cnt[cFlight];
  out->vac = (long long)cnt[cVac];
  out->repl = (long long)cnt[cRepl];
  out->subst = (long long)cnt[cSubst];
  out->inter = (long long)cnt[cInter];
  out->lost = (long long)cnt[cLost];
  out->vacuum = (long long)cnt[cVacuum];
  out->rutherford = (long long)cnt[cRuth];
  out->newton = (long long)cnt[cNewton];
  out->disp = (long long)cnt[cDisp];
  out->range_n = (long long)cnt[cRangeN];
  out->surrogate = (long long)cnt[cSurr];
  out->coll_primary = (long long)cnt[cCollPrimary];
  out->E0 = E0;
  out->Eel = rsum[0] + rsum[1];
  out->Enuc = rsum[2] + rsum[3];
  out->Eunf = en[eEunf];
  out->Elost = en[eElost];
  out->Edrop = en[eEdrop];
  out->Evac = en[eEvac];
  out->range_sum = en[eRangeSum];
  out->range_sum2 = en[eRangeSum2];
  out->queue_peak = (long long)ctrl[kQueuePeak];
  out->queue_ops = (long long)ctrl[kQueueOps];
  out->sclk_mhz = ctrl[kWallSum] > 0 ? double(ctrl[kClkSum]) / double(ctrl[kWallSum]) * clock_khz * 1e-3 : 0.0;
}

// MT_E_DEVICE with the names of the device error flags, if any were raised
static int device_errors(const unsigned long long* ctrl) {
  const unsigned errbits = unsigned(ctrl[kError]);
  if (!errbits) return MT_OK;
  std::string m = "device error flags:";
  if (errbits & kErrQueueFull) m += " queue-full";
  if (errbits & kErrRunaway) m += " runaway-flight";
  if (errbits & kErrBadEntry) m += " bad-queue-entry";
  if (errbits & kErrTimeout) m += " wait-timeout";
  if (errbits & kErrNewton) m += " newton-cap";
  if (errbits & kErrNonFinite) m += " non-finite-energy";
  return fail(MT_E_DEVICE, m);
}

static int run_cascades(mt_ctx* ctx, const mt_primaries* pk, const mt_run_opts* o, mt_log* lg, mt_tallies* out) {
  if (int rc = validate_run(ctx, pk, o, out)) return rc;
  HIPCHECK(hipSetDevice(ctx->device));
  const long long n_hist = pk->n_hist;
  const long long n_entries = pk->pka_uniform ? 1 : n_hist;
  double E0 = 0.0, e_max = 0.0;
  if (int rc = check_primaries(pk, E0, e_max)) return rc;
  std::vector<Species> species;
  std::vector<int> pka_species;
  std::string msg = collect_species(ctx->tgt, pk->Z, pk->m, n_entries, species, pka_species);
  if (!msg.empty()) return fail(MT_E_LIMIT, msg);
  std::vector<unsigned char> blob;
  msg = build_tables(ctx->coef, ctx->tgt, species, blob);
  if (!msg.empty()) return fail(MT_E_LIMIT, msg);

  hipStream_t st = ctx->stream;
  HIPCHECK(ctx->tables.ensure(blob.size()));
  HIPCHECK(hipMemcpyAsync(ctx->tables.p, blob.data(), blob.size(), hipMemcpyHostToDevice, st));
  const size_t ne = size_t(std::max<long long>(n_entries, 1));
  HIPCHECK(ctx->pka_sp.ensure(ne * sizeof(int)));
  HIPCHECK(ctx->pka_E.ensure(ne * sizeof(double)));
  HIPCHECK(ctx->pka_pos.ensure(3 * ne * sizeof(double)));
  HIPCHECK(ctx->pka_dir.ensure(3 * ne * sizeof(double)));
  HIPCHECK(hipMemcpyAsync(ctx->pka_sp.p, pka_species.data(), n_entries * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(ctx->pka_E.p, pk->E, n_entries * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(ctx->pka_pos.p, pk->pos, 3 * n_entries * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(ctx->pka_dir.p, pk->dir, 3 * n_entries * sizeof(double), hipMemcpyHostToDevice, st));

  RunShape sh;
  if (int rc = run_shape(ctx, o, blob.size(), perhist_of(o, lg), sh)) return rc;
  const int grid = sh.grid, block = sh.block, waves = sh.waves;
  const size_t lds = sh.lds;
  const int pot = ctx->tgt.potential;
  const bool log_on = lg && lg->mask != 0;
  const bool perhist = perhist_of(o, lg);
  const double f_init = flights_per_cascade0(ctx, e_max);

  // queue: room for every flight of the batch up to 2^27 entries (80 B each)
  const double want = std::max(double(n_hist), 1.0) * f_init;
  uint64_t qcap = ctx->tune.queue_cap
                      ? uint64_t(ctx->tune.queue_cap)
                      : next_pow2(uint64_t(std::min(std::max(want, 65536.0), double(uint64_t(1) << 27))));
  if (ctx->qcap != qcap) {
    for (int i = 0; i < kQueueArrays; ++i) HIPCHECK(ctx->q[i].ensure(qcap * kQueueFieldBytes[i]));
    ctx->qcap = qcap;
  }
  // tallies
  // control words, counters, energies and the flushed energy sums share one
  // allocation (res); the five histograms share another (hist_all); both are
  // copied back to pinned staging in one transfer each
  const int nb[5] = {o->range_nb + 2, o->lat_nb + 2, o->lat_nb + 2, o->vac_nb + 2, o->vac_nb + 2};
  size_t hoff[6] = {0};
  for (int i = 0; i < 5; ++i) hoff[i + 1] = hoff[i] + size_t(nb[i]);
  const size_t hist_bytes = hoff[5] * sizeof(unsigned long long);
  HIPCHECK(ctx->hist_all.ensure(hist_bytes));
  HIPCHECK(ctx->res.ensure(kResWords * sizeof(unsigned long long)));
  HIPCHECK(ctx->wave_energy.ensure(size_t(waves) * 2 * kNumEnergy * sizeof(double)));
  HIPCHECK(ctx->stage_res.ensure(kResWords * sizeof(unsigned long long)));
  HIPCHECK(ctx->stage_hist.ensure(hist_bytes));
  const size_t nh = size_t(std::max<long long>(n_hist, 1));
  const size_t per_sz[8] = {4, 8, 4, 8, 8, 8, 4, 8};
  ZeroSpans zs{};
  auto span = [&zs](void* p, size_t bytes) {
    zs.p[zs.n] = p;
    zs.bytes[zs.n] = bytes;
    ++zs.n;
  };
  span(ctx->q[kQueueReady].p, qcap * sizeof(uint64_t));
  span(ctx->res.p, kResWords * sizeof(unsigned long long));
  span(ctx->hist_all.p, hist_bytes);
  span(ctx->wave_energy.p, size_t(waves) * 2 * kNumEnergy * sizeof(double));
  HIPCHECK(ctx->h_out.ensure(nh * sizeof(int)));
  span(ctx->h_out.p, nh * sizeof(int));
  if (perhist) {
    for (int i = 0; i < 8; ++i) {
      HIPCHECK(ctx->per[i].ensure(nh * per_sz[i]));
      span(ctx->per[i].p, (nh * per_sz[i] + 3) & ~size_t(3));
    }
  }

  RunMem M;
  M.tables = ctx->tables.as<unsigned char>();
  M.table_bytes = int(blob.size());
  M.pka_sp = ctx->pka_sp.as<int>();
  M.pka_E = ctx->pka_E.as<double>();
  M.pka_pos = ctx->pka_pos.as<double>();
  M.pka_dir = ctx->pka_dir.as<double>();
  for (int i = 0; i < kQueueArrays; ++i) M.q[i] = ctx->q[i].p;
  M.qcap = qcap;
  M.res = ctx->res.as<unsigned long long>();
  M.hist = ctx->hist_all.as<unsigned long long>();
  M.wave_energy = ctx->wave_energy.as<double>();
  M.h_out = ctx->h_out.as<int>();
  Params P;
  fill_params(P, ctx, pk, o, M, sh, f_init, hoff);
  unsigned long long* res = M.res;
  if (perhist) {
    P.h_vac = ctx->per[0].as<int>();
    P.h_coll = ctx->per[1].as<unsigned long long>();
    P.h_flights = ctx->per[2].as<int>();
    P.h_x = ctx->per[3].as<double>();
    P.h_y = ctx->per[4].as<double>();
    P.h_z = ctx->per[5].as<double>();
    P.h_state = ctx->per[6].as<int>();
    P.h_Eel = ctx->per[7].as<double>();
  }
  if (log_on) {
    P.log.mask = lg->mask;
    if (lg->max_records > 0) {
      HIPCHECK(ctx->log_rec.ensure(size_t(lg->max_records) * sizeof(LogRecord)));
      P.log.rec = ctx->log_rec.as<LogRecord>();
      P.log.cap = (unsigned long long)lg->max_records;
    }
    for (int i = 0; i < 3; ++i) P.log.vz[i] = lg->vacmap_Z[i];
    if (lg->mask & MT_LOG_VACMAP) {
      HIPCHECK(ctx->log_map.ensure(sizeof(int) * kVacMapN * kVacMapN * 3));
      P.log.vacmap = ctx->log_map.as<int>();
      span(ctx->log_map.p, sizeof(int) * kVacMapN * kVacMapN * 3);
    }
    if (lg->mask & MT_LOG_TRANSPORT) {
      const size_t nb = sizeof(long long) * size_t(ctx->tgt.geo.n_mat) * kTransportZ;
      HIPCHECK(ctx->log_tr.ensure(nb));
      P.log.transport = ctx->log_tr.as<unsigned long long>();
      span(ctx->log_tr.p, nb);
    }
  }
  if (o->arith == MT_FP32_SURROGATE) {
    P.mlp.p = ctx->mlp_p.as<float>();
    P.mlp.w2bf = ctx->mlp_w2bf.as<uint32_t>();
    P.mlp.n = ctx->mlp_norm;
    P.mlp.bf16 = ctx->mlp_bf16;
  }

  // reset -> [event] -> persistent kernel -> [event] -> energy flush -> copies
  // to pinned staging: replayed from a cached graph, or launched one by one
  double* red = reinterpret_cast<double*>(res + kResReduced);
  if (ctx->tune.graph >= 0 && !log_on) {
    const void* kfn = cascade_function(o->arith, pot, ctx->tgt.geo.kind, perhist);
    const RunGraphKey key{kfn, grid, block, lds, ctx->res.p, ctx->hist_all.p, hist_bytes, ctx->stage_res.p,
                          ctx->stage_hist.p};
    hipKernelNodeParams zp{}, kp{}, fp{};
    void* zargs[1] = {&zs};
    zp.func = const_cast<void*>(zero_spans_function());
    zp.gridDim = dim3(zero_spans_grid(zs));
    zp.blockDim = dim3(256);
    zp.kernelParams = zargs;
    void* kargs[1] = {&P};
    kp.func = const_cast<void*>(kfn);
    kp.gridDim = dim3(grid);
    kp.blockDim = dim3(block);
    kp.sharedMemBytes = (unsigned)lds;
    kp.kernelParams = kargs;
    const double* we_p = ctx->wave_energy.as<double>();
    int waves_arg = waves;
    void* fargs[3] = {&we_p, &waves_arg, &red};
    fp.func = const_cast<void*>(flush_energy_function());
    fp.gridDim = dim3(1);
    fp.blockDim = dim3(64);
    fp.kernelParams = fargs;
    RunGraph& G = ctx->graph;
    if (!G.exec || !(G.key == key)) {
      G.release();
      HIPCHECK(hipGraphCreate(&G.graph, 0));
      hipGraphNode_t prev, n;
      HIPCHECK(hipGraphAddKernelNode(&G.n_zero, G.graph, nullptr, 0, &zp));
      HIPCHECK(hipGraphAddEventRecordNode(&n, G.graph, &G.n_zero, 1, ctx->ev0));
      prev = n;
      HIPCHECK(hipGraphAddKernelNode(&G.n_kern, G.graph, &prev, 1, &kp));
      HIPCHECK(hipGraphAddEventRecordNode(&n, G.graph, &G.n_kern, 1, ctx->ev1));
      prev = n;
      HIPCHECK(hipGraphAddKernelNode(&G.n_flush, G.graph, &prev, 1, &fp));
      HIPCHECK(hipGraphAddMemcpyNode1D(&n, G.graph, &G.n_flush, 1, ctx->stage_res.p, ctx->res.p,
                                       kResWords * sizeof(unsigned long long), hipMemcpyDeviceToHost));
      prev = n;
      HIPCHECK(hipGraphAddMemcpyNode1D(&n, G.graph, &prev, 1, ctx->stage_hist.p, ctx->hist_all.p, hist_bytes,
                                       hipMemcpyDeviceToHost));
      HIPCHECK(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
      G.key = key;
    } else {
      HIPCHECK(hipGraphExecKernelNodeSetParams(G.exec, G.n_zero, &zp));
      HIPCHECK(hipGraphExecKernelNodeSetParams(G.exec, G.n_kern, &kp));
      HIPCHECK(hipGraphExecKernelNodeSetParams(G.exec, G.n_flush, &fp));
    }
    HIPCHECK(hipGraphLaunch(G.exec, st));
  } else {
    HIPCHECK(zero_spans_launch(zs, st));
    HIPCHECK(hipEventRecord(ctx->ev0, st));
    HIPCHECK(cascade_launch(P, o->arith, pot, perhist, grid, block, lds, st));
    HIPCHECK(hipEventRecord(ctx->ev1, st));
    HIPCHECK(flush_energy_launch(ctx->wave_energy.as<double>(), waves, red, st));
    HIPCHECK(hipMemcpyAsync(ctx->stage_res.p, ctx->res.p, kResWords * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(ctx->stage_hist.p, ctx->hist_all.p, hist_bytes, hipMemcpyDeviceToHost, st));
  }
  HIPCHECK(hipStreamSynchronize(st));
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));

  const unsigned long long* sr = static_cast<const unsigned long long*>(ctx->stage_res.p);
  const unsigned long long* ctrl = sr + kResCtrl;
  read_results(sr, E0, ctx->clock_khz, out);
  ctx->last_t0 = ctrl[kT0];
  ctx->last_tadm = ctrl[kTAdmit];
  ctx->last_fadm = ctrl[kFlightsAtAdmit];
  out->admit_ms = ctx->last_tadm > ctx->last_t0 && ctx->last_t0 > 0
                      ? double(ctx->last_tadm - ctx->last_t0) / ctx->clock_khz
                      : 0.0;
  out->flights_at_admit = (long long)ctx->last_fadm;
  ctx->last_thalf = ctrl[kTHalf];
  ctx->last_fhalf = ctrl[kFlightsAtHalf];
  out->half_ms = ctx->last_thalf > ctx->last_t0 && ctx->last_t0 > 0
                     ? double(ctx->last_thalf - ctx->last_t0) / ctx->clock_khz
                     : 0.0;
  out->flights_at_half = (long long)ctx->last_fhalf;
  long long* hdst[5] = {out->range_hist, out->lat_y_hist, out->lat_z_hist, out->vac_hist, out->repl_hist};
  for (int i = 0; i < 5; ++i)
    if (hdst[i])
      std::memcpy(hdst[i], static_cast<const unsigned long long*>(ctx->stage_hist.p) + hoff[i],
                  nb[i] * sizeof(long long));
  if (log_on) {
    const long long n = (long long)ctrl[kLogCount];
    lg->n_records = n;
    const long long keep = std::min(n, lg->max_records);
    if (keep > 0)
      HIPCHECK(hipMemcpy(lg->records, ctx->log_rec.p, size_t(keep) * sizeof(LogRecord), hipMemcpyDeviceToHost));
    if (lg->mask & MT_LOG_VACMAP) {
      std::vector<int> m(kVacMapN * kVacMapN * 3);
      HIPCHECK(hipMemcpy(m.data(), ctx->log_map.p, m.size() * sizeof(int), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < m.size(); ++i) lg->vacmap[i] += m[i];
    }
    if (lg->mask & MT_LOG_TRANSPORT) {
      std::vector<long long> t(size_t(ctx->tgt.geo.n_mat) * kTransportZ);
      HIPCHECK(hipMemcpy(t.data(), ctx->log_tr.p, t.size() * sizeof(long long), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < t.size(); ++i) lg->transport[i] += t[i];
    }
  }
  if (perhist && o->per_history) {
    void* pdst[8] = {out->h_vac, out->h_coll, out->h_flights, out->h_x,
                     out->h_y,   out->h_z,    out->h_state,   out->h_Eel};
    for (int i = 0; i < 8; ++i)
      if (pdst[i]) HIPCHECK(hipMemcpy(pdst[i], ctx->per[i].p, size_t(n_hist) * per_sz[i], hipMemcpyDeviceToHost));
  }
  out->kernel_ms = ms;
  out->grid = grid;
  out->block = block;
  out->waves = waves;
  return device_errors(ctrl);
}

int mt_autotune(mt_ctx* ctx, const mt_primaries* pk, const mt_run_opts* o, long long sample, int reps,
                mt_tuning* best, double* best_rate) {
  if (!ctx || !pk || !o || !best) return fail(MT_E_ARG, "NULL argument");
  if (sample < 1 || reps < 1) return fail(MT_E_ARG, "need sample >= 1 and reps >= 1");
  mt_primaries s = *pk;
  s.n_hist = std::min(sample, pk->n_hist);
  if (s.n_hist < 1) return fail(MT_E_ARG, "no histories to sample");
  mt_run_opts ro = *o;
  ro.per_history = 0;
  const mt_tuning keep = ctx->tune;
  // one timed pass of a sample under tuning t (the best of `reps`): the rate
  // projected for a launch of all pk->n_hist histories (projected_rate: the
  // sample's ramp and drain kept, the rest at its steady rate), in
  // collisions per second; 0 when the shape does not fit (LDS, occupancy),
  // < 0 on any other error (rc set).  Ramp and drain take tens of ms each,
  // depend on the shape (a shape with more waves ramps longer; deep LDS
  // stacks hold work in a few waves at the end) and the drain varies by as
  // much from one launch to the next: a third of a sample launch of 3e9
  // collisions, a few per cent of a launch of 1e6 C2 histories.  On C5 one
  // shape's start-to-last-admission rate was 18.2e9 on 3.3e5 histories and
  // 20.3e9 on 1e6, another's 16.9e9 and 17.2e9.
  auto rate_of = [&](const mt_tuning& t, int& rc, mt_tallies* last, const mt_primaries* ps = nullptr) -> double {
    if (!ps) ps = &s;
    ctx->tune = t;
    double r = 0.0;
    for (int i = 0; i < reps; ++i) {
      mt_tallies out{};
      rc = mt_run_cascades(ctx, ps, &ro, &out);
      if (rc == MT_E_LIMIT) {
        rc = MT_OK;
        return 0.0;
      }
      if (rc != MT_OK) return -1.0;
      LaunchPhases ph;
      ph.kernel_s = 1e-3 * out.kernel_ms;
      ph.half_s = 1e-3 * out.half_ms;
      ph.admit_s = 1e-3 * out.admit_ms;
      ph.n_coll = out.n_coll;
      ph.n_flight = out.n_flight;
      ph.flights_half = out.flights_at_half;
      ph.flights_admit = out.flights_at_admit;
      ph.n_hist = ps->n_hist;
      r = std::max(r, projected_rate(ph, pk->n_hist));
      if (last) *last = out;
    }
    return r;
  };
  // feeder waves per 64 for a multiple f of the primaries' share of the
  // collisions (0: no roles)
  auto feeders = [](double share, double f) {
    const double k = 64.0 * std::min(1.0, f * share);
    return k < 0.5 ? 0 : std::min(63, std::max(1, int(std::lround(k))));
  };
  // the primaries' share of the collisions, from a run at the default shape
  mt_tuning t0{};
  t0.queue_cap = keep.queue_cap;
  t0.lds_hist = keep.lds_hist;
  t0.graph = keep.graph;
  int rc = MT_OK;
  mt_tallies probe{};
  if (rate_of(t0, rc, &probe) < 0.0) {
    ctx->tune = keep;
    return rc;
  }
  const double share = probe.n_coll > 0 ? double(probe.coll_primary) / double(probe.n_coll) : 0.0;
  // candidate shapes (threads per block, resident blocks per CU, LDS stack
  // entries per wave) around the region measured best on MI355X: two or three
  // waves per SIMD (a kernel whose registers or LDS do not allow the requested
  // blocks per CU runs at the most that fit) with 64 to 256 stack entries, one
  // with deep stacks; queue capacity stays at its default.  Each runs without wave roles and with feeder
  // waves for 2.5 times the primaries' share, the region where roles
  // measured within a few per cent of their best on every workload; the
  // feeder count is then refined at the fastest shape with roles.  Roles are
  // kept only when they beat the fastest shape without them by 3 %: a sample
  // is smaller than the launches it tunes, and where roles gain little on
  // the sample (cascades of thousands of flights) they have lost at full
  // size.
  static const int kCand[][3] = {{256, 2, 256}, {256, 2, 192}, {256, 2, 128}, {256, 2, 96}, {256, 2, 64},
                                 {128, 4, 128}, {128, 6, 128}, {128, 6, 96},  {128, 6, 64}, {64, 8, 128},
                                 {256, 1, 384}, {256, 1, 256}, {128, 3, 256}, {128, 3, 128}, {256, 3, 128},
                                 {256, 3, 96},  {128, 5, 128}};
  // below a 2 % share (heavy-ion cascades of thousands of flights) feeders only
  // add queue traffic: on C2 (1.2 %) roles measured within noise on the sample
  // and 1-4 % slower on full launches, three times out of three
  const int k_roles = share >= 0.02 ? feeders(share, 2.5) : 0;
  std::vector<TimedShape> timed[2];  // without, with roles
  // each shape with the depth bins split between LDS and HBM (as many in LDS
  // as fit) and, unless the caller set the split, with every bin in HBM: which
  // is faster depends on the workload and the shape (below)
  const int n_bins = keep.lds_hist == 0 ? 2 : 1;
  for (const auto& c : kCand) {
    for (int with = 0; with < (k_roles > 0 ? 2 : 1); ++with) {
      for (int b = 0; b < n_bins; ++b) {
        mt_tuning t = t0;
        t.block = c[0];
        t.blocks_per_cu = c[1];
        t.stack_cap = c[2];
        t.primary_waves64 = with ? k_roles : 0;
        if (b) t.lds_hist = -1;
        mt_tallies last{};
        const double r = rate_of(t, rc, &last);
        if (r < 0.0) {
          ctx->tune = keep;
          return rc;
        }
        if (r > 0.0) timed[with].push_back(TimedShape{t, r, ctx->cus > 0 ? last.waves / ctx->cus : 0});
      }
    }
  }
  // per group (without, with roles): the pick (pick_shape: the fastest, the
  // deepest stack within 2 % of it), and the pick of each occupancy within
  // 10 % of the group's fastest.  Occupancy is the waves per CU a shape ran
  // at, which the registers and the LDS (stacks, tables, bins) may hold below
  // the requested threads per block x blocks per CU.  Without roles, also the
  // runner-up of another shape at each occupancy: a sample of fewer histories
  // than one round of admissions (64 per wave at the start) is rated on its
  // kernel time, which varies by several per cent.
  double best_r[2] = {0.0, 0.0};
  mt_tuning best_t[2] = {keep, keep};
  std::vector<TimedShape> per_occ[2];
  for (int w = 0; w < 2; ++w) {
    const int p = pick_shape(timed[w]);
    if (p >= 0) {
      best_t[w] = timed[w][size_t(p)].t;
      best_r[w] = timed[w][size_t(p)].rate;
    }
    per_occ[w] = occupancy_finalists(timed[w], w == 0);
  }
  if (best_r[0] <= 0.0 && best_r[1] <= 0.0) {
    ctx->tune = keep;
    return fail(MT_E_LIMIT, "no candidate launch shape fits");
  }
  // with roles, three more feeder counts at each occupancy's pick
  const int ks[3] = {feeders(share, 1.5), feeders(share, 4.0), 63};
  for (auto& e : per_occ[1]) {
    for (int i = 0; i < 3; ++i) {
      bool dup = ks[i] == k_roles;
      for (int j = 0; j < i; ++j) dup = dup || ks[j] == ks[i];
      if (dup) continue;
      mt_tuning t = e.t;
      t.primary_waves64 = ks[i];
      const double r = rate_of(t, rc, nullptr);
      if (r < 0.0) {
        ctx->tune = keep;
        return rc;
      }
      if (r > e.rate) {
        e.t = t;
        e.rate = r;
      }
    }
    if (e.rate > best_r[1]) {
      best_r[1] = e.rate;
      best_t[1] = e.t;
    }
  }
  // Second stage, where the primaries hold at least twice the sample: the
  // finalists are timed on up to four times as many histories, and those
  // timings decide.  A launch's queue keeps growing past the sample's size,
  // and shapes that spill more of their recoils to it lose ground there: on
  // C2 (Cu -> Cu 500 keV), three waves per SIMD led two by 10 % on a
  // 5e4-history sample and trailed by 6-12 % from 2e5 histories on, with
  // 2.6e6 entries per queue shard against 1.5e6, on some MI355X devices
  // (others ran both as on the sample).  The finalists are the picks of each
  // occupancy within 10 % of the fastest, without and with roles (C5's best,
  // 128 x 6 with 14 feeders, trailed 256 x 2 at the first feeder count).
  // Each finalist runs with its own depth-bin split and with the
  // other one, unless the caller set lds_hist: LDS atomics on the bins are in
  // order with the stack's LDS traffic and serialise on a hot bin, HBM
  // atomics are neither but contend in L2 where a few bins take most counts.
  // On C5 (vacancies spread over 1.5 um) the bins in HBM ran 8 % faster at
  // 256 x 3; on C2 (shallow, dense cascades) a third slower.  Otherwise the
  // two picks are timed once more each on the sample (a single timing varies
  // by a few per cent).
  mt_primaries s2 = *pk;
  s2.n_hist = std::min(4 * s.n_hist, pk->n_hist);
  const bool stage2 = s2.n_hist >= 2 * s.n_hist;
  // the better of a finalist with its depth-bin split and with the other one,
  // on the second-stage histories; < 0 on error (rc set)
  auto final_rate = [&](mt_tuning& t) -> double {
    double r = rate_of(t, rc, nullptr, &s2);
    if (r < 0.0 || keep.lds_hist != 0) return r;
    mt_tuning h = t;
    h.lds_hist = t.lds_hist == 0 ? -1 : 0;
    const double rh = rate_of(h, rc, nullptr, &s2);
    if (rh < 0.0) return rh;
    if (rh > r) {
      r = rh;
      t = h;
    }
    return r;
  };
  if (stage2) {
    for (int w = 0; w < 2; ++w) {
      double rw = 0.0;
      for (auto& e : per_occ[w]) {
        const double r = final_rate(e.t);
        if (r < 0.0) {
          ctx->tune = keep;
          return rc;
        }
        if (r > rw) {
          rw = r;
          best_t[w] = e.t;
        }
      }
      best_r[w] = rw;
    }
  } else if (best_r[0] > 0.0 && best_r[1] > 0.0) {
    for (int w = 0; w < 2; ++w) {
      const double r = rate_of(best_t[w], rc, nullptr);
      if (r < 0.0) {
        ctx->tune = keep;
        return rc;
      }
      best_r[w] = std::max(best_r[w], r);
    }
  }
  const int pick = best_r[0] <= 0.0 || (best_r[1] > 0.0 && best_r[1] > 1.03 * best_r[0]) ? 1 : 0;
  const double best_rate_v = best_r[pick];
  const mt_tuning best_v = best_t[pick];
  ctx->tune = best_v;
  *best = best_v;
  if (best_rate) *best_rate = best_rate_v;
  return MT_OK;
}

int mt_trim_flight(mt_ctx* ctx, mt_ion* ion, const double* u, long long n_u, long long* n_used, mt_event* events,
                   int max_events, int* n_events) {
  if (!ctx || !ion || !n_used || !n_events || (n_u > 0 && !u) || (max_events > 0 && !events))
    return fail(MT_E_ARG, "NULL argument");
  if (!ctx->has_target) return fail(MT_E_STATE, "no target set");
  if (ion->Z < 1 || ion->Z > 92 || !(ion->m > 0.0)) return fail(MT_E_ARG, "ion Z/m out of range");
  if (n_u < 1 || max_events < 1) return fail(MT_E_ARG, "need at least one draw and one event slot");
  HIPCHECK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const auto key = std::make_pair(ion->Z, ion->m);
  auto it = ctx->flight_tables.find(key);
  if (it == ctx->flight_tables.end()) {
    std::vector<Species> species;
    std::vector<int> ps;
    std::string msg = collect_species(ctx->tgt, &ion->Z, &ion->m, 1, species, ps);
    if (!msg.empty()) return fail(MT_E_LIMIT, msg);
    std::vector<unsigned char> blob;
    msg = build_tables(ctx->coef, ctx->tgt, species, blob);
    if (!msg.empty()) return fail(MT_E_LIMIT, msg);
    if (blob.size() > size_t(ctx->lds_max)) return fail(MT_E_LIMIT, "the ion's tables exceed the LDS of a workgroup");
    std::unique_ptr<DBuf> b(new DBuf);
    HIPCHECK(b->ensure(blob.size()));
    HIPCHECK(hipMemcpy(b->p, blob.data(), blob.size(), hipMemcpyHostToDevice));
    it = ctx->flight_tables.emplace(key, std::move(b)).first;
    ctx->flight_table_bytes[key] = int(blob.size());
  }
  // one copy in (the ion and the uniforms, staged in pinned memory) and one
  // synchronisation: the kernel writes the status words, the ion and the
  // events straight into mapped pinned memory
  const size_t u_off = (sizeof(mt_ion) + 15) & ~size_t(15);
  const size_t in_bytes = u_off + size_t(n_u) * sizeof(double);
  const size_t ion_off = 32, ev_off = (ion_off + sizeof(mt_ion) + 15) & ~size_t(15);
  const size_t out_bytes = ev_off + size_t(max_events) * sizeof(mt_event);
  HIPCHECK(ctx->f_in.ensure(in_bytes));
  HIPCHECK(ctx->f_stage.ensure(in_bytes));
  HIPCHECK(ctx->f_out.ensure(out_bytes));
  unsigned char* stage = static_cast<unsigned char*>(ctx->f_stage.p);
  std::memcpy(stage, ion, sizeof(mt_ion));
  std::memcpy(stage + u_off, u, size_t(n_u) * sizeof(double));
  HIPCHECK(hipMemcpyAsync(ctx->f_in.p, stage, in_bytes, hipMemcpyHostToDevice, st));
  unsigned char* out_h = static_cast<unsigned char*>(ctx->f_out.p);
  unsigned char* out_d = static_cast<unsigned char*>(ctx->f_out.dev);
  long long* used = reinterpret_cast<long long*>(out_h);
  used[2] = -1;  // overwritten by the kernel
  FlightParams P;
  std::memset(&P, 0, sizeof(P));
  P.tables = it->second->as<unsigned char>();
  P.table_bytes = ctx->flight_table_bytes[key];
  P.geo = ctx->tgt.geo;
  P.tau = ctx->tgt.tau;
  P.cw = ctx->tgt.cw;
  P.invscale = 1.0 / ctx->tgt.scale;
  P.ion = ctx->f_in.as<mt_ion>();
  P.u = reinterpret_cast<const double*>(ctx->f_in.as<unsigned char>() + u_off);
  P.n_u = n_u;
  P.ion_out = reinterpret_cast<mt_ion*>(out_d + ion_off);
  P.ev = reinterpret_cast<mt_event*>(out_d + ev_off);
  P.max_ev = max_events;
  P.n_used = reinterpret_cast<long long*>(out_d);
  HIPCHECK(flight_launch(P, ctx->tgt.potential, st));
  HIPCHECK(hipStreamSynchronize(st));
  *n_used = used[0];
  *n_events = int(used[1]);
  if (used[2] == 1) return fail(MT_E_MORE, "uniform stream exhausted");
  if (used[2] == 2) return fail(MT_E_MORE, "event buffer full");
  if (used[2] == 3) return fail(MT_E_DEVICE, "runaway flight");
  if (used[2] == 4) return fail(MT_E_DEVICE, "non-finite-energy");
  if (used[2] != 0) return fail(MT_E_DEVICE, "single-flight kernel left no status");
  std::memcpy(ion, out_h + ion_off, sizeof(mt_ion));
  if (used[1] > 0) std::memcpy(events, out_h + ev_off, size_t(used[1]) * sizeof(mt_event));
  return MT_OK;
}

// Training triples and scaling of the surrogate for the context's target:
// every (species, material, element) the cascades of `o->ion_*` touch, with
// the flight-energy range that keeps eps = fi E <= eps_max.
static std::string surrogate_setup(const mt_ctx* ctx, const mt_surrogate_opts* o, MlpNorm& nm,
                                   std::vector<MlpTriple>* out) {
  if (!ctx->has_target) return "no target set";
  if (ctx->tgt.potential != MT_UNIVERSAL) return "the surrogate is trained on the universal potential";
  if (o->n_ions < 1 || !o->ion_Z || !o->ion_m) return "need at least one ion species";
  if (!(o->E_min > 0.0) || !(o->E_max > o->E_min)) return "need 0 < E_min < E_max";
  if (!(o->eps_max > 0.0)) return "need eps_max > 0";
  if (o->inputs != 0 && o->inputs != 2 && o->inputs != 3) return "inputs must be 2 or 3";
  if (o->sampling != 0 && o->sampling != 1) return "sampling must be 0 or 1";
  if (o->sampling == 1 && !(o->eps_min > 0.0 && o->eps_min < o->eps_max)) return "need 0 < eps_min < eps_max";
  std::vector<Species> species;
  std::vector<int> ps;
  std::string msg = collect_species(ctx->tgt, o->ion_Z, o->ion_m, o->n_ions, species, ps);
  if (!msg.empty()) return msg;
  std::vector<unsigned char> blob;
  msg = build_tables(ctx->coef, ctx->tgt, species, blob);
  if (!msg.empty()) return msg;
  const TableHeader* H = reinterpret_cast<const TableHeader*>(blob.data());
  const MatConst* mat = reinterpret_cast<const MatConst*>(blob.data() + H->off_mat);
  const ElemConst* el = reinterpret_cast<const ElemConst*>(blob.data() + H->off_el);
  std::vector<MlpTriple> tri;
  double eps_lo = 1e300, eps_hi = 0.0, b_hi = 0.0;
  for (int sp = 0; sp < H->n_species; ++sp)
    for (int k = 0; k < H->n_mat; ++k) {
      const MatConst& mc = mat[sp * H->n_mat + k];
      for (int i = 0; i < mc.n_el; ++i) {
        const ElemConst& ec = el[sp * H->n_el + mc.el_begin + i];
        MlpTriple t{mc.a, mc.f, mc.epsdg, ec.ai, ec.fi, std::min(o->E_max, o->eps_max / ec.fi)};
        if (!(t.E_hi > o->E_min)) continue;
        tri.push_back(t);
        eps_lo = std::min(eps_lo, ec.fi * o->E_min);
        eps_hi = std::max(eps_hi, ec.fi * t.E_hi);
        // pmax falls with energy: the widest impact parameter is at E_min
        const double eeg = std::sqrt(o->E_min * mc.f * mc.epsdg);
        b_hi = std::max(b_hi, mc.a / (eeg + std::sqrt(eeg) + 0.125 * std::pow(eeg, 0.1)) / ec.ai);
      }
    }
  if (tri.empty()) return "no collision of these ions reaches the surrogate's domain";
  if (o->sampling == 1) {  // the rectangle: its own eps range, the cascades' widest b
    eps_lo = o->eps_min;
    eps_hi = o->eps_max;
  }
  const double l0 = std::log10(eps_lo), l1 = std::log10(eps_hi);
  const double b_floor = 1e-4 * b_hi;
  const double m0 = std::log10(b_floor), m1 = std::log10(b_hi);
  nm.e_mid = float(0.5 * (l0 + l1));
  nm.e_inv_half = float(2.0 / std::max(l1 - l0, 1e-6));
  nm.lb_mid = float(0.5 * (m0 + m1));
  // two inputs: the log b input is held at zero (its weights get no gradient)
  nm.lb_inv_half = o->inputs == 2 ? 0.0f : float(2.0 / (m1 - m0));
  nm.b_mid = float(0.5 * b_hi);
  nm.b_inv_half = float(2.0 / b_hi);
  nm.b_floor = float(b_floor);
  nm.y_off = -8.0f;
  nm.y_scale = 8.0f;
  nm.eps_lo = float(eps_lo);
  nm.eps_hi = float(eps_hi);
  nm.b_hi = float(b_hi);
  if (out) out->swap(tri);
  return "";
}

int mt_surrogate_norm(mt_ctx* ctx, const mt_surrogate_opts* o, float* norm) {
  if (!ctx || !o || !norm) return fail(MT_E_ARG, "NULL argument");
  MlpNorm nm;
  const std::string msg = surrogate_setup(ctx, o, nm, nullptr);
  if (!msg.empty()) return fail(MT_E_ARG, msg);
  static_assert(sizeof(MlpNorm) == kMlpNormFloats * sizeof(float), "MlpNorm is 12 floats");
  std::memcpy(norm, &nm, sizeof(nm));
  return MT_OK;
}

int mt_surrogate_train(mt_ctx* ctx, const mt_surrogate_opts* o, float* params, double* epoch_loss) {
  if (!ctx || !o || !params || !epoch_loss) return fail(MT_E_ARG, "NULL argument");
  MlpNorm nm;
  std::vector<MlpTriple> tri;
  std::string msg = surrogate_setup(ctx, o, nm, &tri);
  if (!msg.empty()) return fail(MT_E_ARG, msg);
  if (o->batch < 64 || o->batch > 65536 || o->epochs < 1 || o->n_train < o->batch || o->n_train > (1ll << 31))
    return fail(MT_E_ARG, "bad batch / epochs / n_train");
  if (!(o->lr > 0.0)) return fail(MT_E_ARG, "need lr > 0");
  HIPCHECK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const long long n = o->n_train;
  const int B = o->batch;
  const int steps = int(n / B);
  HIPCHECK(ctx->mlp_p.ensure(kMlpParams * sizeof(float)));
  HIPCHECK(ctx->mlp_X.ensure(size_t(n) * kMlpIn * sizeof(float)));
  HIPCHECK(ctx->mlp_T.ensure(size_t(n) * sizeof(float)));
  HIPCHECK(ctx->mlp_rec.ensure(size_t(B) * kMlpRecord * sizeof(float)));
  HIPCHECK(ctx->mlp_part.ensure(size_t((B + kMlpGradChunk - 1) / kMlpGradChunk) * (kMlpParams + 1) * sizeof(float)));
  HIPCHECK(ctx->mlp_g.ensure((kMlpParams + 1) * sizeof(float)));
  HIPCHECK(ctx->mlp_m.ensure(kMlpParams * sizeof(float)));
  HIPCHECK(ctx->mlp_v.ensure(kMlpParams * sizeof(float)));
  HIPCHECK(ctx->mlp_t.ensure(sizeof(int)));
  HIPCHECK(ctx->mlp_bl.ensure(size_t(steps) * sizeof(float)));
  HIPCHECK(ctx->mlp_tri.ensure(tri.size() * sizeof(MlpTriple)));
  HIPCHECK(hipMemcpyAsync(ctx->mlp_tri.p, tri.data(), tri.size() * sizeof(MlpTriple), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(ctx->mlp_p.p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemsetAsync(ctx->mlp_m.p, 0, kMlpParams * sizeof(float), st));
  HIPCHECK(hipMemsetAsync(ctx->mlp_v.p, 0, kMlpParams * sizeof(float), st));
  HIPCHECK(hipMemsetAsync(ctx->mlp_t.p, 0, sizeof(int), st));
  HIPCHECK(mlp_make_data(o->seed, n, ctx->mlp_tri.as<MlpTriple>(), int(tri.size()), o->E_min, o->sampling, nm,
                         ctx->mlp_X.as<float>(), ctx->mlp_T.as<float>(), st));
  HIPCHECK(mlp_train(ctx->mlp_p.as<float>(), ctx->mlp_X.as<float>(), ctx->mlp_T.as<float>(), n, B, o->epochs,
                     float(o->lr), ctx->mlp_rec.as<float>(), ctx->mlp_part.as<float>(), ctx->mlp_g.as<float>(),
                     ctx->mlp_m.as<float>(),
                     ctx->mlp_v.as<float>(), ctx->mlp_t.as<int>(), ctx->mlp_bl.as<float>(), epoch_loss, st));
  HIPCHECK(hipMemcpyAsync(params, ctx->mlp_p.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return MT_OK;
}

int mt_surrogate_install(mt_ctx* ctx, const mt_surrogate_opts* o, const float* params) {
  if (!ctx || !o || !params) return fail(MT_E_ARG, "NULL argument");
  MlpNorm nm;
  std::string msg = surrogate_setup(ctx, o, nm, nullptr);
  if (!msg.empty()) return fail(MT_E_ARG, msg);
  HIPCHECK(hipSetDevice(ctx->device));
  HIPCHECK(ctx->mlp_p.ensure(kMlpParams * sizeof(float)));
  HIPCHECK(ctx->mlp_w2bf.ensure(64 * 32 * sizeof(uint32_t)));
  HIPCHECK(hipMemcpyAsync(ctx->mlp_p.p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  HIPCHECK(mlp_pack(ctx->mlp_p.as<float>(), ctx->mlp_w2bf.as<uint32_t>(), ctx->stream));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  ctx->mlp_norm = nm;
  ctx->mlp_bf16 = o->bf16 ? 1 : 0;
  ctx->mlp_ready = true;
  return MT_OK;
}

int mt_surrogate_eval(mt_ctx* ctx, const double* eps, const double* b, long long n, double* s2) {
  if (!ctx || (n > 0 && (!eps || !b || !s2))) return fail(MT_E_ARG, "NULL argument");
  if (!ctx->mlp_ready) return fail(MT_E_STATE, "no surrogate installed");
  if (n <= 0) return MT_OK;
  HIPCHECK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  HIPCHECK(ctx->scratch_a.ensure(n * sizeof(double)));
  HIPCHECK(ctx->scratch_b.ensure(n * sizeof(double)));
  HIPCHECK(ctx->scratch_c.ensure(n * sizeof(double)));
  HIPCHECK(hipMemcpyAsync(ctx->scratch_a.p, eps, n * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(ctx->scratch_b.p, b, n * sizeof(double), hipMemcpyHostToDevice, st));
  MlpDev S;
  S.p = ctx->mlp_p.as<float>();
  S.w2bf = ctx->mlp_w2bf.as<uint32_t>();
  S.n = ctx->mlp_norm;
  S.bf16 = ctx->mlp_bf16;
  HIPCHECK(mlp_infer_launch(S, ctx->scratch_a.as<double>(), ctx->scratch_b.as<double>(), n,
                            ctx->scratch_c.as<double>(), st));
  HIPCHECK(hipMemcpyAsync(s2, ctx->scratch_c.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return MT_OK;
}

int mt_surrogate_ablate(mt_ctx* ctx, int mode, const double* eps, const double* b, long long n, int reps,
                        double* s2, double* kernel_ms) {
  if (!ctx || !kernel_ms || (n > 0 && (!eps || !b || !s2))) return fail(MT_E_ARG, "NULL argument");
  if (mode < 0 || mode > 3) return fail(MT_E_ARG, "mode must be 0 (MAGIC), 1 (FP32), 2 (BF16 dot) or 3 (MFMA)");
  if (reps < 1) return fail(MT_E_ARG, "reps must be >= 1");
  if (mode > 0 && !ctx->mlp_ready) return fail(MT_E_STATE, "no surrogate installed");
  *kernel_ms = 0.0;
  if (n <= 0) return MT_OK;
  HIPCHECK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  HIPCHECK(ctx->scratch_a.ensure(n * sizeof(double)));
  HIPCHECK(ctx->scratch_b.ensure(n * sizeof(double)));
  HIPCHECK(ctx->scratch_c.ensure(n * sizeof(double)));
  HIPCHECK(hipMemcpyAsync(ctx->scratch_a.p, eps, n * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(ctx->scratch_b.p, b, n * sizeof(double), hipMemcpyHostToDevice, st));
  MlpDev S{};
  if (ctx->mlp_ready) {
    S.p = ctx->mlp_p.as<float>();
    S.w2bf = ctx->mlp_w2bf.as<uint32_t>();
    S.n = ctx->mlp_norm;
  }
  // one untimed launch, then `reps` timed back to back
  HIPCHECK(ablate_launch(mode, S, ctx->scratch_a.as<double>(), ctx->scratch_b.as<double>(), n,
                         ctx->scratch_c.as<double>(), st));
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  hipError_t e = hipEventRecord(e0, st);
  for (int r = 0; r < reps && e == hipSuccess; ++r)
    e = ablate_launch(mode, S, ctx->scratch_a.as<double>(), ctx->scratch_b.as<double>(), n,
                      ctx->scratch_c.as<double>(), st);
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHECK(e);
  *kernel_ms = double(ms) / reps;
  HIPCHECK(hipMemcpyAsync(s2, ctx->scratch_c.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return MT_OK;
}

int mt_surrogate_grad(mt_ctx* ctx, const mt_surrogate_opts* o, const float* params, const double* eps,
                      const double* b, const double* ln_s2, int n, float* grad, double* loss) {
  if (!ctx || !o || !params || !eps || !b || !ln_s2 || !grad || !loss) return fail(MT_E_ARG, "NULL argument");
  if (n < 1 || n > 65536) return fail(MT_E_ARG, "batch out of range");
  MlpNorm nm;
  std::string msg = surrogate_setup(ctx, o, nm, nullptr);
  if (!msg.empty()) return fail(MT_E_ARG, msg);
  std::vector<float> X(kMlpIn * size_t(n)), T(n);
  for (int i = 0; i < n; ++i) {
    mlp_inputs(nm, float(eps[i]), float(b[i]), X[3 * i], X[3 * i + 1], X[3 * i + 2]);
    T[i] = float((ln_s2[i] - nm.y_off) / nm.y_scale);
  }
  HIPCHECK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  HIPCHECK(ctx->scratch_a.ensure(kMlpParams * sizeof(float)));
  HIPCHECK(ctx->scratch_b.ensure(X.size() * sizeof(float)));
  HIPCHECK(ctx->scratch_c.ensure(T.size() * sizeof(float) + 16));
  HIPCHECK(ctx->scratch_d.ensure(size_t(n) * kMlpRecord * sizeof(float)));
  HIPCHECK(ctx->scratch_e.ensure((kMlpParams + 1) * sizeof(float)));
  HIPCHECK(ctx->mlp_part.ensure(size_t((n + kMlpGradChunk - 1) / kMlpGradChunk) * (kMlpParams + 1) * sizeof(float)));
  HIPCHECK(hipMemcpyAsync(ctx->scratch_a.p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(ctx->scratch_b.p, X.data(), X.size() * sizeof(float), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(ctx->scratch_c.p, T.data(), T.size() * sizeof(float), hipMemcpyHostToDevice, st));
  float* g = ctx->scratch_e.as<float>();
  HIPCHECK(mlp_batch_grad(ctx->scratch_a.as<float>(), ctx->scratch_b.as<float>(), ctx->scratch_c.as<float>(), n,
                          ctx->scratch_d.as<float>(), ctx->mlp_part.as<float>(), g, g + kMlpParams, st));
  float lf = 0.0f;
  HIPCHECK(hipMemcpyAsync(grad, g, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(&lf, g + kMlpParams, sizeof(float), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  *loss = lf;
  return MT_OK;
}

// mt_run_cascades with every device buffer carved from the device's static
// scratch: no allocation of any kind (no device, pinned or graph memory); the
// kernels are launched one by one and the results copied straight into the
// caller's arrays.  Hot kernel only (no per-history records or streams).
int mt_run_cascades_noalloc(mt_ctx* ctx, const mt_primaries* pk, const mt_run_opts* o, mt_tallies* out) {
  if (int rc = validate_run(ctx, pk, o, out)) return rc;
  if (o->per_history) return fail(MT_E_ARG, "mt_run_cascades_noalloc keeps no per-history records");
  if (pk->n_hist > kNaMaxHist) return fail(MT_E_LIMIT, "more than 2^22 histories for the static scratch");
  const long long n_entries = pk->pka_uniform ? 1 : pk->n_hist;
  if (n_entries > kNaMaxEntries)
    return fail(MT_E_LIMIT, "more than 2^18 per-history primaries for the static scratch");
  const int nb[5] = {o->range_nb + 2, o->lat_nb + 2, o->lat_nb + 2, o->vac_nb + 2, o->vac_nb + 2};
  size_t hoff[6] = {0};
  for (int i = 0; i < 5; ++i) hoff[i + 1] = hoff[i] + size_t(nb[i]);
  if ((size_t(kResWords) + hoff[5]) * sizeof(unsigned long long) > kNaResHistBytes)
    return fail(MT_E_LIMIT, "histograms exceed the static scratch (about 130 000 bins)");
  if (!ctx->na_scratch) return fail(MT_E_STATE, "no static scratch on this device");
  HIPCHECK(hipSetDevice(ctx->device));
  double E0 = 0.0, e_max = 0.0;
  if (int rc = check_primaries(pk, E0, e_max)) return rc;
  Species sp[kMaxSpecies];
  int ns = 0;
  std::string msg = collect_species_into(ctx->tgt, pk->Z, pk->m, n_entries, sp, &ns);
  if (!msg.empty()) return fail(MT_E_LIMIT, msg);
  size_t tb = 0;
  msg = build_tables_into(ctx->coef, ctx->tgt, sp, ns, ctx->na_blob.data(), ctx->na_blob.size(), &tb);
  if (!msg.empty()) return fail(MT_E_LIMIT, msg);
  RunShape sh;
  if (int rc = run_shape(ctx, o, tb, false, sh)) return rc;
  if (sh.waves > kNaMaxWaves) return fail(MT_E_LIMIT, "more than 16384 waves for the static scratch");
  const double f_init = flights_per_cascade0(ctx, e_max);

  unsigned char* base = static_cast<unsigned char*>(ctx->na_scratch);
  int* d_sp = reinterpret_cast<int*>(base + kNaOffPka);
  int* d_Z = d_sp + kNaMaxEntries;
  double* d_m = reinterpret_cast<double*>(d_Z + kNaMaxEntries);
  double* d_E = d_m + kNaMaxEntries;
  double* d_pos = d_E + kNaMaxEntries;
  double* d_dir = d_pos + 3 * kNaMaxEntries;
  RunMem M;
  M.tables = base + kNaOffTables;
  M.table_bytes = int(tb);
  M.pka_sp = d_sp;
  M.pka_E = d_E;
  M.pka_pos = d_pos;
  M.pka_dir = d_dir;
  size_t qo = kNaOffQueue;
  for (int i = 0; i < kQueueArrays; ++i) {
    M.q[i] = base + qo;
    qo += size_t(kNaQueueSlots) * kQueueFieldBytes[i];
  }
  M.qcap = kNaQueueSlots;
  M.res = reinterpret_cast<unsigned long long*>(base + kNaOffRes);
  M.hist = M.res + kResWords;
  M.wave_energy = reinterpret_cast<double*>(base + kNaOffWave);
  M.h_out = reinterpret_cast<int*>(base + kNaOffHout);

  hipStream_t st = ctx->stream;
  HIPCHECK(hipMemcpyAsync(base + kNaOffTables, ctx->na_blob.data(), tb, hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(d_Z, pk->Z, size_t(n_entries) * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(d_m, pk->m, size_t(n_entries) * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(d_E, pk->E, size_t(n_entries) * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(d_pos, pk->pos, 3 * size_t(n_entries) * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(d_dir, pk->dir, 3 * size_t(n_entries) * sizeof(double), hipMemcpyHostToDevice, st));
  SpeciesList L;
  std::memset(&L, 0, sizeof(L));
  L.n = ns;
  for (int k = 0; k < ns; ++k) {
    L.Z[k] = sp[k].Z;
    L.m[k] = sp[k].m;
  }
  HIPCHECK(primary_species_launch(L, d_Z, d_m, n_entries, d_sp, st));

  // what must start at zero: the previous call's contents are still there
  ZeroSpans zs{};
  auto span = [&zs](void* p, size_t bytes) {
    zs.p[zs.n] = p;
    zs.bytes[zs.n] = bytes;
    ++zs.n;
  };
  span(M.q[kQueueReady], size_t(kNaQueueSlots) * sizeof(uint64_t));
  span(M.res, (size_t(kResWords) + hoff[5]) * sizeof(unsigned long long));
  span(M.wave_energy, size_t(sh.waves) * 2 * kNumEnergy * sizeof(double));
  span(M.h_out, size_t(std::max<long long>(pk->n_hist, 1)) * sizeof(int));
  Params P;
  fill_params(P, ctx, pk, o, M, sh, f_init, hoff);
  double* red = reinterpret_cast<double*>(M.res + kResReduced);
  HIPCHECK(zero_spans_launch(zs, st));
  HIPCHECK(hipEventRecord(ctx->ev0, st));
  HIPCHECK(cascade_launch(P, o->arith, ctx->tgt.potential, false, sh.grid, sh.block, sh.lds, st));
  HIPCHECK(hipEventRecord(ctx->ev1, st));
  HIPCHECK(flush_energy_launch(M.wave_energy, sh.waves, red, st));
  unsigned long long resh[kResWords];
  HIPCHECK(hipMemcpyAsync(resh, M.res, sizeof(resh), hipMemcpyDeviceToHost, st));
  long long* hdst[5] = {out->range_hist, out->lat_y_hist, out->lat_z_hist, out->vac_hist, out->repl_hist};
  for (int i = 0; i < 5; ++i)
    if (hdst[i])
      HIPCHECK(hipMemcpyAsync(hdst[i], M.hist + hoff[i], size_t(nb[i]) * sizeof(long long), hipMemcpyDeviceToHost,
                              st));
  HIPCHECK(hipStreamSynchronize(st));
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  read_results(resh, E0, ctx->clock_khz, out);
  out->kernel_ms = ms;
  out->grid = sh.grid;
  out->block = sh.block;
  out->waves = sh.waves;
  return device_errors(resh + kResCtrl);
}

int mt_stopping_batch_arith(mt_ctx* ctx, int arith, int Z1, double m1, int mat, const double* E, long long n,
                            double* se_out) {
  if (!ctx || (n > 0 && (!E || !se_out))) return fail(MT_E_ARG, "NULL argument");
  if (arith < MT_FP64 || arith > MT_FP32_SURROGATE) return fail(MT_E_ARG, "unknown arithmetic");
  if (!ctx->has_target) return fail(MT_E_STATE, "no target set");
  if (mat < 0 || mat >= int(ctx->tgt.mat.size())) return fail(MT_E_ARG, "material index out of range");
  if (Z1 < 1 || Z1 > 92) return fail(MT_E_ARG, "Z1 out of range 1..92");
  if (n <= 0) return MT_OK;
  HIPCHECK(hipSetDevice(ctx->device));
  std::vector<Species> species;
  std::vector<int> ps;
  std::string msg = collect_species(ctx->tgt, &Z1, &m1, 1, species, ps);
  if (!msg.empty()) return fail(MT_E_LIMIT, msg);
  std::vector<unsigned char> blob;
  msg = build_tables(ctx->coef, ctx->tgt, species, blob);
  if (!msg.empty()) return fail(MT_E_LIMIT, msg);
  hipStream_t st = ctx->stream;
  HIPCHECK(ctx->tables.ensure(blob.size()));
  HIPCHECK(ctx->scratch_a.ensure(n * sizeof(double)));
  HIPCHECK(ctx->scratch_b.ensure(n * sizeof(double)));
  HIPCHECK(hipMemcpyAsync(ctx->tables.p, blob.data(), blob.size(), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(ctx->scratch_a.p, E, n * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHECK(stopping_launch(arith, ctx->tables.as<unsigned char>(), ps[0], mat, ctx->scratch_a.as<double>(), n,
                           ctx->scratch_b.as<double>(), st));
  HIPCHECK(hipMemcpyAsync(se_out, ctx->scratch_b.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return MT_OK;
}

int mt_stopping_batch(mt_ctx* ctx, int Z1, double m1, int mat, const double* E, long long n, double* se_out) {
  return mt_stopping_batch_arith(ctx, MT_FP64, Z1, m1, mat, E, n, se_out);
}

int mt_scatter_batch_arith(mt_ctx* ctx, int arith, int potential, const double* eps, const double* b, long long n,
                           double* s2, double* ct, double* st_out, int* iters) {
  if (!ctx || (n > 0 && (!eps || !b || !s2 || !ct || !st_out || !iters))) return fail(MT_E_ARG, "NULL argument");
  if (arith < MT_FP64 || arith > MT_FP32_SURROGATE) return fail(MT_E_ARG, "unknown arithmetic");
  if (potential < 0 || potential > 2) return fail(MT_E_ARG, "unknown potential");
  if (arith == MT_FP32_SURROGATE && potential != MT_UNIVERSAL)
    return fail(MT_E_ARG, "the surrogate replaces the universal potential's scattering");
  if (arith == MT_FP32_SURROGATE && !ctx->mlp_ready) return fail(MT_E_STATE, "no surrogate installed");
  if (n <= 0) return MT_OK;
  HIPCHECK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  HIPCHECK(ctx->scratch_a.ensure(n * sizeof(double)));
  HIPCHECK(ctx->scratch_b.ensure(n * sizeof(double)));
  HIPCHECK(ctx->scratch_c.ensure(n * sizeof(double)));
  HIPCHECK(ctx->scratch_d.ensure(n * sizeof(double)));
  HIPCHECK(ctx->scratch_e.ensure(n * 2 * sizeof(double)));
  HIPCHECK(hipMemcpyAsync(ctx->scratch_a.p, eps, n * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(ctx->scratch_b.p, b, n * sizeof(double), hipMemcpyHostToDevice, st));
  double* d_s2 = ctx->scratch_c.as<double>();
  double* d_ct = ctx->scratch_d.as<double>();
  double* d_st = ctx->scratch_e.as<double>();
  int* d_it = reinterpret_cast<int*>(d_st + n);
  MlpDev S;
  std::memset(&S, 0, sizeof(S));
  if (arith == MT_FP32_SURROGATE) {
    S.p = ctx->mlp_p.as<float>();
    S.w2bf = ctx->mlp_w2bf.as<uint32_t>();
    S.n = ctx->mlp_norm;
    S.bf16 = ctx->mlp_bf16;
  }
  HIPCHECK(scatter_launch(arith, potential, S, ctx->scratch_a.as<double>(), ctx->scratch_b.as<double>(), n, d_s2,
                          d_ct, d_st, d_it, st));
  HIPCHECK(hipMemcpyAsync(s2, d_s2, n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(ct, d_ct, n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(st_out, d_st, n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(iters, d_it, n * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return MT_OK;
}

int mt_scatter_batch(mt_ctx* ctx, int potential, const double* eps, const double* b, long long n, double* s2,
                     double* ct, double* st_out, int* iters) {
  return mt_scatter_batch_arith(ctx, MT_FP64, potential, eps, b, n, s2, ct, st_out, iters);
}

}  // extern "C"
