// Internal: device contexts of the C++ API (see device.h).
#include "device.h"
#include "api_error.h"

#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>

#include "mytrim_gpu.h"
#include "sample_clusters.h"
#include "sample_layers.h"

namespace MyTRIM_NS {
namespace device {

int ordinal() {
  const char* d = std::getenv("MYTRIM_DEVICE");
  return d ? std::atoi(d) : 0;
}

void check(int rc, const char* what) {
  if (rc != MT_OK) api_error(std::string(what) + ": " + mt_last_error());
}

mt_ctx* create_context(SimconfType* simconf) {
  mt_ctx* c = nullptr;
  check(mt_ctx_create(ordinal(), &c), "mt_ctx_create");
  const std::vector<double> rows = simconf->tableRows();
  check(mt_ctx_set_tables(c, rows.data(), 92, 16), "mt_ctx_set_tables");
  return c;
}

StoppingContext& stopping_context(SimconfType* simconf) {
  static std::mutex mu;
  static std::map<SimconfType*, std::unique_ptr<StoppingContext>> all;
  std::lock_guard<std::mutex> lock(mu);
  auto& p = all[simconf];
  if (!p) {
    p.reset(new StoppingContext);
    p->ctx = create_context(simconf);
  }
  return *p;
}

void StoppingContext::upload(const MaterialBase& m) {
  std::vector<double> s{m._rho, double(m._element.size())};
  for (const auto& e : m._element) {
    s.push_back(e._Z);
    s.push_back(e._m);
    s.push_back(e._t);
  }
  if (s == sig) return;
  const double rho = m._rho;
  const int off[2] = {0, int(m._element.size())};
  std::vector<int> Z;
  std::vector<double> mm, t, ed, eb;
  for (const auto& e : m._element) {
    Z.push_back(e._Z);
    mm.push_back(e._m);
    t.push_back(e._t);
    ed.push_back(e._Edisp);
    eb.push_back(e._Elbind);
  }
  mt_target tg = {};
  tg.n_mat = 1;
  tg.mat_rho = &rho;
  tg.mat_el_off = off;
  tg.el_Z = Z.data();
  tg.el_m = mm.data();
  tg.el_t = t.data();
  tg.el_Edisp = ed.data();
  tg.el_Elbind = eb.data();
  tg.geo = MT_GEO_SOLID;
  tg.w[0] = tg.w[1] = tg.w[2] = 1.0;
  tg.tmin = 0.2;
  tg.cw = 0.001;
  tg.scale = 1.0;
  check(mt_ctx_set_target(ctx, &tg), "mt_ctx_set_target");
  sig = s;
}

std::vector<double> target_signature(SampleBase* s, SimconfType* simconf, int potential) {
  std::vector<double> v{double(s->geometry()), s->w[0], s->w[1], s->w[2], double(s->bc[0]), double(s->bc[1]),
                        double(s->bc[2]), simconf->tmin, simconf->tau, simconf->cw, simconf->lengthScale(),
                        double(potential), double(s->material.size())};
  for (MaterialBase* m : s->material) {
    v.push_back(m->_rho);
    v.push_back(double(m->_element.size()));
    for (const auto& e : m->_element) {
      v.push_back(e._Z);
      v.push_back(e._m);
      v.push_back(e._t);
      v.push_back(e._Edisp);
      v.push_back(e._Elbind);
    }
  }
  if (SampleLayers* l = dynamic_cast<SampleLayers*>(s))
    for (Real t : l->layerThickness) v.push_back(t);
  if (sampleClusters* c = dynamic_cast<sampleClusters*>(s)) {
    v.push_back(c->cn);
    for (int i = 0; i < 3; ++i) v.push_back(c->kn[i]);
    for (int i = 0; i < 4; ++i) v.push_back(double(reinterpret_cast<uintptr_t>(c->c[i])));
    if (c->cn > 0) v.push_back(c->c[0][c->cn - 1]);
  }
  return v;
}

void upload_target(mt_ctx* ctx, SampleBase* s, SimconfType* simconf, int potential) {
  const int kind = s->geometry();
  if (kind < 0)
    api_error(
        "sample kind has no device geometry (use SampleSolid, SampleLayers, sampleClusters, SampleWire or "
        "SampleBurriedWire)");
  std::vector<double> rho;
  std::vector<int> off{0}, Z;
  std::vector<double> mm, t, ed, eb;
  for (MaterialBase* m : s->material) {
    rho.push_back(m->_rho);
    for (const auto& e : m->_element) {
      Z.push_back(e._Z);
      mm.push_back(e._m);
      t.push_back(e._t);
      ed.push_back(e._Edisp);
      eb.push_back(e._Elbind);
    }
    off.push_back(int(Z.size()));
  }
  mt_target tg = {};
  tg.n_mat = int(rho.size());
  tg.mat_rho = rho.data();
  tg.mat_el_off = off.data();
  tg.el_Z = Z.data();
  tg.el_m = mm.data();
  tg.el_t = t.data();
  tg.el_Edisp = ed.data();
  tg.el_Elbind = eb.data();
  tg.geo = kind;
  for (int i = 0; i < 3; ++i) {
    tg.w[i] = s->w[i];
    tg.bc[i] = int(s->bc[i]);
  }
  std::vector<double> layers;
  if (SampleLayers* l = dynamic_cast<SampleLayers*>(s)) {
    layers.assign(l->layerThickness.begin(), l->layerThickness.end());
    tg.n_layers = int(layers.size());
    tg.layer_thick = layers.data();
  }
  if (sampleClusters* c = dynamic_cast<sampleClusters*>(s)) {
    for (int i = 0; i < 3; ++i) tg.kn[i] = c->kn[i];
    tg.n_cl = c->cn;
    tg.cl_x = c->c[0];
    tg.cl_y = c->c[1];
    tg.cl_z = c->c[2];
    tg.cl_r = c->c[3];
  }
  tg.tmin = simconf->tmin;
  tg.tau = simconf->tau;
  tg.cw = simconf->cw;
  tg.scale = simconf->lengthScale();
  tg.potential = potential;
  check(mt_ctx_set_target(ctx, &tg), "mt_ctx_set_target");
}

}  // namespace device
}  // namespace MyTRIM_NS
