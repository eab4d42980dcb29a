// Host-side preparation of the device tables (see mt_host.h).  The
// arithmetic below is the reference's, in its evaluation order, so the
// constants are the doubles MaterialBase::average() would produce.
#include "mt_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mt {

static const double kDefaultRows[92][17] = {
#include "mt_scoef_default.inc"
};

const std::vector<CoefRow>& default_coefs() {
  static std::vector<CoefRow> rows = [] {
    std::vector<CoefRow> r(92);
    for (int i = 0; i < 92; ++i) {
      const double* v = kDefaultRows[i];
      CoefRow& c = r[i];
      c.mm1 = v[1];
      c.m1 = v[2];
      c.mnat = v[3];
      c.rho = v[4];
      c.atrho = v[5];
      c.vfermi = v[6];
      c.heat = v[7];
      for (int j = 0; j < 8; ++j) c.pcoef[j] = v[8 + j];
      c.lfctr = v[16];
    }
    return r;
  }();
  return rows;
}

void HostMaterial::prepare() {  // material.C:37-74
  double tt = 0.0;
  for (auto& e : el) {
    if (e.t < 0.0) e.t = 0.0;
    tt += e.t;
  }
  for (auto& e : el) e.t /= tt;
  am = 0.0;
  az = 0.0;
  for (auto& e : el) {
    am += e.m * e.t;
    az += double(e.Z) * e.t;
  }
  arho = rho * 0.6022 / am;
}

double host_rpstop(const std::vector<CoefRow>& coef, int z2p, double e) {  // material.C:134-158
  const double* c = coef[z2p - 1].pcoef;
  const double pe0 = 25.0;
  const double pe = std::max(pe0, e);
  const double sl = (c[0] * std::pow(pe, c[1])) + (c[2] * std::pow(pe, c[3]));
  const double sh = c[4] / std::pow(pe, c[5]) * std::log(c[6] / pe + c[7] * pe);
  double sp = sl * sh / (sl + sh);
  if (e <= pe0) sp *= std::pow(e / pe0, (z2p <= 6) ? 0.25 : 0.45);
  return sp;
}

static StopConst make_stop(const std::vector<CoefRow>& coef, int z1, double m_ion, int z2) {
  StopConst c;
  std::memset(&c, 0, sizeof(c));
  const double fz1 = double(z1), fz2 = double(z2);
  c.m1 = (m_ion == 0.0) ? coef[z1 - 1].mm1 : m_ion;  // material.C:256-259
  for (int j = 0; j < 8; ++j) c.pc[j] = coef[z2 - 1].pcoef[j];
  c.velpwr = (z2 <= 6) ? 0.25 : 0.45;
  c.vfermi = coef[z2 - 1].vfermi;
  c.lfctr = coef[z1 - 1].lfctr;
  c.fz1 = fz1;
  c.kind = (z1 == 1) ? 1 : (z1 == 2) ? 2 : 3;
  c.he_k = 0.007 + 0.00005 * fz2;  // material.C:283
  // heavy ions, material.C:213-279
  const double yrmin = 0.13, vrmin = 1.0, vfermi = c.vfermi;
  c.cbrt_fz1 = std::cbrt(fz1);
  c.cbrt2_fz1 = c.cbrt_fz1 * c.cbrt_fz1;
  c.yr_lo = vrmin / c.cbrt2_fz1;
  c.yr_branch = std::max(yrmin, vrmin / c.cbrt2_fz1);
  c.vk = 3.0 * vfermi / 4.0;
  c.bq = (std::min(0.43, std::max(0.32, 0.12 + 0.025 * fz1))) / c.cbrt_fz1;
  c.l0k = std::min(1.2, 0.6 + fz1 / 30.0);
  c.q1 = std::max(0.0, 0.9 - 0.025 * fz1);
  c.d1 = std::abs(std::max(0.0, 0.9 - 0.025 * fz1) - 0.2000001);
  c.q2 = std::max(0.0, 1.0 - 0.025 * std::min(16.0, fz1));
  c.d2 = 0.025 * std::min(16.0, fz1);
  c.zk = 1.0 / (2.0 * vfermi * vfermi);
  c.z3 = (1.0 / (fz1 * fz1)) * (0.18 + 0.0015 * fz2);
  const double vr2 = std::max(vrmin, yrmin * c.cbrt2_fz1);
  const double vmin = 0.5 * (vr2 + std::sqrt(std::max(0.0, vr2 * vr2 - 0.8 * vfermi * vfermi)));
  c.eee = 25.0 * vmin * vmin;
  c.sp_low = host_rpstop(coef, z2, c.eee);
  c.power = (z2 == 6 || ((z2 == 14 || z2 == 32) && z1 <= 19)) ? 0.375 : 0.5;
  c.inv_m1 = 1.0 / c.m1;
  c.inv_vfermi = 1.0 / c.vfermi;
  c.inv_cbrt_fz1 = 1.0 / c.cbrt_fz1;
  c.inv_cbrt2_fz1 = 1.0 / c.cbrt2_fz1;
  c.inv_d1 = c.d1 > 0.0 ? 1.0 / c.d1 : 0.0;
  c.inv_d2 = c.d2 > 0.0 ? 1.0 / c.d2 : 0.0;
  c.ln_eee = std::log(c.eee);
  c.ln_emul = std::log(0.001 * c.inv_m1);
  {  // material.C:231-257 at yr = yr_branch, where the low-velocity branch has it
    const double yr = c.yr_branch;
    const double yr03 = std::pow(yr, 0.3);
    const double a = -0.803 * yr03 + 1.3167 * yr03 * yr03 + 0.38157 * yr + 0.008983 * yr * yr;
    const double q = std::min(1.0, std::max(0.0, 1.0 - std::exp(-std::min(a, 50.0))));
    const double l0 = (0.8 - q * std::min(1.2, 0.6 + fz1 / 30.0)) / c.cbrt_fz1;
    double l1;
    if (q < 0.2)
      l1 = 0.0;
    else if (q < std::max(0.0, 0.9 - 0.025 * fz1))
      l1 = c.bq * (q - 0.2) / std::abs(std::max(0.0, 0.9 - 0.025 * fz1) - 0.2000001);
    else if (q < std::max(0.0, 1.0 - 0.025 * std::min(16.0, fz1)))
      l1 = c.bq;
    else
      l1 = c.bq * (1.0 - q) / (0.025 * std::min(16.0, fz1));
    const double l = std::max(l1, l0 * c.lfctr);
    const double s = 4.0 * l * vfermi / 1.919;
    c.zeta_lo = q + (1.0 / (2.0 * vfermi * vfermi)) * (1.0 - q) * std::log(1.0 + s * s);
  }
  return c;
}

std::string collect_species_into(const HostTarget& tgt, const int* pZ, const double* pm, long long n, Species* sp,
                                 int* ns) {
  int k = 0;
  auto add = [&](int Z, double m) {
    for (int i = 0; i < k; ++i)
      if (sp[i].Z == Z && sp[i].m == m) return true;
    if (k >= kMaxSpecies) return false;
    sp[k++] = Species{Z, m};
    return true;
  };
  for (long long i = 0; i < n; ++i) {
    if (pZ[i] < 1 || pZ[i] > 92) return "primary Z out of range 1..92";
    if (!add(pZ[i], pm[i])) return "too many ion species";
  }
  for (const auto& m : tgt.mat)
    for (const auto& e : m.el)
      if (!add(e.Z, e.m)) return "too many ion species";
  *ns = k;
  return "";
}

int species_index(std::vector<Species>& species, int Z, double m) {
  for (size_t i = 0; i < species.size(); ++i)
    if (species[i].Z == Z && species[i].m == m) return int(i);
  species.push_back(Species{Z, m});
  return int(species.size()) - 1;
}

std::string collect_species(const HostTarget& tgt, const int* pZ, const double* pm, long long n,
                            std::vector<Species>& species, std::vector<int>& pka_species) {
  species.clear();
  pka_species.resize(size_t(n));
  for (long long i = 0; i < n; ++i) {
    if (pZ[i] < 1 || pZ[i] > 92) return "primary Z out of range 1..92";
    pka_species[i] = species_index(species, pZ[i], pm[i]);
    if (int(species.size()) > kMaxSpecies) return "too many ion species";
  }
  for (const auto& m : tgt.mat)
    for (const auto& e : m.el) {
      species_index(species, e.Z, e.m);
      if (int(species.size()) > kMaxSpecies) return "too many ion species";
    }
  return "";
}

namespace {
size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

TableHeader table_header(const HostTarget& tgt, int ns, int ne) {
  TableHeader h;
  std::memset(&h, 0, sizeof(h));
  const int nm = int(tgt.mat.size());
  h.n_species = ns;
  h.n_mat = nm;
  h.n_el = ne;
  h.potential = tgt.potential;
  h.off_mat = int(align16(sizeof(TableHeader)));
  h.off_el = int(align16(h.off_mat + sizeof(MatConst) * ns * nm));
  h.off_stop = int(align16(h.off_el + sizeof(ElemConst) * ns * ne));
  h.bytes = int(align16(h.off_stop + sizeof(StopConst) * ns * ne));
  return h;
}
}  // namespace

std::string build_tables(const std::vector<CoefRow>& coef, const HostTarget& tgt,
                         const std::vector<Species>& species, std::vector<unsigned char>& blob) {
  int ne = 0;
  for (const auto& m : tgt.mat) ne += int(m.el.size());
  const int ns = int(species.size());
  if (ns > kMaxSpecies || int(tgt.mat.size()) > kMaxMaterials || ne > kMaxElements) return "table limits exceeded";
  blob.assign(size_t(table_header(tgt, ns, ne).bytes), 0);
  size_t bytes = 0;
  return build_tables_into(coef, tgt, species.data(), ns, blob.data(), blob.size(), &bytes);
}

std::string build_tables_into(const std::vector<CoefRow>& coef, const HostTarget& tgt, const Species* species,
                              int ns, unsigned char* dst, size_t cap, size_t* bytes) {
  const int nm = int(tgt.mat.size());
  int ne = 0;
  for (const auto& m : tgt.mat) ne += int(m.el.size());
  if (ns > kMaxSpecies || nm > kMaxMaterials || ne > kMaxElements) return "table limits exceeded";
  TableHeader h = table_header(tgt, ns, ne);
  if (size_t(h.bytes) > cap) return "table buffer too small";
  *bytes = size_t(h.bytes);
  for (int s = 0; s < ns; ++s) h.species[s] = species[s];
  std::memset(dst, 0, size_t(h.bytes));
  std::memcpy(dst, &h, sizeof(h));
  MatConst* M = reinterpret_cast<MatConst*>(dst + h.off_mat);
  ElemConst* E = reinterpret_cast<ElemConst*>(dst + h.off_el);
  StopConst* S = reinterpret_cast<StopConst*>(dst + h.off_stop);

  for (int s = 0; s < ns; ++s) {
    const double m = species[s].m;
    const double fZ = double(species[s].Z);
    const double fZ023 = std::pow(fZ, 0.23);
    int eb = 0;
    for (int k = 0; k < nm; ++k) {  // MaterialBase::average (material.C:78-110)
      const HostMaterial& mat = tgt.mat[k];
      const double mu = m / mat.am;
      const double a = .5292 * .8853 / (fZ023 + std::pow(mat.az, 0.23));
      const double f = a * mat.am / (mat.az * fZ * 14.4 * (m + mat.am));
      const double epsdg = tgt.tmin * f * ((1.0 + mu) * (1.0 + mu)) / (4.0 * mu);
      MatConst& mc = M[s * nm + k];
      mc.f = f;
      mc.epsdg = epsdg;
      mc.a = a;
      mc.ln_f_epsdg = std::log(f * epsdg);
      mc.arho = mat.arho;
      mc.az = mat.az;
      mc.am = mat.am;
      mc.el_begin = eb;
      mc.n_el = int(mat.el.size());
      for (size_t i = 0; i < mat.el.size(); ++i) {
        const HostElement& he = mat.el[i];
        ElemConst& ec = E[s * ne + eb + int(i)];
        ec.t = he.t;
        ec.my = m / he.m;
        ec.ec = 4.0 * ec.my / ((1.0 + ec.my) * (1.0 + ec.my));
        ec.ai = .5292 * .8853 / (fZ023 + std::pow(double(he.Z), 0.23));
        ec.fi = ec.ai * he.m / (fZ * double(he.Z) * 14.4 * (m + he.m));
        ec.ln_fi = std::log(ec.fi);
        ec.Edisp = he.Edisp;
        ec.Elbind = he.Elbind;
        ec.Z = he.Z;
        int rs = -1;
        for (int q = 0; q < ns; ++q)
          if (species[q].Z == he.Z && species[q].m == he.m) rs = q;
        if (rs < 0) return "recoil species missing";
        ec.recoil_species = rs;
        S[s * ne + eb + int(i)] = make_stop(coef, species[s].Z, m, he.Z);
      }
      eb += int(mat.el.size());
    }
  }
  return "";
}

}  // namespace mt
