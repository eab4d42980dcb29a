// Device-side constant tables of the cascade engine.
//
// Everything MaterialBase::average() (material.C:78-110) derives for one ion
// species is computed once on the host, per (species, material) and per
// (species, element), and kept here; the reference recomputes it before
// every flight (apps/runmytrim.C:85).  The stopping entries also carry the
// sub-expressions of rstop/rpstop (material.C:134-282) that depend only on
// (ion Z, target Z2), evaluated on the host in the reference's order so the
// device arithmetic stays the same expression tree.
#pragma once
#include <stdint.h>

namespace mt {

constexpr int kMaxSpecies = 32;
constexpr int kMaxMaterials = 16;
constexpr int kMaxElements = 32;   // summed over all materials
constexpr int kMaxLayers = 16;

struct MatConst {   // per (species, material)
  double f, epsdg, a, arho;
  double ln_f_epsdg;  // log(f epsdg): log eeg = (log E + ln_f_epsdg) / 2
  double az, am;    // prepare() averages, for Kinchin-Pease
  int el_begin, n_el;
};

struct ElemConst {  // per (species, element)
  double t;         // normalised stoichiometric fraction
  double fi, ai, ec, my;
  double ln_fi;     // log fi: log eps = log E + ln_fi
  double Edisp, Elbind;
  int Z;
  int recoil_species;
};

struct StopConst {  // per (species, element): rstop(ion, Z2)
  double m1;                    // ion mass used by rstop (table mm1 if m == 0)
  double pc[8];                 // pcoef of Z2
  double velpwr;                // 0.25 (Z2 <= 6) else 0.45
  double vfermi, lfctr, fz1;
  double cbrt_fz1, cbrt2_fz1;
  double yr_lo;                 // vrmin / cbrt2_fz1
  double yr_branch;             // max(yrmin, vrmin / cbrt2_fz1)
  double vk;                    // 3 vfermi / 4
  double bq;                    // min(0.43, max(0.32, 0.12 + 0.025 Z1)) / cbrt(Z1)
  double l0k;                   // min(1.2, 0.6 + Z1/30)
  double q1, d1, q2, d2;        // branch limits of l1
  double zk;                    // 1 / (2 vfermi^2)
  double z3;                    // (1/Z1^2) (0.18 + 0.0015 Z2)
  double eee, sp_low, power;    // low-velocity branch
  double he_k;                  // 0.007 + 0.00005 Z2
  // reciprocals and logarithms of the constants above, so that the kernel
  // multiplies where the reference divides
  double inv_m1, inv_vfermi, inv_cbrt_fz1, inv_cbrt2_fz1, inv_d1, inv_d2;
  double ln_eee;                // log(eee)
  double ln_emul;               // log(0.001 / m1): log e = log E + ln_emul
  double zeta_lo;               // heavy ions, low-velocity branch (yr = yr_branch): the
                                // effective-charge factor before its z1^3 term
  int kind;                     // 1 proton, 2 helium, 3 heavy ion
  int pad;
};

struct Species {
  int Z;
  double m;
};

// Table blob: header followed by MatConst[ns*nm], ElemConst[ns*ne],
// StopConst[ns*ne] (indices [s * nm + k] and [s * ne + e]).  The cascade
// kernel copies the whole blob into LDS.
struct TableHeader {
  int n_species, n_mat, n_el, potential;
  int off_mat, off_el, off_stop, bytes;   // byte offsets from the blob start
  Species species[kMaxSpecies];
};

struct TableView {
  int n_species, n_mat, n_el, potential;
  const MatConst* mat;
  const ElemConst* el;
  const StopConst* stop;
};

struct Geometry {
  int kind;                     // 0 solid, 1 layers, 2 clusters, 3 wire, 4 buried wire
  int bc[3];
  double w[3];
  int n_layers, n_mat;
  double layer_end[kMaxLayers]; // running sums of the thicknesses
  // clusters (sample_clusters.C): cell heads, next links, centres and radii
  int kn[3];
  double kd[3];
  double cmr;
  int ks[3];        // cells scanned on each side: int(cmr / kd[i]) + 1
  double inv_w[3];  // 1 / w[i], for the minimum image
  const int* sh;
  const int* cl;
  const double* cx;
  const double* cy;
  const double* cz;
  const double* cr;
  // occupancy bits of a grid of on[0] x on[1] x on[2] cells over the sample:
  // a cell's bit is set when some cluster sphere (radius padded by
  // 1e-9 of the sample size) meets it, so a position in a clear cell is in
  // no cluster and needs no hash scan
  int on[3];
  double on_w[3];  // on[i] / w[i]
  const uint32_t* occ;
  // candidate lists of the occupied cells: the k-th set bit of occ (k =
  // occ_rank[word] + the set bits below it in its word) owns clusters
  // occ_idx[occ_off[k] .. occ_off[k + 1]), every cluster whose padded sphere
  // meets the cell.  When two of them overlap, the list is the single entry
  // -1 and the lookup scans the hash as the reference does.
  const uint32_t* occ_rank;
  const int* occ_off;
  const int* occ_idx;
};

}  // namespace mt
