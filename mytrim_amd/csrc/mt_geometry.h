// Position -> material lookups of the three sample kinds, on the device.
//   SampleSolid::lookupMaterial     sample_solid.C:25-29
//   SampleLayers::lookupLayer       sample_layers.C:26-43 (x < 0 -> first
//                                   layer, beyond the last -> last material)
//   sampleClusters::lookupCluster   sample_clusters.C:43-133 (cell hash with a
//                                   ks-cell neighbourhood and minimum-image
//                                   distances along periodic axes)
//   SampleWire::lookupMaterial      sample_wire.C:37-46 (inside the cylinder
//                                   inscribed in the w[0] x w[1] box)
//   SampleBurriedWire::lookupMaterial  sample_burried_wire.C:34-55
// GEO is the kernel's sample template (0 solid and the two wire kinds, told
// apart by Geometry::kind at run time; 1 layers; 2 clusters).  Returns the
// material index, or -1 for vacuum (a CUT axis left by a clusters sample,
// outside a wire); `tag` receives the cluster index when inside one.
#pragma once
#include <hip/hip_runtime.h>

#include "mt_tables.h"

namespace mt {

// The occupancy cell of a flight's last cluster lookup and what it holds:
// o < 0 a clear cell, e < 0 a cell whose clusters overlap (the neighbourhood
// scan decides), else its candidate list occ_idx[o, e).  Consecutive
// collisions of a flight are a few A apart, mostly inside one cell, so the
// bitmap word, rank and list offsets are loaded once per cell, not per
// collision.  kNoCell marks an empty cache (a flight's start).
constexpr uint32_t kNoCell = 0xffffffffu;
struct CellCache {
  uint32_t cell = kNoCell;
  int o = -1, e = -1;
};

__device__ __forceinline__ int lookup_cluster(const Geometry& g, const double pos[3], CellCache& cc) {
  int k[3], k1[3], k2[3];
  int oc[3];
  bool out[3];
  for (int i = 0; i < 3; ++i) {
    double pw = pos[i];
    out[i] = pos[i] < 0.0 || pos[i] >= g.w[i];
    if (out[i]) {
      if (g.bc[i] == 2) return -2;
      if (g.bc[i] == 1) return -1;
      pw -= floor(pw / g.w[i]) * g.w[i];  // periodic image inside the sample
    }
    // occupancy cell by a multiplication with on/w: a position within
    // rounding of a cell face is within the 1e-9 w padding of any sphere that
    // contains it, so both cells at that face carry the sphere's bit
    oc[i] = min(max((int)floor(pw * g.on_w[i]), 0), g.on[i] - 1);
  }
  // clear occupancy cell: no cluster contains the position (the scan below
  // would find none); this skips the (2 ks + 1)^3-cell scan for most positions
  {
    // the cell count is at most 2^30, so a cell index fits 32 bits
    const uint32_t bit = (uint32_t)oc[0] + (uint32_t)g.on[0] * ((uint32_t)oc[1] + (uint32_t)g.on[1] * (uint32_t)oc[2]);
    if (bit != cc.cell) {
      cc.cell = bit;
      const uint32_t word = g.occ[bit >> 5];
      const uint32_t m = 1u << (bit & 31u);
      if ((word & m) == 0u) {
        cc.o = -1;
      } else {
        // occupied cell: its listed clusters are every cluster that can
        // contain a position in it.  When they do not overlap at most one
        // contains the position, and the reference's scan returns that one
        // (or none): test them with the scan's own distance.  A list of
        // overlapping clusters is the single entry -1: scan as the reference
        // does.
        const uint32_t k = g.occ_rank[bit >> 5] + (uint32_t)__popc(word & (m - 1u));
        cc.o = g.occ_off[k];
        cc.e = g.occ_idx[cc.o] >= 0 ? g.occ_off[k + 1] : -1;
      }
    }
    if (cc.o < 0) return -1;
    if (cc.e >= 0) {
      for (int q = cc.o; q < cc.e; ++q) {
        const int l = g.occ_idx[q];
        double d0 = pos[0] - g.cx[l];
        double d1 = pos[1] - g.cy[l];
        double d2 = pos[2] - g.cz[l];
        if (g.bc[0] == 0) d0 -= round(d0 * g.inv_w[0]) * g.w[0];
        if (g.bc[1] == 0) d1 -= round(d1 * g.inv_w[1]) * g.w[1];
        if (g.bc[2] == 0) d2 -= round(d2 * g.inv_w[2]) * g.w[2];
        const double rc = g.cr[l];
        if (d0 * d0 + d1 * d1 + d2 * d2 < rc * rc) return l;
      }
      return -1;
    }
  }
  // hash cell as the reference computes it (sample_clusters.C:59-81)
  for (int i = 0; i < 3; ++i) {
    k[i] = (int)floor((pos[i] * g.kn[i]) / g.w[i]);
    if (out[i]) {
      k[i] = k[i] % g.kn[i];
      if (k[i] < 0) k[i] += g.kn[i];
    }
  }
  int m1[3];  // k1 mod kn: the first cell index of the scan on each axis
  for (int i = 0; i < 3; ++i) {
    k1[i] = k[i] - g.ks[i];
    k2[i] = k[i] + g.ks[i];
    if (k1[i] < 0 && g.bc[i] != 0) k1[i] = 0;
    if (k2[i] >= g.kn[i] && g.bc[i] != 0) k2[i] = g.kn[i] - 1;
    m1[i] = k1[i] % g.kn[i];
    if (m1[i] < 0) m1[i] += g.kn[i];
  }
  // cells in the reference's order; the periodic cell index advances with
  // the loop counter and wraps at kn (one modulo per axis and lookup)
  int c0 = m1[0];
  for (int j0 = k1[0]; j0 <= k2[0]; ++j0) {
    int c1 = m1[1];
    for (int j1 = k1[1]; j1 <= k2[1]; ++j1) {
      int c2 = m1[2];
      for (int j2 = k1[2]; j2 <= k2[2]; ++j2) {
        int l = g.sh[c0 + g.kn[0] * (c1 + g.kn[1] * c2)];
        while (l >= 0) {
          double d0 = pos[0] - g.cx[l];
          double d1 = pos[1] - g.cy[l];
          double d2 = pos[2] - g.cz[l];
          // minimum image with 1/w: it can differ from d / w only where
          // |d| is half the box, far outside any cluster smaller than that
          if (g.bc[0] == 0) d0 -= round(d0 * g.inv_w[0]) * g.w[0];
          if (g.bc[1] == 0) d1 -= round(d1 * g.inv_w[1]) * g.w[1];
          if (g.bc[2] == 0) d2 -= round(d2 * g.inv_w[2]) * g.w[2];
          const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
          const double rc = g.cr[l];
          if (r2 < rc * rc) return l;
          l = g.cl[l];
        }
        c2 = (c2 + 1 == g.kn[2]) ? 0 : c2 + 1;
      }
      c1 = (c1 + 1 == g.kn[1]) ? 0 : c1 + 1;
    }
    c0 = (c0 + 1 == g.kn[0]) ? 0 : c0 + 1;
  }
  return -1;
}

// inside the wire: ((x / w0) 2 - 1)^2 + ((y / w1) 2 - 1)^2 <= 1
__device__ __forceinline__ bool in_wire(const Geometry& g, const double pos[3]) {
  const double x = (pos[0] / g.w[0]) * 2.0 - 1.0;
  const double y = (pos[1] / g.w[1]) * 2.0 - 1.0;
  return !((x * x + y * y) > 1.0);
}

__device__ __forceinline__ int lookup_wire(const Geometry& g, const double pos[3]) {
  if (g.kind == 3) return in_wire(g, pos) ? 0 : -1;
  if (pos[2] < 0.0 && pos[2] >= -250.0) return 1;  // cover layer
  if (pos[2] > g.w[2] || pos[2] < -250.0) return -1;
  return in_wire(g, pos) ? 0 : 1;
}

template <int GEO>
__device__ __forceinline__ int lookup_material(const Geometry& g, const double pos[3], int& tag, CellCache& cc) {
  tag = -1;
  if (GEO == 0) return g.kind >= 3 ? lookup_wire(g, pos) : 0;
  if (GEO == 1) {
    // first layer with x < its running end (the ends are non-decreasing, so
    // this counts a prefix); uniform loop bound keeps layer_end in scalars
    int i = 0;
    for (int j = 0; j < g.n_layers; ++j) i += !(pos[0] < g.layer_end[j]);
    return i >= g.n_mat ? g.n_mat - 1 : i;
  }
  const int l = lookup_cluster(g, pos, cc);
  if (l == -2) return -1;  // vacuum beyond a CUT face
  if (l == -1) return 0;
  tag = l;
  return 1;
}

// a single lookup (no flight to keep the cell for)
template <int GEO>
__device__ __forceinline__ int lookup_material(const Geometry& g, const double pos[3], int& tag) {
  CellCache cc;
  return lookup_material<GEO>(g, pos, tag, cc);
}

}  // namespace mt
