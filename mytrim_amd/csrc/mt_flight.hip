// One TrimBase::trim() call (trim.C:35-425) on the device, driven by the
// caller's uniform stream in the reference's draw order: r1 at flight start,
// r2 and hh per collision, then the unit-ball rejection loop for the
// azimuth (trim.C:71, 143, 147, 322-333).  Every collision is logged as an
// mt_event so the host can replay the virtual hooks (trim.h:72-84) in order.
// A single lane runs the flight; this is the compatibility path for callers
// that drive trim() one ion at a time.
#include <hip/hip_runtime.h>

#include "mt_flight.h"
#include "mt_geometry.h"
#include "mt_physics.h"
#include "mt_tables.h"
#include "mytrim_gpu.h"

namespace mt {

enum { kFlightDone = 0, kFlightNeedDraws = 1, kFlightNeedEvents = 2, kFlightRunaway = 3, kFlightNonFinite = 4 };

template <int POT>
__global__ void __launch_bounds__(64) flight_kernel(FlightParams P) {
  // the wave copies the ion's tables to LDS; each collision then reads its
  // constants there (about a hundred cycles) instead of from L2
  extern __shared__ __align__(16) unsigned char lds_tables[];
  {
    const uint4* src = (const uint4*)P.tables;
    uint4* dst = (uint4*)lds_tables;
    for (int i = threadIdx.x; i < P.table_bytes / 16; i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const unsigned char* blob = lds_tables;
  const TableHeader* H = (const TableHeader*)blob;
  TableView T;
  T.n_species = H->n_species;
  T.n_mat = H->n_mat;
  T.n_el = H->n_el;
  T.potential = H->potential;
  T.mat = (const MatConst*)(blob + H->off_mat);
  T.el = (const ElemConst*)(blob + H->off_el);
  T.stop = (const StopConst*)(blob + H->off_stop);
  const Geometry& G = P.geo;

  mt_ion ion = *P.ion;
  const int sp = 0;  // the flight's ion is species 0 of its table
  long long iu = 0;
  int status = kFlightDone;
  auto draw = [&]() -> double {
    if (iu < P.n_u) return P.u[iu++];
    status = kFlightNeedDraws;
    return 0.5;
  };
  int nev = 0;
  ion.state = 0;
  const double r1 = draw();
  int ic = 0;
  do {
    if (status != kFlightDone) break;
    ++ic;
    int mtag;
    const double pos[3] = {ion.pos[0], ion.pos[1], ion.pos[2]};
    const int k = (G.kind == 0 || G.kind >= 3) ? lookup_material<0>(G, pos, mtag)
                  : G.kind == 1 ? lookup_material<1>(G, pos, mtag)
                                : lookup_material<2>(G, pos, mtag);
    if (k < 0) break;  // vacuum: state stays MOVING (trim.C:80-82)
    {
      const double inv = 1.0 / sqrt(ion.dir[0] * ion.dir[0] + ion.dir[1] * ion.dir[1] + ion.dir[2] * ion.dir[2]);
      for (int i = 0; i < 3; ++i) ion.dir[i] *= inv;
    }
    const MatConst& mc = T.mat[sp * T.n_mat + k];
    double eps = ion.E * mc.f;
    const double eeg = sqrt(eps * mc.epsdg);
    const double pmax = mc.a / (eeg + sqrt(eeg) + 0.125 * pow(eeg, 0.1));
    double ls = 1.0 / (M_PI * (pmax * pmax) * mc.arho);
    if (ic == 1) ls = r1 * mn(ls, P.cw);
    const double p = pmax * sqrt(draw());
    double hh = draw();
    const int eb = sp * T.n_el + mc.el_begin;
    int nn = 0;
    for (; nn < mc.n_el; ++nn) {
      hh -= T.el[eb + nn].t;
      if (hh <= 0.0) break;
    }
    if (nn >= mc.n_el) nn = mc.n_el - 1;
    const ElemConst& ec = T.el[eb + nn];
    eps = ec.fi * ion.E;
    const double b = p / ec.ai;
    double dee = ls * getrstop<double>(T, sp, k, ion.E);
    double s2, ct, st;
    int it;
    scatter<double, POT>(eps, b, s2, ct, st, it);
    double den = ec.ec * s2 * ion.E;
    if (dee > ion.E) dee = ion.E;
    ion.E -= dee;
    const double p1 = sqrt(2.0 * ion.m * ion.E);
    if (den > ion.E) den = ion.E;
    ion.E -= den;
    const double p2 = sqrt(2.0 * ion.m * ion.E);
    double rpos[3], rdir[3];
    for (int i = 0; i < 3; ++i) rpos[i] = ion.pos[i];
    for (int i = 0; i < 3; ++i) ion.pos[i] += ion.dir[i] * (ls - P.tau) * P.invscale;
    for (int i = 0; i < 3; ++i) rdir[i] = ion.dir[i] * p1;
    const double rE = den - ec.Elbind;
    // random unit vector perpendicular to dir (trim.C:322-333)
    double perp[3], norm;
    int guard = 0;
    do {
      double rd[3];
      do {
        for (int i = 0; i < 3; ++i) rd[i] = 2.0 * draw() - 1.0;
      } while (rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2] > 1.0 && status == kFlightDone);
      for (int i = 0; i < 3; ++i)
        perp[i] = ion.dir[(i + 1) % 3] * rd[(i + 2) % 3] - ion.dir[(i + 2) % 3] * rd[(i + 1) % 3];
      norm = sqrt(perp[0] * perp[0] + perp[1] * perp[1] + perp[2] * perp[2]);
    } while (norm == 0.0 && status == kFlightDone && ++guard < 1000);
    if (status != kFlightDone) break;
    for (int i = 0; i < 3; ++i) perp[i] /= norm;
    const double psi = atan2(st, ct + ec.my);
    const double cpsi = cos(psi), spsi = sin(psi);
    for (int i = 0; i < 3; ++i) ion.dir[i] *= cpsi;
    for (int i = 0; i < 3; ++i) ion.dir[i] += perp[i] * spsi;
    for (int i = 0; i < 3; ++i) rdir[i] -= ion.dir[i] * p2;
    int fate, vac = 0;
    bool lost = false;
    for (int i = 0; i < 3; ++i)
      if (G.bc[i] == 2 && (ion.pos[i] > G.w[i] || ion.pos[i] < 0.0)) lost = true;
    if (lost) {
      ion.state = 4;
      fate = 2;
    } else if (rE > ec.Edisp - ec.Elbind) {
      fate = 1;
      if (ion.E > ec.Edisp) {
        vac = 1;
      } else {
        ion.state = (ion.Z == ec.Z) ? 1 : 2;
      }
    } else {
      fate = 0;
      if (ion.E < ion.Ef) ion.state = 3;
    }
    if (nev >= P.max_ev) {
      status = kFlightNeedEvents;
      break;
    }
    mt_event& e = P.ev[nev++];
    e.pka_E = ion.E;
    for (int i = 0; i < 3; ++i) {
      e.pka_pos[i] = ion.pos[i];
      e.pka_dir[i] = ion.dir[i];
      e.rec_pos[i] = rpos[i];
      e.rec_dir[i] = rdir[i];
    }
    e.ls = ls;
    e.dee = dee;
    e.den = den;
    e.rec_E = rE;
    e.material = k;
    e.element = nn;
    e.pka_state = ion.state;
    e.fate = fate;
    e.vacancy = vac;
    e.mat_tag = mtag;
    if (ic >= 100000000) {
      status = kFlightRunaway;
      break;
    }
    // a NaN or infinite energy or position would fail every threshold and
    // run to the collision cap (x - x is 0 exactly for finite x)
    if (!(ion.E - ion.E == 0.0) || !(ion.pos[0] - ion.pos[0] == 0.0)) {
      status = kFlightNonFinite;
      break;
    }
  } while (ion.state == 0);
  if (status == kFlightDone) *P.ion_out = ion;
  P.n_used[0] = iu;
  P.n_used[1] = nev;
  P.n_used[2] = status;
}

hipError_t flight_launch(const FlightParams& P, int potential, hipStream_t stream) {
  const unsigned lds = unsigned(P.table_bytes);
  if (potential == 1)
    hipLaunchKernelGGL(flight_kernel<1>, dim3(1), dim3(64), lds, stream, P);
  else if (potential == 2)
    hipLaunchKernelGGL(flight_kernel<2>, dim3(1), dim3(64), lds, stream, P);
  else
    hipLaunchKernelGGL(flight_kernel<0>, dim3(1), dim3(64), lds, stream, P);
  return hipGetLastError();
}

}  // namespace mt
