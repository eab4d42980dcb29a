// Arguments of the single-flight kernel (mt_flight.hip) and its launcher, shared
// by the kernel and the C ABI that fills them (mt_trim_flight).
#pragma once
#include <hip/hip_runtime.h>

#include "mt_tables.h"
#include "mytrim_gpu.h"

namespace mt {

struct FlightParams {
  const unsigned char* tables;  // the ion's table blob (a multiple of 16 bytes)
  int table_bytes;
  Geometry geo;
  double tau, cw, invscale;
  const mt_ion* ion;  // the flight's ion
  const double* u;    // the caller's uniforms, n_u of them
  long long n_u;
  mt_ion* ion_out;    // the ion at the flight's end (written when it ends normally)
  mt_event* ev;       // one event per collision, max_ev of them
  int max_ev;
  long long* n_used;  // [0] draws used, [1] events written, [2] status
};

hipError_t flight_launch(const FlightParams& P, int potential, hipStream_t stream);

}  // namespace mt
