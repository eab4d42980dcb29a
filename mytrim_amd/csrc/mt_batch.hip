// Per-function batch kernels: electronic stopping (MaterialBase::getrstop,
// material.C:114-122) over an energy array, and the scattering block of
// trim.C:172-272 over (eps, b) pairs.  One lane per element, grid-stride.
#include <hip/hip_runtime.h>

#include "mt_physics.h"
#include "mt_tables.h"

namespace mt {

__global__ void __launch_bounds__(256) stopping_kernel(const unsigned char* blob, int s, int k, const double* E,
                                                       long long n, double* out) {
  const TableHeader* H = (const TableHeader*)blob;
  TableView T;
  T.n_species = H->n_species;
  T.n_mat = H->n_mat;
  T.n_el = H->n_el;
  T.potential = H->potential;
  T.mat = (const MatConst*)(blob + H->off_mat);
  T.el = (const ElemConst*)(blob + H->off_el);
  T.stop = (const StopConst*)(blob + H->off_stop);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = getrstop<double>(T, s, k, E[i]);
}

template <int POT>
__global__ void __launch_bounds__(256) scatter_kernel(const double* eps, const double* b, long long n, double* s2,
                                                      double* ct, double* st, int* iters) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double a, c, d;
    int it;
    scatter<double, POT>(eps[i], b[i], a, c, d, it);
    s2[i] = a;
    ct[i] = c;
    st[i] = d;
    iters[i] = it;
  }
}

static int grid_for(long long n) {
  const long long g = (n + 255) / 256;
  return (int)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

hipError_t stopping_launch(const unsigned char* tables, int species, int mat, const double* E, long long n,
                           double* out, hipStream_t stream) {
  hipLaunchKernelGGL(stopping_kernel, dim3(grid_for(n)), dim3(256), 0, stream, tables, species, mat, E, n, out);
  return hipGetLastError();
}

hipError_t scatter_launch(int potential, const double* eps, const double* b, long long n, double* s2, double* ct,
                          double* st, int* iters, hipStream_t stream) {
  if (potential == 1)
    hipLaunchKernelGGL(scatter_kernel<1>, dim3(grid_for(n)), dim3(256), 0, stream, eps, b, n, s2, ct, st, iters);
  else if (potential == 2)
    hipLaunchKernelGGL(scatter_kernel<2>, dim3(grid_for(n)), dim3(256), 0, stream, eps, b, n, s2, ct, st, iters);
  else
    hipLaunchKernelGGL(scatter_kernel<0>, dim3(grid_for(n)), dim3(256), 0, stream, eps, b, n, s2, ct, st, iters);
  return hipGetLastError();
}

}  // namespace mt
