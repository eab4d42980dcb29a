// Per-function batch kernels: electronic stopping (MaterialBase::getrstop,
// material.C:114-122) over an energy array, and the scattering block of
// trim.C:172-272 over (eps, b) pairs.  One lane per element, grid-stride.
// ARITH selects the device functions of a cascade precision mode: 0 the FP64
// twins of the reference, 1 the single-precision ones, 2 the single-precision
// ones with the learned scattering map in front (the stopping of modes 1 and
// 2 is the same).  Inputs and outputs are doubles in every mode.
#include <hip/hip_runtime.h>

#include "mt_mlp.h"
#include "mt_physics.h"
#include "mt_scatter32.h"
#include "mt_tables.h"

namespace mt {

template <int ARITH>
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
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    if constexpr (ARITH == 0) {
      out[i] = getrstop<double>(T, s, k, E[i]);
    } else {
      const float Ef = (float)E[i];
      out[i] = (double)getrstop<float>(T, s, k, Ef, m_log(Ef));
    }
  }
}

template <int ARITH, int POT>
__global__ void __launch_bounds__(256) scatter_kernel(MlpDev mlp, const double* eps, const double* b, long long n,
                                                      double* s2, double* ct, double* st, int* iters) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    int it;
    if constexpr (ARITH == 0) {
      double a, c, d;
      scatter<double, POT>(eps[i], b[i], a, c, d, it);
      s2[i] = a;
      ct[i] = c;
      st[i] = d;
    } else {
      const float ef = (float)eps[i], bf = (float)b[i];
      float a, c, d;
      scatter_single<ARITH, POT>(mlp, ef, bf, m_log(ef), a, c, d, it);
      s2[i] = a;
      ct[i] = c;
      st[i] = d;
    }
    iters[i] = it;
  }
}

static int grid_for(long long n) {
  const long long g = (n + 255) / 256;
  return (int)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

hipError_t stopping_launch(int arith, const unsigned char* tables, int species, int mat, const double* E, long long n,
                           double* out, hipStream_t stream) {
  if (arith == 0)
    hipLaunchKernelGGL(stopping_kernel<0>, dim3(grid_for(n)), dim3(256), 0, stream, tables, species, mat, E, n, out);
  else
    hipLaunchKernelGGL(stopping_kernel<1>, dim3(grid_for(n)), dim3(256), 0, stream, tables, species, mat, E, n, out);
  return hipGetLastError();
}

template <int ARITH>
static void scatter_launch_a(int potential, const MlpDev& mlp, const double* eps, const double* b, long long n,
                             double* s2, double* ct, double* st, int* iters, hipStream_t stream) {
  if (potential == 1)
    hipLaunchKernelGGL((scatter_kernel<ARITH, 1>), dim3(grid_for(n)), dim3(256), 0, stream, mlp, eps, b, n, s2, ct,
                       st, iters);
  else if (potential == 2)
    hipLaunchKernelGGL((scatter_kernel<ARITH, 2>), dim3(grid_for(n)), dim3(256), 0, stream, mlp, eps, b, n, s2, ct,
                       st, iters);
  else
    hipLaunchKernelGGL((scatter_kernel<ARITH, 0>), dim3(grid_for(n)), dim3(256), 0, stream, mlp, eps, b, n, s2, ct,
                       st, iters);
}

hipError_t scatter_launch(int arith, int potential, const MlpDev& mlp, const double* eps, const double* b, long long n,
                          double* s2, double* ct, double* st, int* iters, hipStream_t stream) {
  if (arith == 0)
    scatter_launch_a<0>(potential, mlp, eps, b, n, s2, ct, st, iters, stream);
  else if (arith == 1)
    scatter_launch_a<1>(potential, mlp, eps, b, n, s2, ct, st, iters, stream);
  else
    scatter_launch_a<2>(0, mlp, eps, b, n, s2, ct, st, iters, stream);
  return hipGetLastError();
}

// Inference ablation of the learned scattering: the FP32 MAGIC solve
// (universal potential) it replaces (MODE 0), and the installed network with
// layer 2 in FP32 FMAs (1), BF16 dot products (2) or on the matrix cores (3),
// over the same (eps, b) pairs, without the domain check.  The loop runs
// whole waves (base is a multiple of 64, lanes past n evaluate a dummy pair and
// store nothing), as mlp_eval_mfma requires.
template <int MODE>
__global__ void __launch_bounds__(256) ablate_kernel(MlpDev S, const double* eps, const double* b, long long n,
                                                     double* s2) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long base = blockIdx.x * (long long)blockDim.x + (threadIdx.x & ~63); base < n; base += stride) {
    const long long i = base + (threadIdx.x & 63);
    const bool in = i < n;
    const float e = in ? (float)eps[i] : 1.0f, bb = in ? (float)b[i] : 1.0f;
    float r;
    if constexpr (MODE == 0) {
      float ct, st;
      int it;
      scatter_f32<0>(e, bb, m_log(e), r, ct, st, it);
    } else {
      float x0, x1, x2;
      mlp_inputs(S.n, e, bb, x0, x1, x2);
      MlpDev T = S;
      float y;
      if constexpr (MODE == 3) {
        y = mlp_eval_mfma(T, x0, x1, x2);
      } else {
        T.bf16 = MODE == 2;
        y = mlp_eval(T, x0, x1, x2);
      }
      r = fminf(__expf(y * S.n.y_scale + S.n.y_off), 1.0f);
    }
    if (in) s2[i] = r;
  }
}

hipError_t ablate_launch(int mode, const MlpDev& S, const double* eps, const double* b, long long n, double* s2,
                         hipStream_t stream) {
  const dim3 g(grid_for(n)), blk(256);
  if (mode == 0)
    hipLaunchKernelGGL(ablate_kernel<0>, g, blk, 0, stream, S, eps, b, n, s2);
  else if (mode == 1)
    hipLaunchKernelGGL(ablate_kernel<1>, g, blk, 0, stream, S, eps, b, n, s2);
  else if (mode == 2)
    hipLaunchKernelGGL(ablate_kernel<2>, g, blk, 0, stream, S, eps, b, n, s2);
  else
    hipLaunchKernelGGL(ablate_kernel<3>, g, blk, 0, stream, S, eps, b, n, s2);
  return hipGetLastError();
}

}  // namespace mt
