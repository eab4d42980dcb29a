// Training and evaluation kernels of the scattering surrogate (mt_mlp.h).
//
//   mlp_data      training pairs sampled as the cascade samples (eps, b),
//                 or on a rectangle in (log eps, b); target ln s2 from
//                 scatter_stable<double> (universal potential), one Philox
//                 stream per sample
//   mlp_fwdbwd    one lane per sample: forward pass and the per-sample
//                 back-propagated deltas (stored for the gradient kernel)
//   mlp_grad_*    per parameter: batch sum of delta x activation in fixed
//                 chunks, then the chunks in order (deterministic, no
//                 atomics); one extra column sums the loss
//   mlp_adam      Adam update with a cosine step-size schedule on device
// An epoch (every batch: fwdbwd, grad, adam) is captured once in a hipGraph
// and replayed.
#include <hip/hip_runtime.h>

#include <vector>

#include "mt_mlp.h"
#include "mt_philox.h"
#include "mt_physics.h"

namespace mt {

// per-sample record written by mlp_fwdbwd: x0 x1 x2 dy h1[64] dh1[64] h2[64] dh2[64] e2
constexpr int kRec = kMlpRecord;
constexpr int kRecH1 = 4, kRecD1 = 68, kRecH2 = 132, kRecD2 = 196, kRecDy = 3;

// maximum impact parameter of a triple at flight energy E (trim.C:88-95)
__device__ __forceinline__ double mlp_pmax(const MlpTriple& t, double E) {
  const double eeg = sqrt(E * t.f * t.epsdg);
  return t.a / (eeg + sqrt(eeg) + 0.125 * pow(eeg, 0.1));
}

// Training pairs, one Philox stream per sample; labels from the FP64 MAGIC
// solve.  Two samplings:
//   rect = 0: the way the cascade draws them: a random triple, a flight
//             energy (half log-uniform, half uniform on [E_lo, E_hi of the
//             triple]), eps = fi E and b = pmax(E) sqrt(u) / ai;
//   rect = 1: a rectangle: eps log-uniform on [eps_lo, eps_hi], b uniform on
//             [0, b_hi].
__global__ void __launch_bounds__(256) mlp_data(uint32_t seed, long long n, const MlpTriple* tri, int n_tri,
                                                double E_lo, int rect, double eps_lo, double eps_hi, double b_hi,
                                                MlpNorm nm, float* X, float* T) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const P4 r = philox10((uint32_t)i, (uint32_t)(i >> 32), 0x5a17u, 0u, seed, 0x6d6c70u);
    const P4 r2 = philox10((uint32_t)i, (uint32_t)(i >> 32), 0x5a18u, 0u, seed, 0x6d6c70u);
    const double u = u53(r.y, r.z);
    double eps, b;
    if (rect) {
      eps = eps_lo * pow(eps_hi / eps_lo, u);
      b = b_hi * u32d(r.w);
    } else {
      const int k = min((int)(u32d(r.x) * n_tri), n_tri - 1);
      const MlpTriple t = tri[k];
      const double E = (r2.x & 1u) ? E_lo * pow(t.E_hi / E_lo, u) : E_lo + (t.E_hi - E_lo) * u;
      eps = t.fi * E;
      b = mlp_pmax(t, E) * sqrt(u32d(r.w)) / t.ai;
    }
    double s2, ct, st;
    int it;
    scatter_stable<double, 0>(eps, b, s2, ct, st, it);
    float x0, x1, x2;
    mlp_inputs(nm, (float)eps, (float)b, x0, x1, x2);
    X[3 * i] = x0;
    X[3 * i + 1] = x1;
    X[3 * i + 2] = x2;
    T[i] = (float)((log(fmax(s2, 1e-300)) - nm.y_off) / nm.y_scale);
  }
}

__global__ void __launch_bounds__(256) mlp_fwdbwd(const float* __restrict__ p, const float* X, const float* T,
                                                  long long base, int B, float* rec) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const long long i = base + s;
  const float x0 = X[3 * i], x1 = X[3 * i + 1], x2 = X[3 * i + 2];
  float h1[64], h2[64];
#pragma unroll
  for (int j = 0; j < 64; ++j) h1[j] = mlp_h1(p, j, x0, x1, x2);
  float y = p[kOffB3];
  for (int k = 0; k < 64; ++k) {
    float a = p[kOffB2 + k];
#pragma unroll
    for (int j = 0; j < 64; ++j) a += p[kOffW2 + 64 * k + j] * h1[j];
    h2[k] = mlp_tanh(a);
    y += p[kOffW3 + k] * h2[k];
  }
  const float e = y - T[i];
  const float dy = 2.0f * e / (float)B;  // d(mean squared error)/dy
  float* r = rec + (size_t)s * kRec;
  r[0] = x0;
  r[1] = x1;
  r[2] = x2;
  r[kRecDy] = dy;
  float dh1[64];
#pragma unroll
  for (int j = 0; j < 64; ++j) dh1[j] = 0.0f;
  for (int k = 0; k < 64; ++k) {
    const float d2 = dy * p[kOffW3 + k] * (1.0f - h2[k] * h2[k]);
    r[kRecH2 + k] = h2[k];
    r[kRecD2 + k] = d2;
#pragma unroll
    for (int j = 0; j < 64; ++j) dh1[j] += p[kOffW2 + 64 * k + j] * d2;
  }
#pragma unroll
  for (int j = 0; j < 64; ++j) {
    r[kRecH1 + j] = h1[j];
    r[kRecD1 + j] = dh1[j] * (1.0f - h1[j] * h1[j]);
  }
  r[kRec - 1] = e * e;
}

constexpr int kGradChunk = kMlpGradChunk;

// record offsets of the delta and the input a parameter's gradient pairs
// (input -1: a bias)
__device__ __forceinline__ void grad_operands(int q, int& a_off, int& b_off) {
  if (q < kOffB1) {
    a_off = kRecD1 + q / kMlpIn;
    b_off = q % kMlpIn;
  } else if (q < kOffW2) {
    a_off = kRecD1 + (q - kOffB1);
    b_off = -1;
  } else if (q < kOffB2) {
    const int k = (q - kOffW2) / 64, j = (q - kOffW2) % 64;
    a_off = kRecD2 + k;
    b_off = kRecH1 + j;
  } else if (q < kOffW3) {
    a_off = kRecD2 + (q - kOffB2);
    b_off = -1;
  } else if (q < kOffB3) {
    a_off = kRecDy;
    b_off = kRecH2 + (q - kOffW3);
  } else if (q == kOffB3) {
    a_off = kRecDy;
    b_off = -1;
  } else {  // q == kMlpParams: the squared errors
    a_off = kRec - 1;
    b_off = -1;
  }
}

// dL/dparam = sum over samples of (delta of the unit it feeds) x (its input):
// lane q of block column c sums samples [c * kGradChunk, (c + 1) * kGradChunk)
// into part[c][q]; the last "parameter" collects the squared errors.  Fixed
// summation order, no atomics: the result is deterministic.
__global__ void __launch_bounds__(256) mlp_grad_part(const float* rec, int B, float* part) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > kMlpParams) return;
  int a_off, b_off;
  grad_operands(q, a_off, b_off);
  const int s0 = blockIdx.y * kGradChunk, s1 = min(B, s0 + kGradChunk);
  float acc = 0.0f;
  for (int s = s0; s < s1; ++s) {
    const float* r = rec + (size_t)s * kRec;
    acc += r[a_off] * (b_off < 0 ? 1.0f : r[b_off]);
  }
  part[(size_t)blockIdx.y * (kMlpParams + 1) + q] = acc;
}

__global__ void __launch_bounds__(256) mlp_grad_sum(const float* part, int n_chunks, int B, float* g,
                                                    float* loss_slot) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > kMlpParams) return;
  float acc = 0.0f;
  for (int c = 0; c < n_chunks; ++c) acc += part[(size_t)c * (kMlpParams + 1) + q];
  if (q == kMlpParams)
    *loss_slot = acc / (float)B;
  else
    g[q] = acc;
}

// Adam (Kingma & Ba 2015) with beta1 0.9, beta2 0.999; the step size follows
// a cosine from lr0 down to lr0 / 100 over the run; t counts steps on the device
__global__ void __launch_bounds__(256) mlp_adam(float* p, const float* g, float* m, float* v, int* t, float lr0,
                                                int total_steps) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int step = *t + 1;
  if (q < kMlpParams) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float frac = fminf((float)step / (float)total_steps, 1.0f);
    const float lr = lr0 * (0.01f + 0.495f * (1.0f + cosf(3.14159265f * frac)));
    m[q] = b1 * m[q] + (1.0f - b1) * g[q];
    v[q] = b2 * v[q] + (1.0f - b2) * g[q] * g[q];
    const float mh = m[q] / (1.0f - powf(b1, (float)step));
    const float vh = v[q] / (1.0f - powf(b2, (float)step));
    p[q] -= lr * mh / (sqrtf(vh) + 1e-8f);
  }
}

__global__ void mlp_tick(int* t) { *t += 1; }

__global__ void __launch_bounds__(256) mlp_infer(MlpDev S, const double* eps, const double* b, long long n,
                                                 double* s2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = mlp_s2(S, (float)eps[i], (float)b[i]);
    s2[i] = v < 0.0f ? -1.0 : (double)v;
  }
}

__global__ void mlp_pack_bf16(const float* p, uint32_t* w2bf) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < 64 * 32) {
    const int k = q / 32, jp = q % 32;
    w2bf[q] = bf16_bits(p[kOffW2 + 64 * k + 2 * jp]) | (bf16_bits(p[kOffW2 + 64 * k + 2 * jp + 1]) << 16);
  }
}

// ------------------------------------------------------------------ host side
static int grid_of(long long n) { return (int)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536); }

hipError_t mlp_make_data(uint32_t seed, long long n, const MlpTriple* tri, int n_tri, double E_lo, int rect,
                         const MlpNorm& nm, float* X, float* T, hipStream_t st) {
  hipLaunchKernelGGL(mlp_data, dim3(grid_of(n)), dim3(256), 0, st, seed, n, tri, n_tri, E_lo, rect,
                     (double)nm.eps_lo, (double)nm.eps_hi, (double)nm.b_hi, nm, X, T);
  return hipGetLastError();
}

// part: (B / kGradChunk rounded up) x (kMlpParams + 1) floats
// one optimisation step on samples [base, base + B)
static void grad_launch(const float* rec, int B, float* part, float* g, float* loss_slot, hipStream_t st) {
  const int chunks = (B + kGradChunk - 1) / kGradChunk;
  hipLaunchKernelGGL(mlp_grad_part, dim3((kMlpParams + 1 + 255) / 256, chunks), dim3(256), 0, st, rec, B, part);
  hipLaunchKernelGGL(mlp_grad_sum, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, st, part, chunks, B, g,
                     loss_slot);
}

static void step_launch(float* p, const float* X, const float* T, long long base, int B, float* rec, float* part,
                        float* g, float* m, float* v, int* t, float lr0, int total, float* loss_slot, hipStream_t st) {
  hipLaunchKernelGGL(mlp_fwdbwd, dim3((B + 255) / 256), dim3(256), 0, st, p, X, T, base, B, rec);
  grad_launch(rec, B, part, g, loss_slot, st);
  hipLaunchKernelGGL(mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, p, g, m, v, t, lr0, total);
  hipLaunchKernelGGL(mlp_tick, dim3(1), dim3(1), 0, st, t);
}

// Train for `epochs` over n samples in batches of B; one epoch captured as a
// graph and replayed.  losses[e] = mean batch loss of epoch e.
hipError_t mlp_train(float* p, const float* X, const float* T, long long n, int B, int epochs, float lr0,
                     float* rec, float* part, float* g, float* m, float* v, int* t, float* batch_loss,
                     double* losses, hipStream_t st) {
  const int steps = (int)(n / B);
  if (steps < 1) return hipErrorInvalidValue;
  const int total = steps * epochs;
  hipGraph_t graph;
  hipGraphExec_t exec;
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  for (int s = 0; s < steps; ++s)
    step_launch(p, X, T, (long long)s * B, B, rec, part, g, m, v, t, lr0, total, batch_loss + s, st);
  e = hipStreamEndCapture(st, &graph);
  if (e != hipSuccess) return e;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return e;
  std::vector<float> bl(steps);
  for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
    e = hipGraphLaunch(exec, st);
    if (e == hipSuccess) e = hipMemcpyAsync(bl.data(), batch_loss, steps * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    double acc = 0.0;
    for (float x : bl) acc += x;
    losses[ep] = acc / steps;
  }
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  return e;
}

// gradient and loss of one batch (for checks against the CPU)
hipError_t mlp_batch_grad(const float* p, const float* X, const float* T, int B, float* rec, float* part,
                          float* g, float* loss, hipStream_t st) {
  hipLaunchKernelGGL(mlp_fwdbwd, dim3((B + 255) / 256), dim3(256), 0, st, p, X, T, 0ll, B, rec);
  grad_launch(rec, B, part, g, loss, st);
  return hipGetLastError();
}

hipError_t mlp_pack(const float* p, uint32_t* w2bf, hipStream_t st) {
  hipLaunchKernelGGL(mlp_pack_bf16, dim3(8), dim3(256), 0, st, p, w2bf);
  return hipGetLastError();
}

hipError_t mlp_infer_launch(const MlpDev& S, const double* eps, const double* b, long long n, double* s2,
                            hipStream_t st) {
  hipLaunchKernelGGL(mlp_infer, dim3(grid_of(n)), dim3(256), 0, st, S, eps, b, n, s2);
  return hipGetLastError();
}

}  // namespace mt
