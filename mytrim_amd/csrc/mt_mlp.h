// Learned surrogate of the CM scattering: an MLP 2 -> 64 -> 64 -> 1 (tanh)
// mapping (log10 eps, b) to ln sin^2(theta/2), trained on the engine's own
// universal-potential scattering (scatter_stable<double>).  Inference runs per
// lane on the VALU: layer 2 either in FP32 FMAs or in BF16 dot products
// (v_dot2_f32_bf16, no MFMA), weights read with wave-uniform (scalar) loads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mt {

constexpr int kMlpH = 64;
constexpr int kOffW1 = 0;                     // W1[j][i], j < 64, i < 2
constexpr int kOffB1 = 128;                   // b1[j]
constexpr int kOffW2 = 192;                   // W2[k][j] (row k = output neuron)
constexpr int kOffB2 = kOffW2 + 64 * 64;      // b2[k]
constexpr int kOffW3 = kOffB2 + 64;           // W3[k]
constexpr int kOffB3 = kOffW3 + 64;           // b3
constexpr int kMlpParams = kOffB3 + 1;        // 4417

struct MlpNorm {
  float e_mid, e_inv_half;   // x0 = (log10 eps - e_mid) * e_inv_half
  float b_mid, b_inv_half;   // x1 = (b - b_mid) * b_inv_half
  float y_off, y_scale;      // ln s2 = y * y_scale + y_off
  float eps_lo, eps_hi, b_hi;  // domain; outside it the MAGIC path runs
};

struct MlpDev {
  const float* p;        // kMlpParams FP32 parameters
  const uint32_t* w2bf;  // W2 as BF16 pairs along j: [k][j/2]
  MlpNorm n;
  int bf16;
};

__device__ __forceinline__ float mlp_tanh(float x) {
  const float e = __expf(-2.0f * fabsf(x));
  const float r = (1.0f - e) / (1.0f + e);
  return copysignf(r, x);
}

__host__ __device__ __forceinline__ uint32_t bf16_bits(float f) {  // round to nearest even
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

typedef __bf16 mt_bf16x2 __attribute__((ext_vector_type(2)));

// network output y (normalised ln s2) for normalised inputs (x0, x1)
__device__ __forceinline__ float mlp_eval(const MlpDev& S, float x0, float x1) {
  const float* __restrict__ p = S.p;
  float y = p[kOffB3];
  if (S.bf16) {
    uint32_t hp[32];
#pragma unroll
    for (int jp = 0; jp < 32; ++jp) {
      const int j = 2 * jp;
      const float ha = mlp_tanh(p[kOffW1 + 2 * j] * x0 + p[kOffW1 + 2 * j + 1] * x1 + p[kOffB1 + j]);
      const float hb = mlp_tanh(p[kOffW1 + 2 * j + 2] * x0 + p[kOffW1 + 2 * j + 3] * x1 + p[kOffB1 + j + 1]);
      hp[jp] = bf16_bits(ha) | (bf16_bits(hb) << 16);
    }
    const uint32_t* __restrict__ w2 = S.w2bf;
#pragma unroll 4
    for (int k = 0; k < 64; ++k) {
      float acc = p[kOffB2 + k];
#pragma unroll
      for (int jp = 0; jp < 32; ++jp)
        acc = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(mt_bf16x2, w2[k * 32 + jp]),
                                              __builtin_bit_cast(mt_bf16x2, hp[jp]), acc, false);
      y += p[kOffW3 + k] * mlp_tanh(acc);
    }
  } else {
    float h[64];
#pragma unroll
    for (int j = 0; j < 64; ++j)
      h[j] = mlp_tanh(p[kOffW1 + 2 * j] * x0 + p[kOffW1 + 2 * j + 1] * x1 + p[kOffB1 + j]);
#pragma unroll 4
    for (int k = 0; k < 64; ++k) {
      float acc = p[kOffB2 + k];
#pragma unroll
      for (int j = 0; j < 64; ++j) acc += p[kOffW2 + 64 * k + j] * h[j];
      y += p[kOffW3 + k] * mlp_tanh(acc);
    }
  }
  return y;
}

// s2 = sin^2(theta/2) from the surrogate, or a negative value outside its domain
__device__ __forceinline__ float mlp_s2(const MlpDev& S, float eps, float b) {
  if (!(eps >= S.n.eps_lo && eps <= S.n.eps_hi && b <= S.n.b_hi)) return -1.0f;
  const float x0 = (__log10f(eps) - S.n.e_mid) * S.n.e_inv_half;
  const float x1 = (b - S.n.b_mid) * S.n.b_inv_half;
  const float y = mlp_eval(S, x0, x1);
  return fminf(__expf(y * S.n.y_scale + S.n.y_off), 1.0f);
}

}  // namespace mt
