// Learned surrogate of the CM scattering: an MLP 3 -> 64 -> 64 -> 1 (tanh)
// mapping (log10 eps, log10 b, b) to ln sin^2(theta/2), trained on the
// engine's own universal-potential scattering (scatter_stable<double>) over
// the (eps, b) pairs a cascade in the current target actually draws.
// Inference in the cascades runs per lane on the VALU: layer 2 either in FP32
// FMAs or in BF16 dot products (v_dot2_f32_bf16), weights read with
// wave-uniform (scalar) loads.  A matrix-core form of layer 2 (mlp_eval_mfma)
// needs the whole wave in step and serves the inference ablation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mt {

constexpr int kMlpIn = 3;
constexpr int kMlpH = 64;
constexpr int kOffW1 = 0;                     // W1[j][i], j < 64, i < 3
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;  // b1[j]
constexpr int kOffW2 = kOffB1 + kMlpH;        // W2[k][j] (row k = output neuron)
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;  // b2[k]
constexpr int kOffW3 = kOffB2 + kMlpH;        // W3[k]
constexpr int kOffB3 = kOffW3 + kMlpH;        // b3
constexpr int kMlpParams = kOffB3 + 1;        // 4481
// floats per training sample written by the forward/backward kernel
constexpr int kMlpRecord = 4 + 4 * kMlpH + 1;
// samples per partial sum of the gradient reduction
constexpr int kMlpGradChunk = 128;

// input / output scaling and the trained domain
struct MlpNorm {
  float e_mid, e_inv_half;    // x0 = (log10 eps - e_mid) * e_inv_half
  float lb_mid, lb_inv_half;  // x1 = (log10 max(b, b_floor) - lb_mid) * lb_inv_half
  float b_mid, b_inv_half;    // x2 = (b - b_mid) * b_inv_half
  float b_floor;
  float y_off, y_scale;       // ln s2 = y * y_scale + y_off
  float eps_lo, eps_hi, b_hi; // domain; outside it the MAGIC path runs
};
constexpr int kMlpNormFloats = 12;

// one (ion species, material, element) combination the cascade samples:
// eps = fi E, b = pmax(E) sqrt(u) / ai, pmax from (a, f, epsdg) (trim.C:88-95, 143-160)
struct MlpTriple {
  double a, f, epsdg, ai, fi;
  double E_hi;  // largest flight energy with fi E <= eps_hi
};

struct MlpDev {
  const float* p;        // kMlpParams FP32 parameters
  const uint32_t* w2bf;  // W2 as BF16 pairs along j: [k][j/2]
  MlpNorm n;
  int bf16;
};

// tanh(x) = (1 - e) / (1 + e) with e = exp(-2|x|) in (0, 1], the sign of x
// restored; 1 / (1 + e) from the hardware reciprocal (about 1 ulp), not an
// IEEE division (a Newton-refined sequence of a dozen instructions)
__device__ __forceinline__ float mlp_tanh(float x) {
  const float e = __expf(-2.0f * fabsf(x));
  const float r = (1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e);
  return copysignf(r, x);
}

__host__ __device__ __forceinline__ uint32_t bf16_bits(float f) {  // round to nearest even
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

typedef __bf16 mt_bf16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float mlp_h1(const float* __restrict__ p, int j, float x0, float x1, float x2) {
  return mlp_tanh(p[kOffW1 + 3 * j] * x0 + p[kOffW1 + 3 * j + 1] * x1 + p[kOffW1 + 3 * j + 2] * x2 + p[kOffB1 + j]);
}

// network output y (normalised ln s2) for normalised inputs
__device__ __forceinline__ float mlp_eval(const MlpDev& S, float x0, float x1, float x2) {
  const float* __restrict__ p = S.p;
  float y = p[kOffB3];
  if (S.bf16) {
    uint32_t hp[32];
#pragma unroll
    for (int jp = 0; jp < 32; ++jp)
      hp[jp] = bf16_bits(mlp_h1(p, 2 * jp, x0, x1, x2)) | (bf16_bits(mlp_h1(p, 2 * jp + 1, x0, x1, x2)) << 16);
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
    for (int j = 0; j < 64; ++j) h[j] = mlp_h1(p, j, x0, x1, x2);
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

typedef __bf16 mt_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mt_f32x16 __attribute__((ext_vector_type(16)));

// The same network with layer 2 on the matrix cores: the wave's 64 evaluations
// form one 64 x 64 by 64 x 64 product, H2^T = W2 H1^T, as 2 x 2 tiles of
// v_mfma_f32_32x32x16_bf16 over 4 K-steps (16 instructions).  Tile layouts of
// that instruction (lane l, element j): A[i][k] holds A[l % 32][8 (l / 32) + j],
// B[k][n] holds B[8 (l / 32) + j][l % 32], D[i][n] holds element r at
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].  Column n of an M-tile t is
// evaluation (lane) n + 32 t, so lane l computes the layer-1 activations of its
// own evaluation and of lane l ^ 32's for its half of K, and after layer 2 holds
// 16 of the 64 rows of those two columns; lane l ^ 32 holds the other 16 of
// each tile, and one exchange completes layer 3.  W2 is read as A in BF16
// (the w2bf pairs), the layer-1 activations are rounded to BF16, accumulation
// is FP32.  All 64 lanes of the wave must call it together (uniform control
// flow): every evaluation's product uses operands from two lanes.
__device__ __forceinline__ float mlp_eval_mfma(const MlpDev& S, float x0, float x1, float x2) {
  const int l = threadIdx.x & 63;
  const int half = l >> 5, col = l & 31;
  const float q0 = __shfl_xor(x0, 32), q1 = __shfl_xor(x1, 32), q2 = __shfl_xor(x2, 32);
  float y_own = 0.0f;
  // one M-tile at a time, and within it one K-step at a time: the step's 8
  // layer-1 activations feed both output tiles at once, so the live set is
  // the two accumulators (32 registers) and one step's operands
  const float* p = S.p;
  const uint32_t* w2 = S.w2bf;
#pragma unroll 1
  for (int t = 0; t < 2; ++t) {
    // the weights are reloaded (from cache) for each M-tile: opaque copies of
    // the pointers keep the compiler from holding all 224 per-lane weights of
    // both tiles in registers (300 VGPRs, one wave per SIMD)
    asm volatile("" : "+s"(p), "+s"(w2));
    // evaluation n + 32 t of column n = col: this lane's own when t == half
    const float a0 = t == half ? x0 : q0, a1 = t == half ? x1 : q1, a2 = t == half ? x2 : q2;
    mt_f32x16 acc0, acc1;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r >> 2) + 4 * half + (r & 3);
      acc0[r] = p[kOffB2 + row];
      acc1[r] = p[kOffB2 + 32 + row];
    }
    const uint4* w0 = (const uint4*)(w2 + col * 32);
    const uint4* w1 = (const uint4*)(w2 + (32 + col) * 32);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      uint32_t w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = 16 * c + 8 * half + 2 * j;
        w[j] = bf16_bits(mlp_h1(p, k, a0, a1, a2)) | (bf16_bits(mlp_h1(p, k + 1, a0, a1, a2)) << 16);
      }
      const mt_bf16x8 hb = __builtin_bit_cast(mt_bf16x8, w);
      // W2[32 nt + col][16 c + 8 half .. + 7]
      acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mt_bf16x8, w0[2 * c + half]), hb, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mt_bf16x8, w1[2 * c + half]), hb, acc1, 0, 0, 0);
    }
    float part = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r >> 2) + 4 * half + (r & 3);
      part += p[kOffW3 + row] * mlp_tanh(acc0[r]) + p[kOffW3 + 32 + row] * mlp_tanh(acc1[r]);
    }
    const float y = part + __shfl_xor(part, 32);
    if (t == half) y_own = y;
  }
  return S.p[kOffB3] + y_own;
}

__host__ __device__ __forceinline__ void mlp_inputs(const MlpNorm& n, float eps, float b, float& x0, float& x1,
                                                    float& x2) {
  x0 = (log10f(eps) - n.e_mid) * n.e_inv_half;
  x1 = (log10f(fmaxf(b, n.b_floor)) - n.lb_mid) * n.lb_inv_half;
  x2 = (b - n.b_mid) * n.b_inv_half;
}

// s2 = sin^2(theta/2) from the surrogate, or a negative value outside its domain
__device__ __forceinline__ float mlp_s2(const MlpDev& S, float eps, float b) {
  if (!(eps >= S.n.eps_lo && eps <= S.n.eps_hi && b >= 0.0f && b <= S.n.b_hi)) return -1.0f;
  float x0, x1, x2;
  mlp_inputs(S.n, eps, b, x0, x1, x2);
  const float y = mlp_eval(S, x0, x1, x2);
  return fminf(__expf(y * S.n.y_scale + S.n.y_off), 1.0f);
}

}  // namespace mt
