// Per-collision physics of the cascade kernels, templated on the working
// precision R (double for the FP64 path, float for the FP32 path).
//
// Expression trees follow the reference line by line so that the FP64 path
// differs from it only by the device's libm and FMA contraction:
//   electronic stopping   rpstop material.C:134-158, rstop material.C:161-282,
//                         getrstop material.C:114-122
//   scattering            trim.C:172-272 (Rutherford branch, Newton solve for
//                         the apsis, MAGIC angle) for the three potentials
// Constants that depend only on (ion, target) come precomputed in StopConst.
#pragma once
#include <hip/hip_runtime.h>

#include "mt_fmath.h"
#include "mt_tables.h"

namespace mt {

constexpr int kMaxNewton = 200;  // the reference loop is unbounded; 2-6 in practice

template <typename R>
__device__ __forceinline__ R mx(R a, R b) { return (a < b) ? b : a; }  // std::max
template <typename R>
__device__ __forceinline__ R mn(R a, R b) { return (b < a) ? b : a; }  // std::min
template <typename R>
__device__ __forceinline__ R sq(R x) { return x * x; }

// precision-dispatched math: FP64 exp/log/pow from mt_fmath.h (a few ulp,
// a third to a quarter of the library's instructions), single precision
// the hardware-assisted ocml variants
__device__ __forceinline__ double m_exp(double x) { return fx_exp(x); }
__device__ __forceinline__ double m_log(double x) { return fx_log(x); }
__device__ __forceinline__ double m_pow(double x, double y) { return fx_pow(x, y); }
__device__ __forceinline__ double m_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double m_abs(double x) { return fabs(x); }
__device__ __forceinline__ float m_exp(float x) { return __expf(x); }
__device__ __forceinline__ float m_log(float x) { return __logf(x); }
__device__ __forceinline__ float m_pow(float x, float y) { return __powf(x, y); }
__device__ __forceinline__ float m_sqrt(float x) { return __fsqrt_rn(x); }
__device__ __forceinline__ float m_abs(float x) { return fabsf(x); }

#define K(x) R(x)

template <typename R>
__device__ __forceinline__ R rpstop(const StopConst& c, R e) {  // material.C:134-158
  const R pe0 = K(25.0);
  const R pe = mx(pe0, e);
  const R sl = (R(c.pc[0]) * m_pow(pe, R(c.pc[1]))) + (R(c.pc[2]) * m_pow(pe, R(c.pc[3])));
  const R sh = R(c.pc[4]) / m_pow(pe, R(c.pc[5])) * m_log(R(c.pc[6]) / pe + R(c.pc[7]) * pe);
  R sp = sl * sh / (sl + sh);
  if (e <= pe0) sp *= m_pow(e / pe0, R(c.velpwr));
  return sp;
}

template <typename R>
__device__ __forceinline__ R rstop(const StopConst& c, R E) {  // material.C:161-282
  const R e = (K(0.001) * E) / R(c.m1);
  R se;
  if (c.kind == 1) {
    se = rpstop(c, e);
  } else if (c.kind == 2) {
    const R he0 = K(1.0);
    const R he = mx(he0, e);
    const R b = m_log(he);
    const R b2 = b * b;
    const R b4 = b2 * b2;
    R a = K(0.2865) + K(0.1266) * b - K(0.001429) * b2 + K(0.02402) * b * b2 - K(0.01135) * b4 +
          K(0.001475) * b4 * b;
    R heh = K(1.0) - m_exp(-mn(K(30.0), a));
    a = K(1.0) + R(c.he_k) * m_exp(-sq(K(7.6) - b));
    heh *= a * a;
    se = rpstop(c, he) * heh * K(4.0);
    if (e <= he0) se *= m_sqrt(e / he0);
  } else {
    const R vf = R(c.vfermi);
    const R v = m_sqrt(e / K(25.0)) / vf;
    const R v2 = v * v;
    R vr;
    if (v >= K(1.0))
      vr = v * vf * (K(1.0) + K(1.0) / (K(5.0) * v2));
    else
      vr = R(c.vk) * (K(1.0) + (K(2.0) * v2 / K(3.0)) - v2 * v2 / K(15.0));
    R yr = mx(K(0.13), vr / R(c.cbrt2_fz1));
    yr = mx(yr, R(c.yr_lo));
    const R yr03 = m_pow(yr, K(0.3));
    R a = K(-0.803) * yr03 + K(1.3167) * yr03 * yr03 + K(0.38157) * yr + K(0.008983) * yr * yr;
    const R q = mn(K(1.0), mx(K(0.0), K(1.0) - m_exp(-mn(a, K(50.0)))));
    const R l0 = (K(0.8) - q * R(c.l0k)) / R(c.cbrt_fz1);
    R l1;
    if (q < K(0.2))
      l1 = K(0.0);
    else if (q < R(c.q1))
      l1 = R(c.bq) * (q - K(0.2)) / R(c.d1);
    else if (q < R(c.q2))
      l1 = R(c.bq);
    else
      l1 = R(c.bq) * (K(1.0) - q) / R(c.d2);
    const R l = mx(l1, l0 * R(c.lfctr));
    R zeta = q + R(c.zk) * (K(1.0) - q) * m_log(K(1.0) + sq(K(4.0) * l * vf / K(1.919)));
    a = -sq(K(7.6) - mx(K(0.0), m_log(e)));
    zeta *= K(1.0) + R(c.z3) * m_exp(a);
    if (yr <= R(c.yr_branch))
      se = R(c.sp_low) * sq(zeta * R(c.fz1)) * m_pow(e / R(c.eee), R(c.power));
    else
      se = rpstop(c, e) * sq(zeta * R(c.fz1));
  }
  return se * K(10.0);
}

// MaterialBase::getrstop for species s in material k [eV/A]
template <typename R>
__device__ __forceinline__ R getrstop(const TableView& T, int s, int k, R E) {
  const MatConst& mc = T.mat[s * T.n_mat + k];
  R se = K(0.0);
  for (int i = 0; i < mc.n_el; ++i) {
    const int e = s * T.n_el + mc.el_begin + i;
    se += rstop(T.stop[e], E) * R(T.el[e].t);
  }
  return se * R(mc.arho);
}

// trim.C:172-272
template <typename R, int POT>
__device__ __forceinline__ void scatter(R eps, R b, R& s2, R& ct, R& st, int& iters) {
  if (eps > K(10.0)) {
    s2 = K(1.0) / (K(1.0) + (K(1.0) + b * (K(1.0) + b)) * sq(K(2.0) * eps * b));
    const R c2 = K(1.0) - s2;
    ct = K(2.0) * c2 - K(1.0);
    st = m_sqrt(K(1.0) - ct * ct);
    iters = 0;
    return;
  }
  R r = b;
  R rr = K(-2.7) * m_log(eps * b);
  if (rr >= b) {
    rr = K(-2.7) * m_log(eps * rr);
    if (rr >= b) r = rr;
  }
  R v, v1, q;
  int it = 0;
  do {
    if (POT == 1) {  // Moliere
      const R ex1 = m_exp(K(-0.3) * r);
      const R ex2 = sq(ex1) * sq(ex1);
      const R ex3 = ex2 * (sq(ex2) * sq(ex2));
      v = (K(0.35) * ex1 + K(0.55) * ex2 + K(0.1) * ex3) / r;
      v1 = -(v + K(0.105) * ex1 + K(0.66) * ex2 + K(0.6) * ex3) / r;
    } else if (POT == 2) {  // C-Kr
      const R ex1 = m_exp(K(-0.279) * r);
      const R ex2 = m_exp(K(-0.637) * r);
      const R ex3 = m_exp(K(-1.1919) * r);
      v = (K(0.191) * ex1 + K(0.474) * ex2 + K(0.335) * ex3) / r;
      v1 = -(v + K(0.531865) * ex1 + K(0.30181) * ex2 + K(0.6437) * ex3) / r;
    } else {  // universal
      const R ex1 = K(0.18175) * m_exp(K(-3.1998) * r);
      const R ex2 = K(0.50986) * m_exp(K(-0.94229) * r);
      const R ex3 = K(0.28022) * m_exp(K(-0.4029) * r);
      const R ex4 = K(0.028171) * m_exp(K(-0.20162) * r);
      v = (ex1 + ex2 + ex3 + ex4) / r;
      v1 = -(v + K(3.1998) * ex1 + K(0.94229) * ex2 + K(0.4029) * ex3 + K(0.20162) * ex4) / r;
    }
    const R fr = b * b / r + v * r / eps - r;
    const R fr1 = -b * b / (r * r) + (v + v1 * r) / eps - K(1.0);
    q = fr / fr1;
    r -= q;
    ++it;
  } while (m_abs(q / r) > K(0.001) && it < kMaxNewton);
  const R roc = K(-2.0) * (eps - v) / v1;
  const R sqe = m_sqrt(eps);
  R cc, aa, ff;
  if (POT == 1) {
    cc = (K(0.009611) + sqe) / (K(0.005175) + sqe);
    aa = K(2.0) * eps * (K(1.0) + (K(0.6743) / sqe)) * m_pow(b, cc);
    ff = (m_sqrt(aa * aa + K(1.0)) - aa) * ((K(6.314) + eps) / (K(10.0) + eps));
  } else if (POT == 2) {
    cc = (K(0.235809) + sqe) / (K(0.126000) + sqe);
    aa = K(2.0) * eps * (K(1.0) + (K(1.0144) / sqe)) * m_pow(b, cc);
    ff = (m_sqrt(aa * aa + K(1.0)) - aa) * ((K(6935.0) + eps) / (K(83550.0) + eps));
  } else {
    cc = (K(0.011615) + sqe) / (K(0.0071222) + sqe);
    aa = K(2.0) * eps * (K(1.0) + (K(0.99229) / sqe)) * m_pow(b, cc);
    ff = (m_sqrt(aa * aa + K(1.0)) - aa) * ((K(9.3066) + eps) / (K(14.813) + eps));
  }
  const R delta = (r - b) * aa * ff / (ff + K(1.0));
  const R co = (b + delta + roc) / (r + roc);
  const R c2 = co * co;
  s2 = K(1.0) - c2;
  ct = K(2.0) * c2 - K(1.0);
  st = m_sqrt(K(1.0) - ct * ct);
  iters = it;
}

// Scattering without the small-angle cancellation (the FP32 path, and the
// surrogate's training targets in FP64).  Same Newton solve and
// MAGIC constants as scatter(), but the small-angle end is evaluated without
// cancellation: at the apsis b^2/r + v r/eps - r = 0, so
//   r - b = v r^2 / (eps (r + b)),
//   1 - co = (r - b)(1 - A) / (r + roc),   A = aa ff / (ff + 1),
//   s2 = (1 - co)(2 - (1 - co)),  ct = 1 - 2 s2,  st = 2 sqrt(s2 (1 - s2)),
// which is algebraically the reference's co = (b + delta + roc)/(r + roc),
// s2 = 1 - co^2 but keeps full relative precision when s2 << 1 (glancing
// collisions transfer s2 ~ tmin / (ec E), below single-precision epsilon).
template <typename R, int POT>
__device__ __forceinline__ void scatter_stable(R eps, R b, R& s2, R& ct, R& st, int& iters) {
  if (eps > K(10.0)) {
    s2 = K(1.0) / (K(1.0) + (K(1.0) + b * (K(1.0) + b)) * sq(K(2.0) * eps * b));
    ct = K(1.0) - K(2.0) * s2;
    st = K(2.0) * m_sqrt(s2 * (K(1.0) - s2));
    iters = 0;
    return;
  }
  R r = b;
  R rr = K(-2.7) * m_log(eps * b);
  if (rr >= b) {
    rr = K(-2.7) * m_log(eps * rr);
    if (rr >= b) r = rr;
  }
  R v, v1, q;
  int it = 0;
  auto pot = [](R x, R& vv, R& vv1) {
    if (POT == 1) {
      const R ex1 = m_exp(K(-0.3) * x);
      const R ex2 = sq(ex1) * sq(ex1);
      const R ex3 = ex2 * (sq(ex2) * sq(ex2));
      vv = (K(0.35) * ex1 + K(0.55) * ex2 + K(0.1) * ex3) / x;
      vv1 = -(vv + K(0.105) * ex1 + K(0.66) * ex2 + K(0.6) * ex3) / x;
    } else if (POT == 2) {
      const R ex1 = m_exp(K(-0.279) * x);
      const R ex2 = m_exp(K(-0.637) * x);
      const R ex3 = m_exp(K(-1.1919) * x);
      vv = (K(0.191) * ex1 + K(0.474) * ex2 + K(0.335) * ex3) / x;
      vv1 = -(vv + K(0.531865) * ex1 + K(0.30181) * ex2 + K(0.6437) * ex3) / x;
    } else {
      const R ex1 = K(0.18175) * m_exp(K(-3.1998) * x);
      const R ex2 = K(0.50986) * m_exp(K(-0.94229) * x);
      const R ex3 = K(0.28022) * m_exp(K(-0.4029) * x);
      const R ex4 = K(0.028171) * m_exp(K(-0.20162) * x);
      vv = (ex1 + ex2 + ex3 + ex4) / x;
      vv1 = -(vv + K(3.1998) * ex1 + K(0.94229) * ex2 + K(0.4029) * ex3 + K(0.20162) * ex4) / x;
    }
  };
  do {
    pot(r, v, v1);
    const R fr = b * b / r + v * r / eps - r;
    const R fr1 = -b * b / (r * r) + (v + v1 * r) / eps - K(1.0);
    q = fr / fr1;
    r -= q;
    ++it;
  } while (m_abs(q / r) > K(0.001) && it < kMaxNewton);
  const R roc = K(-2.0) * (eps - v) / v1;
  const R sqe = m_sqrt(eps);
  R cc, aa, ff;
  if (POT == 1) {
    cc = (K(0.009611) + sqe) / (K(0.005175) + sqe);
    aa = K(2.0) * eps * (K(1.0) + (K(0.6743) / sqe)) * m_pow(b, cc);
    ff = (m_sqrt(aa * aa + K(1.0)) - aa) * ((K(6.314) + eps) / (K(10.0) + eps));
  } else if (POT == 2) {
    cc = (K(0.235809) + sqe) / (K(0.126000) + sqe);
    aa = K(2.0) * eps * (K(1.0) + (K(1.0144) / sqe)) * m_pow(b, cc);
    ff = (m_sqrt(aa * aa + K(1.0)) - aa) * ((K(6935.0) + eps) / (K(83550.0) + eps));
  } else {
    cc = (K(0.011615) + sqe) / (K(0.0071222) + sqe);
    aa = K(2.0) * eps * (K(1.0) + (K(0.99229) / sqe)) * m_pow(b, cc);
    ff = (m_sqrt(aa * aa + K(1.0)) - aa) * ((K(9.3066) + eps) / (K(14.813) + eps));
  }
  R vr, v1r;
  pot(r, vr, v1r);
  const R rmb = vr * r * r / (eps * (r + b));
  const R A = aa * ff / (ff + K(1.0));
  const R omc = rmb * (K(1.0) - A) / (r + roc);
  s2 = omc * (K(2.0) - omc);
  ct = K(1.0) - K(2.0) * s2;
  st = K(2.0) * m_sqrt(mx(K(0.0), s2 * (K(1.0) - s2)));
  iters = it;
}

template <int POT>
__device__ __forceinline__ void scatter_f32(float eps, float b, float& s2, float& ct, float& st, int& iters) {
  scatter_stable<float, POT>(eps, b, s2, ct, st, iters);
}

#undef K

}  // namespace mt
