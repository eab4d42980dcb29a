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

// precision-dispatched math: FP64 from mt_fmath.h (within a few ulp, a
// third to a half of the library's instructions; the physics passes finite
// arguments, positive ones to log and pow), single precision the
// hardware-assisted variants
__device__ __forceinline__ double m_exp(double x) { return fx_exp_fin(x); }
__device__ __forceinline__ double m_log(double x) { return fx_log_pos(x); }
__device__ __forceinline__ double m_pow(double x, double y) { return fx_pow_pos(x, y); }
__device__ __forceinline__ double m_sqrt(double x) { return fx_sqrt(x); }
__device__ __forceinline__ double m_rcp(double x) { return fx_rcp(x); }
__device__ __forceinline__ float m_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double m_rsqrt(double x) { return fx_rsqrt(x); }
__device__ __forceinline__ float m_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
// sin and cos of 2 pi u, u in [0, 1]; in single precision the hardware
// sine / cosine, whose argument is in revolutions
__device__ __forceinline__ void m_sincos2pi(double u, double& s, double& c) { fx_sincos2pi(u, s, c); }
__device__ __forceinline__ void m_sincos2pi(double u, float& s, float& c) {
  s = __builtin_amdgcn_sinf((float)u);
  c = __builtin_amdgcn_cosf((float)u);
}
__device__ __forceinline__ double m_abs(double x) { return fabs(x); }
// single precision: the hardware exp2 / log2 (about 1 ulp for normal
// arguments; the physics' arguments are normal), pow as exp2(y log2 x)
__device__ __forceinline__ float m_exp(float x) { return __expf(x); }
__device__ __forceinline__ float m_log(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
__device__ __forceinline__ float m_pow(float x, float y) {
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
__device__ __forceinline__ float m_sqrt(float x) { return __fsqrt_rn(x); }
__device__ __forceinline__ float m_abs(float x) { return fabsf(x); }

#define K(x) R(x)

// The stopping fits below keep the reference's formulas; divisions by
// per-(ion, element) constants become multiplications by their reciprocals
// (StopConst.inv_*), powers share one logarithm of their base, and
// (e / eee)^power = exp(power (log e - log eee)) reuses log e.  The caller
// passes log E (shared with the free path): log e = log E + log(0.001 / m1).
template <typename R>
__device__ __forceinline__ R rpstop(const StopConst& c, R e, R le) {  // material.C:134-158
  const R pe = mx(K(25.0), e);
  const R lpe = (e > K(25.0)) ? le : K(3.2188758248682007492);  // log 25
  const R sl = (R(c.pc[0]) * m_exp(R(c.pc[1]) * lpe)) + (R(c.pc[2]) * m_exp(R(c.pc[3]) * lpe));
  const R sh = R(c.pc[4]) * m_exp(-R(c.pc[5]) * lpe) * m_log(R(c.pc[6]) * m_rcp(pe) + R(c.pc[7]) * pe);
  R sp = sl * sh * m_rcp(sl + sh);
  if (e <= K(25.0)) sp *= m_exp(R(c.velpwr) * (le - K(3.2188758248682007492)));  // (e / 25)^velpwr
  return sp;
}

template <typename R>
__device__ __forceinline__ R rstop(const StopConst& c, R E, R lE) {  // material.C:161-282
  const R e = (K(0.001) * E) * R(c.inv_m1);
  const R le = lE + R(c.ln_emul);
  R se;
  if (c.kind == 1) {
    se = rpstop(c, e, le);
  } else if (c.kind == 2) {
    const R he0 = K(1.0);
    const R he = mx(he0, e);
    const R b = (e > he0) ? le : K(0.0);  // log he
    const R b2 = b * b;
    const R b4 = b2 * b2;
    R a = K(0.2865) + K(0.1266) * b - K(0.001429) * b2 + K(0.02402) * b * b2 - K(0.01135) * b4 +
          K(0.001475) * b4 * b;
    R heh = K(1.0) - m_exp(-mn(K(30.0), a));
    a = K(1.0) + R(c.he_k) * m_exp(-sq(K(7.6) - b));
    heh *= a * a;
    se = rpstop(c, he, b) * heh * K(4.0);
    if (e <= he0) se *= m_sqrt(e);
  } else {
    const R vf = R(c.vfermi);
    const R v = m_sqrt(e * K(0.04)) * R(c.inv_vfermi);
    const R v2 = v * v;
    R vr;
    if (v >= K(1.0))
      vr = v * vf * (K(1.0) + m_rcp(K(5.0) * v2));
    else
      vr = R(c.vk) * (K(1.0) + v2 * K(2.0 / 3.0) - v2 * v2 * K(1.0 / 15.0));
    R yr = mx(K(0.13), vr * R(c.inv_cbrt2_fz1));
    yr = mx(yr, R(c.yr_lo));
    // yr is at least yr_branch = max(yrmin, vrmin / cbrt2 fz1); the
    // low-velocity branch (yr <= yr_branch) has yr = yr_branch, where q, l
    // and the effective-charge factor are per-(ion, element) constants
    const bool low = yr <= R(c.yr_branch);
    R zeta;
    if (low) {
      zeta = R(c.zeta_lo);
    } else {
      const R yr03 = m_exp(K(0.3) * m_log(yr));
      const R a = K(-0.803) * yr03 + K(1.3167) * yr03 * yr03 + K(0.38157) * yr + K(0.008983) * yr * yr;
      const R q = mn(K(1.0), mx(K(0.0), K(1.0) - m_exp(-mn(a, K(50.0)))));
      const R l0 = (K(0.8) - q * R(c.l0k)) * R(c.inv_cbrt_fz1);
      R l1;
      if (q < K(0.2))
        l1 = K(0.0);
      else if (q < R(c.q1))
        l1 = R(c.bq) * (q - K(0.2)) * R(c.inv_d1);
      else if (q < R(c.q2))
        l1 = R(c.bq);
      else
        l1 = R(c.bq) * (K(1.0) - q) * R(c.inv_d2);
      const R l = mx(l1, l0 * R(c.lfctr));
      zeta = q + R(c.zk) * (K(1.0) - q) * m_log(K(1.0) + sq(l * vf * K(4.0 / 1.919)));
    }
    const R a = -sq(K(7.6) - mx(K(0.0), le));
    zeta *= K(1.0) + R(c.z3) * m_exp(a);
    if (low)
      se = R(c.sp_low) * sq(zeta * R(c.fz1)) * m_exp(R(c.power) * (le - R(c.ln_eee)));
    else
      se = rpstop(c, e, le) * sq(zeta * R(c.fz1));
  }
  return se * K(10.0);
}

// MaterialBase::getrstop for species s in material k [eV/A]
template <typename R>
__device__ __forceinline__ R getrstop(const TableView& T, int s, int k, R E, R lE) {
  const MatConst& mc = T.mat[s * T.n_mat + k];
  R se = K(0.0);
  for (int i = 0; i < mc.n_el; ++i) {
    const int e = s * T.n_el + mc.el_begin + i;
    se += rstop(T.stop[e], E, lE) * R(T.el[e].t);
  }
  return se * R(mc.arho);
}

template <typename R>
__device__ __forceinline__ R getrstop(const TableView& T, int s, int k, R E) {
  return getrstop<R>(T, s, k, E, m_log(E));
}

// Screening sums of the interatomic potential V(r) = S(r) / r with
// S = sum c_i exp(-d_i r) and D = sum c_i d_i exp(-d_i r) = -S'(r), for the
// universal (POT 0), Moliere (1) and C-Kr (2) screenings of trim.C:205-232.
// The reference's v = V and v1 = V' follow as v = S / r, v1 = -(v + D) / r.
template <typename R, int POT>
__device__ __forceinline__ void screening(R r, R& S, R& D) {
  if (POT == 1) {
    const R ex1 = m_exp(K(-0.3) * r);
    const R ex2 = sq(ex1) * sq(ex1);
    const R ex3 = ex2 * (sq(ex2) * sq(ex2));
    S = K(0.35) * ex1 + K(0.55) * ex2 + K(0.1) * ex3;
    D = K(0.105) * ex1 + K(0.66) * ex2 + K(0.6) * ex3;
  } else if (POT == 2) {
    const R ex1 = m_exp(K(-0.279) * r);
    const R ex2 = m_exp(K(-0.637) * r);
    const R ex3 = m_exp(K(-1.1919) * r);
    S = K(0.191) * ex1 + K(0.474) * ex2 + K(0.335) * ex3;
    D = K(0.531865) * ex1 + K(0.30181) * ex2 + K(0.6437) * ex3;
  } else {
    const R ex1 = K(0.18175) * m_exp(K(-3.1998) * r);
    const R ex2 = K(0.50986) * m_exp(K(-0.94229) * r);
    const R ex3 = K(0.28022) * m_exp(K(-0.4029) * r);
    const R ex4 = K(0.028171) * m_exp(K(-0.20162) * r);
    S = ex1 + ex2 + ex3 + ex4;
    D = K(3.1998) * ex1 + K(0.94229) * ex2 + K(0.4029) * ex3 + K(0.20162) * ex4;
  }
}

// Newton solve for the apsis (trim.C:200-240): f(r) = b^2/r + v r/eps - r,
// f'(r) = -b^2/r^2 + (v + v1 r)/eps - 1, r -= f / f' until |q / r| <= 0.001.
// With v = S/r and v + v1 r = -D the iteration needs one reciprocal of r and
// one of f' (the reference divides seven times):
//   f = b^2 (1/r) + S / eps - r,  f' = -b^2 (1/r)^2 - D / eps - 1.
// Returns the iteration count; v and v1 are those of the last r evaluated
// (before its update), as in the reference.
template <typename R, int POT>
__device__ __forceinline__ int apsis(R eps, R b, R& r, R& v, R& v1) {
  const R bb = b * b, ieps = m_rcp(eps);
  R q;
  int it = 0;
  do {
    R S, D;
    screening<R, POT>(r, S, D);
    const R ir = m_rcp(r);
    v = S * ir;
    v1 = -(v + D) * ir;
    const R fr = bb * ir + S * ieps - r;
    const R fr1 = -(bb * (ir * ir)) - D * ieps - K(1.0);
    q = fr * m_rcp(fr1);
    r -= q;
    ++it;
  } while (m_abs(q) > K(0.001) * m_abs(r) && it < kMaxNewton);
  return it;
}

// MAGIC coefficients (trim.C:236-262): cc, aa = 2 eps (1 + c/sqrt(eps)) b^cc
// and ff for the three screenings; 1/sqrt(eps) comes from one reciprocal
// square root, the quotients from reciprocals, b^cc = exp(cc log b) from the
// log b of the apsis' starting point.  ff carries sqrt(aa^2 + 1) - aa, which
// cancels for large aa (glancing collisions: aa ~ 500 leaves 3 % of a single
// precision result); in single precision it is formed as the identical
// 1 / (sqrt(aa^2 + 1) + aa), in FP64 as the reference writes it.
template <typename R>
__device__ __forceinline__ R magic_root_gap(R aa) {  // sqrt(aa^2 + 1) - aa
  if constexpr (sizeof(R) == sizeof(float))
    return m_rcp(m_sqrt(aa * aa + K(1.0)) + aa);
  else
    return m_sqrt(aa * aa + K(1.0)) - aa;
}

template <typename R, int POT>
__device__ __forceinline__ void magic(R eps, R lb, R& aa, R& ff) {  // lb = log b
  const R ise = m_rsqrt(eps);
  const R sqe = eps * ise;
  R cc;
  if (POT == 1) {
    cc = (K(0.009611) + sqe) * m_rcp(K(0.005175) + sqe);
    aa = K(2.0) * eps * (K(1.0) + K(0.6743) * ise) * m_exp(cc * lb);
    ff = magic_root_gap(aa) * ((K(6.314) + eps) * m_rcp(K(10.0) + eps));
  } else if (POT == 2) {
    cc = (K(0.235809) + sqe) * m_rcp(K(0.126000) + sqe);
    aa = K(2.0) * eps * (K(1.0) + K(1.0144) * ise) * m_exp(cc * lb);
    ff = magic_root_gap(aa) * ((K(6935.0) + eps) * m_rcp(K(83550.0) + eps));
  } else {
    cc = (K(0.011615) + sqe) * m_rcp(K(0.0071222) + sqe);
    aa = K(2.0) * eps * (K(1.0) + K(0.99229) * ise) * m_exp(cc * lb);
    ff = magic_root_gap(aa) * ((K(9.3066) + eps) * m_rcp(K(14.813) + eps));
  }
}

// trim.C:172-272
template <typename R, int POT>
__device__ __forceinline__ void scatter(R eps, R b, R leps, R& s2, R& ct, R& st, int& iters) {
  if (eps > K(10.0)) {
    s2 = m_rcp(K(1.0) + (K(1.0) + b * (K(1.0) + b)) * sq(K(2.0) * eps * b));
    const R c2 = K(1.0) - s2;
    ct = K(2.0) * c2 - K(1.0);
    st = m_sqrt(K(1.0) - ct * ct);
    iters = 0;
    return;
  }
  R r = b;
  const R lb = m_log(b);
  R rr = K(-2.7) * (leps + lb);  // log(eps b)
  if (rr >= b) {
    rr = K(-2.7) * m_log(eps * rr);
    if (rr >= b) r = rr;
  }
  R v, v1;
  const int it = apsis<R, POT>(eps, b, r, v, v1);
  const R roc = K(-2.0) * (eps - v) * m_rcp(v1);
  R aa, ff;
  magic<R, POT>(eps, lb, aa, ff);
  const R delta = (r - b) * aa * ff * m_rcp(ff + K(1.0));
  const R co = (b + delta + roc) * m_rcp(r + roc);
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
__device__ __forceinline__ void scatter_stable(R eps, R b, R leps, R& s2, R& ct, R& st, int& iters) {
  if (eps > K(10.0)) {
    s2 = m_rcp(K(1.0) + (K(1.0) + b * (K(1.0) + b)) * sq(K(2.0) * eps * b));
    ct = K(1.0) - K(2.0) * s2;
    st = K(2.0) * m_sqrt(s2 * (K(1.0) - s2));
    iters = 0;
    return;
  }
  R r = b;
  const R lb = m_log(b);
  R rr = K(-2.7) * (leps + lb);  // log(eps b)
  if (rr >= b) {
    rr = K(-2.7) * m_log(eps * rr);
    if (rr >= b) r = rr;
  }
  R v, v1;
  const int it = apsis<R, POT>(eps, b, r, v, v1);
  const R roc = K(-2.0) * (eps - v) * m_rcp(v1);
  R aa, ff;
  magic<R, POT>(eps, lb, aa, ff);
  R Sr, Dr;
  screening<R, POT>(r, Sr, Dr);
  const R rmb = Sr * r / (eps * (r + b));  // v(r) r^2 = S(r) r
  const R A = aa * ff / (ff + K(1.0));
  const R omc = rmb * (K(1.0) - A) / (r + roc);
  s2 = omc * (K(2.0) - omc);
  ct = K(1.0) - K(2.0) * s2;
  st = K(2.0) * m_sqrt(mx(K(0.0), s2 * (K(1.0) - s2)));
  iters = it;
}

template <typename R, int POT>
__device__ __forceinline__ void scatter(R eps, R b, R& s2, R& ct, R& st, int& iters) {
  scatter<R, POT>(eps, b, m_log(eps), s2, ct, st, iters);
}

template <typename R, int POT>
__device__ __forceinline__ void scatter_stable(R eps, R b, R& s2, R& ct, R& st, int& iters) {
  scatter_stable<R, POT>(eps, b, m_log(eps), s2, ct, st, iters);
}

template <int POT>
__device__ __forceinline__ void scatter_f32(float eps, float b, float leps, float& s2, float& ct, float& st,
                                            int& iters) {
  scatter_stable<float, POT>(eps, b, leps, s2, ct, st, iters);
}

#undef K

}  // namespace mt
