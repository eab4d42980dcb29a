// ORACLE — test infrastructure only.
//
// A double that counts the floating-point operations done on it, used to
// freeze the algorithmic work per collision of the reference's per-collision
// physics (the numerator of the cascade kernel's roofline).  Only values that
// depend on the ion's state are `Counted`; constants of the (ion, target)
// pair stay plain doubles, so their arithmetic is not counted.
#pragma once
#include <cmath>

namespace orc {

struct OpCount {
  long long add = 0, mul = 0, div = 0, cmp = 0, sqrt_ = 0, exp_ = 0, log_ = 0, pow_ = 0,
            cbrt_ = 0, sin_ = 0, cos_ = 0, atan2_ = 0;
  // FP64 flop-equivalents: add/sub/mul/cmp/min/max = 1, div = sqrt = 10,
  // exp = log = 20, pow = 40, cbrt = 30, sin = cos = 25, atan2 = 40
  double flop_eq() const {
    return double(add + mul + cmp) + 10.0 * double(div + sqrt_) + 20.0 * double(exp_ + log_) +
           40.0 * double(pow_ + atan2_) + 30.0 * double(cbrt_) + 25.0 * double(sin_ + cos_);
  }
};

extern thread_local OpCount* g_ops;

struct Counted {
  double v;
  Counted() : v(0.0) {}
  Counted(double x) : v(x) {}  // NOLINT: implicit from constants
  explicit operator double() const { return v; }
};

inline Counted operator+(Counted a, Counted b) { g_ops->add++; return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { g_ops->add++; return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { g_ops->mul++; return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { g_ops->div++; return Counted(a.v / b.v); }
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted& operator+=(Counted& a, Counted b) { a = a + b; return a; }
inline Counted& operator-=(Counted& a, Counted b) { a = a - b; return a; }
inline Counted& operator*=(Counted& a, Counted b) { a = a * b; return a; }
inline Counted& operator/=(Counted& a, Counted b) { a = a / b; return a; }
inline bool operator<(Counted a, Counted b) { g_ops->cmp++; return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { g_ops->cmp++; return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { g_ops->cmp++; return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { g_ops->cmp++; return a.v >= b.v; }
inline Counted exp(Counted a) { g_ops->exp_++; return Counted(std::exp(a.v)); }
inline Counted log(Counted a) { g_ops->log_++; return Counted(std::log(a.v)); }
inline Counted sqrt(Counted a) { g_ops->sqrt_++; return Counted(std::sqrt(a.v)); }
inline Counted pow(Counted a, Counted b) { g_ops->pow_++; return Counted(std::pow(a.v, b.v)); }
inline Counted abs(Counted a) { return Counted(std::fabs(a.v)); }

}  // namespace orc
