// Vector helpers of the MyTRIM API (functions.h).
#ifndef MYTRIM_AMD_FUNCTIONS_H
#define MYTRIM_AMD_FUNCTIONS_H

#include "point.h"

namespace MyTRIM_NS {

inline void v_cross(const Point& a, const Point& b, Point& c) {
  for (unsigned int i = 0; i < 3; ++i) c(i) = a((i + 1) % 3) * b((i + 2) % 3) - a((i + 2) % 3) * b((i + 1) % 3);
}
inline void v_cross(const Real* a, const Real* b, Real* c) {
  for (unsigned int i = 0; i < 3; ++i) c[i] = a[(i + 1) % 3] * b[(i + 2) % 3] - a[(i + 2) % 3] * b[(i + 1) % 3];
}
inline Real v_dot(const Real* a, const Real* b) {
  Real s = 0.0;
  for (int i = 0; i < 3; ++i) s += a[i] * b[i];
  return s;
}
inline void v_scale(Real* a, Real b) {
  for (int i = 0; i < 3; ++i) a[i] = a[i] * b;
}
inline void v_norm(Real* a, Real b = 1.0) { v_scale(a, b / std::sqrt(v_dot(a, a))); }
inline void v_norm(Point& a, Real b = 1.0) { a *= b / a.norm(); }
inline Real sqr(Real a) { return a * a; }
inline Real cub(Real a) { return a * a * a; }

}  // namespace MyTRIM_NS

#endif
