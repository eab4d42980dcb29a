// Point and Real of the MyTRIM API (shim/point.h, simconf.h:39): a 3-vector
// with component access and component-wise arithmetic.  Built inside
// MOOSE/Magpie (MYTRIM_ENABLED) they, and Utility::pow, come from MOOSE and
// libMesh instead, as in the reference (simconf.h:31-36); the API uses only
// the members both provide.
#ifndef MYTRIM_AMD_POINT_H
#define MYTRIM_AMD_POINT_H

#ifdef MYTRIM_ENABLED
#include "MooseError.h"
#include "MooseTypes.h"
#include "libmesh/point.h"
#include "libmesh/utility.h"
#else

#include <cmath>
#include <type_traits>

typedef double Real;

class Point {
 public:
  Point() : data{0.0, 0.0, 0.0} {}
  Point(Real x, Real y, Real z) : data{x, y, z} {}
  Real& operator()(unsigned int i) { return data[i]; }
  const Real& operator()(unsigned int i) const { return data[i]; }
  Real norm_sq() const { return data[0] * data[0] + data[1] * data[1] + data[2] * data[2]; }
  Real norm() const { return std::sqrt(norm_sq()); }
  Point operator+(const Point& r) const { return Point(data[0] + r.data[0], data[1] + r.data[1], data[2] + r.data[2]); }
  Point operator-(const Point& r) const { return Point(data[0] - r.data[0], data[1] - r.data[1], data[2] - r.data[2]); }
  Point operator*(Real r) const { return Point(data[0] * r, data[1] * r, data[2] * r); }
  Point operator/(Real r) const { return Point(data[0] / r, data[1] / r, data[2] / r); }
  Point operator-() const { return Point(-data[0], -data[1], -data[2]); }
  Point& operator+=(const Point& r) {
    for (int i = 0; i < 3; ++i) data[i] += r.data[i];
    return *this;
  }
  Point& operator-=(const Point& r) {
    for (int i = 0; i < 3; ++i) data[i] -= r.data[i];
    return *this;
  }
  Point& operator*=(Real r) {
    for (int i = 0; i < 3; ++i) data[i] *= r;
    return *this;
  }
  Point& operator/=(Real r) {
    for (int i = 0; i < 3; ++i) data[i] /= r;
    return *this;
  }

 private:
  Real data[3];
};

namespace Utility {
namespace detail {
// x^N in the order of shim/pow.h: an odd exponent takes one factor of x off,
// an even one squares x^(N/2), and x^6 is x^4 x^2 (C++11: overloads on the
// exponent instead of if constexpr, so that the API's C++11 callers build)
inline Real ipow(const Real&, std::integral_constant<int, 0>) { return 1.0; }
inline Real ipow(const Real& x, std::integral_constant<int, 1>) { return x; }
inline Real ipow(const Real& x, std::integral_constant<int, 6>) {
  const Real x2 = x * x, x4 = x2 * x2;
  return x4 * x2;
}
template <int N>
inline Real ipow(const Real& x, std::integral_constant<int, N>);
template <int N>
inline Real ipow_step(const Real& x, std::true_type) {  // N odd
  return x * ipow(x, std::integral_constant<int, N - 1>());
}
template <int N>
inline Real ipow_step(const Real& x, std::false_type) {  // N even
  const Real h = ipow(x, std::integral_constant<int, N / 2>());
  return h * h;
}
template <int N>
inline Real ipow(const Real& x, std::integral_constant<int, N>) {
  return ipow_step<N>(x, std::integral_constant<bool, (N % 2 == 1)>());
}
}  // namespace detail

template <int N>
inline Real pow(const Real& x) {
  return detail::ipow(x, std::integral_constant<int, N>());
}
}  // namespace Utility

#endif  // MYTRIM_ENABLED

#endif
