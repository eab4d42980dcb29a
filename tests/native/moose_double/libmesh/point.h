// Test double of libMesh's Point (a TypeVector<Real> of three components) with
// the members the MyTRIM API uses: component access, norm, norm_sq and the
// component-wise arithmetic.  Only the interface matters to the test.
#ifndef MT_TEST_LIBMESH_POINT_H
#define MT_TEST_LIBMESH_POINT_H

#include <cmath>

namespace libMesh {

typedef double Real;

class Point {
 public:
  Point(Real x = 0.0, Real y = 0.0, Real z = 0.0) : _coords{x, y, z} {}
  Real& operator()(unsigned int i) { return _coords[i]; }
  const Real& operator()(unsigned int i) const { return _coords[i]; }
  Real norm_sq() const { return _coords[0] * _coords[0] + _coords[1] * _coords[1] + _coords[2] * _coords[2]; }
  Real norm() const { return std::sqrt(norm_sq()); }
  Point operator+(const Point& p) const { return Point(_coords[0] + p(0), _coords[1] + p(1), _coords[2] + p(2)); }
  Point operator-(const Point& p) const { return Point(_coords[0] - p(0), _coords[1] - p(1), _coords[2] - p(2)); }
  Point operator*(Real f) const { return Point(_coords[0] * f, _coords[1] * f, _coords[2] * f); }
  Point operator/(Real f) const { return Point(_coords[0] / f, _coords[1] / f, _coords[2] / f); }
  Point operator-() const { return Point(-_coords[0], -_coords[1], -_coords[2]); }
  Point& operator+=(const Point& p) { return *this = *this + p; }
  Point& operator-=(const Point& p) { return *this = *this - p; }
  Point& operator*=(Real f) { return *this = *this * f; }
  Point& operator/=(Real f) { return *this = *this / f; }

 private:
  Real _coords[3];
};

}  // namespace libMesh

#endif
