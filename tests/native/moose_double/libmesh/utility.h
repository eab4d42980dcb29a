// Test double of libMesh's Utility::pow<N> (an integer power).
#ifndef MT_TEST_LIBMESH_UTILITY_H
#define MT_TEST_LIBMESH_UTILITY_H

namespace libMesh {
namespace Utility {

template <int N, typename T>
inline T pow(const T& x) {
  T r = 1;
  for (int i = 0; i < N; ++i) r *= x;
  return r;
}

}  // namespace Utility
}  // namespace libMesh

#endif
