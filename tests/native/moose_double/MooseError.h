// Test double of MOOSE's error interface, for building the C++ API as inside
// MOOSE/Magpie (MYTRIM_ENABLED) where MOOSE is absent: mooseError(args...)
// with the message streamed from its arguments.  MOOSE prints the message and
// aborts; this double throws std::runtime_error so a test can read it.
#ifndef MT_TEST_MOOSE_ERROR_H
#define MT_TEST_MOOSE_ERROR_H

#include <sstream>
#include <stdexcept>

template <typename... Args>
[[noreturn]] void mooseError(Args&&... args) {
  std::ostringstream s;
  (s << ... << args);
  throw std::runtime_error("mooseError: " + s.str());
}

#endif
