// Errors of the MyTRIM API.  Standalone they are C++ exceptions
// (std::runtime_error), where the reference prints and exits; built inside
// MOOSE/Magpie (MYTRIM_ENABLED) they go to mooseError, as the reference's do
// there (material.C:50-71, simconf.C:142-165).
#ifndef MYTRIM_AMD_API_ERROR_H
#define MYTRIM_AMD_API_ERROR_H

#include <stdexcept>
#include <string>

#ifdef MYTRIM_ENABLED
#include "MooseError.h"
#endif

namespace MyTRIM_NS {

[[noreturn]] inline void api_error(const std::string& msg) {
#ifdef MYTRIM_ENABLED
  mooseError(msg);
#else
  throw std::runtime_error(msg);
#endif
}

}  // namespace MyTRIM_NS

#endif
