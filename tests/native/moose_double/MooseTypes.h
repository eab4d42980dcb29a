// Test double of MOOSE's type header: libMesh's Real and Point visible
// unqualified, as they are to code built inside MOOSE.
#ifndef MT_TEST_MOOSE_TYPES_H
#define MT_TEST_MOOSE_TYPES_H

#include "libmesh/point.h"
#include "libmesh/utility.h"

using namespace libMesh;

#endif
