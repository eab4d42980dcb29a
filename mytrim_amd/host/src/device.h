// Internal: device contexts of the C++ API and the upload of a sample to
// the C ABI (mt_target).  Not part of the public headers.
#ifndef MYTRIM_AMD_DEVICE_H
#define MYTRIM_AMD_DEVICE_H

#include <mutex>
#include <vector>

#include "material.h"
#include "sample.h"
#include "simconf.h"

struct mt_ctx;

namespace MyTRIM_NS {
namespace device {

// device ordinal used by every context of the API ($MYTRIM_DEVICE, default 0)
int ordinal();
// throws std::runtime_error with the C ABI's message when rc != 0
void check(int rc, const char* what);
mt_ctx* create_context(SimconfType* simconf);

// single-material context for MaterialBase::getrstop
struct StoppingContext {
  mt_ctx* ctx = nullptr;
  std::mutex mu;
  std::vector<double> sig;
  void upload(const MaterialBase& m);
};
StoppingContext& stopping_context(SimconfType* simconf);

// signature of everything the device tables depend on
std::vector<double> target_signature(SampleBase* s, SimconfType* simconf, int potential);
// upload `s` to ctx (throws if the sample kind has no device geometry)
void upload_target(mt_ctx* ctx, SampleBase* s, SimconfType* simconf, int potential);

}  // namespace device
}  // namespace MyTRIM_NS

#endif
