// common.hpp — host-side plumbing shared by the libgcq translation units:
// error reporting across the C ABI and HIP call checking.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/gcq.h"

namespace gcqh {

void set_error(const char* fmt, ...);

// Evaluate a HIP call; on failure record the message and return GCQ_EHIP
// (or GCQ_ENOMEM for allocation failures) from the enclosing function.
#define GCQ_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      gcqh::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return e_ == hipErrorOutOfMemory ? GCQ_ENOMEM : GCQ_EHIP;                        \
    }                                                                                  \
  } while (0)

#define GCQ_REQUIRE(cond, ...)          \
  do {                                  \
    if (!(cond)) {                      \
      gcqh::set_error(__VA_ARGS__);     \
      return GCQ_EINVAL;                \
    }                                   \
  } while (0)

inline double ms_between(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  return (double)ms;
}

}  // namespace gcqh
