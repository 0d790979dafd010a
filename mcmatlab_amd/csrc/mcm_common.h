// Shared host utilities of libmcm: thread-local error text, HIP call checks,
// device selection.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "mcm_abi.h"

namespace mcm {

void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

// Evaluate a HIP call; on failure record the error and return MCM_ERR_DEVICE
// (or MCM_ERR_OUT_OF_MEMORY) from the enclosing function.
#define MCM_HIP(call)                                                                               \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) {                                                                         \
      return ::mcm::fail(e_ == hipErrorOutOfMemory ? MCM_ERR_OUT_OF_MEMORY : MCM_ERR_DEVICE,        \
                         "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
    }                                                                                               \
  } while (0)

// Select `device` and check it is a gfx950 part.
int use_device(int device);
int cu_count(int device);

}  // namespace mcm
