// Library-wide entry points of libmcm (errors, version, devices).
#include "mcm_common.h"

#include <cstdlib>
#include <cstring>

namespace mcm {

static thread_local char g_error[1024];

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof g_error, fmt, ap);
  va_end(ap);
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof g_error, fmt, ap);
  va_end(ap);
  return code;
}

int use_device(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(MCM_ERR_NO_DEVICE, "no HIP device visible");
  if (device < 0 || device >= n) return fail(MCM_ERR_INVALID_ARGUMENT, "GPUdevice %d out of range (0..%d)", device, n - 1);
  MCM_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  MCM_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(MCM_ERR_NO_DEVICE, "device %d is %s; this build targets gfx950", device, prop.gcnArchName);
  return MCM_OK;
}

int cu_count(int device) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 256;
  return cus;
}

}  // namespace mcm

extern "C" {

const char* mcm_last_error(void) { return mcm::g_error; }

int mcm_version(void) { return MCM_ABI_VERSION; }

int mcm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  int good = 0;
  for (int d = 0; d < n; ++d) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) ++good;
  }
  return good;
}

void mcm_free(void* p) { std::free(p); }

}  // extern "C"
