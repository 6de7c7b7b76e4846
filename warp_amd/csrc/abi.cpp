// Library-wide pieces of the C ABI: per-thread error messages and the
// device query.
#include "wm_internal.h"
#include "../../include/warpmi.h"

#include <cstring>
#include <string>

static thread_local std::string g_last_error;

void wm_set_error(const std::string &msg) { g_last_error = msg; }

extern "C" {

const char *wm_last_error(void) { return g_last_error.c_str(); }

int wm_device_info(char *name, int64_t namelen, int64_t *cus) {
  WM_ENTRY({
    int dev = 0;
    WM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    WM_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && namelen > 0) {
      std::strncpy(name, prop.gcnArchName, (size_t)namelen - 1);
      name[namelen - 1] = 0;
    }
    if (cus) *cus = prop.multiProcessorCount;
  });
}

}  // extern "C"
