// host_common.h -- error reporting and device selection shared by the host
// code of libpaintop_hip (solver and warm-start network).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/paintop_hip.h"

namespace ph {

// A failure carried to the C ABI as (error class, code, message).
struct Fail {
  int cls;
  std::string code;
  std::string msg;
};

Fail fail(int cls, const char *code, const char *fmt, ...);
void put_error(ph_error *err, const Fail &f);
void clear_error(ph_error *err);

#define HIP_OK(expr)                                                                   \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      throw ::ph::fail(PH_CLASS_EXECUTION, "E_HIP_RUNTIME", "%s failed: %s (%s:%d)", #expr, \
                       hipGetErrorString(_e), __FILE__, __LINE__);                     \
  } while (0)

// One device per process; every C-ABI call holds g_mu.
struct Ctx {
  bool probed = false;
  bool ok = false;
  std::string why;
  int dev = 0;
  hipDeviceProp_t prop{};
};
extern std::mutex g_mu;
extern int g_forced;
extern Ctx g_ctx;
void ensure_device();

}  // namespace ph
