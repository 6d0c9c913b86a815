// Host-side helpers shared by the C ABI translation units: error state, HIP error checking,
// device allocation.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/b12x_c_api.h"

namespace b12x {

inline std::string& lastError() {
  static thread_local std::string e;
  return e;
}

inline int setErr(int code, const std::string& msg) {
  lastError() = msg;
  return code;
}

struct HipFail {
  int code;
  std::string msg;
  HipFail(int c, std::string m) : code(c), msg(std::move(m)) {}
};

#define HIPCHK(call)                                                                            \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      throw ::b12x::HipFail(e_ == hipErrorOutOfMemory ? B12X_E_NOMEM : B12X_E_HIP,              \
                            std::string(#call) + " failed: " + hipGetErrorString(e_));          \
    }                                                                                           \
  } while (0)

template <class F>
int guarded(F f) {
  try {
    return f();
  } catch (const HipFail& e) {
    return setErr(e.code, e.msg);
  } catch (const std::exception& e) {
    return setErr(B12X_E_ARG, e.what());
  }
}

inline int ensureDevice() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return setErr(B12X_E_NODEVICE, "no HIP device available");
  return B12X_OK;
}

template <class T>
T* dalloc(size_t count) {
  void* p = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc(&p, count * sizeof(T)));
  return static_cast<T*>(p);
}

inline void dfree(void* p) {
  if (p) (void)hipFree(p);
}

inline uint32_t ceilDiv(uint64_t a, uint64_t b) { return uint32_t((a + b - 1) / b); }

}  // namespace b12x
