// The C ABI's error reporting, shared by the library's entry points: a failure is a Failure (an error
// code of include/celv2hip.h and a message); guarded() runs an entry point's body and turns what it
// throws into the code, keeping the message for celv2hip_last_error (per thread).
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "celv2hip.h"

namespace celv2host {

inline thread_local std::string g_error;

struct Failure : std::runtime_error {
  int code;
  Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Failure(e == hipErrorOutOfMemory ? CELV2_ERR_MEMORY : CELV2_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return CELV2_OK;
  } catch (const Failure& e) {
    g_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_error = e.what();
    return CELV2_ERR_STATE;
  }
}

}  // namespace celv2host
