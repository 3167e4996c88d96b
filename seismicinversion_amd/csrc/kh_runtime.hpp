// Error reporting and launch helpers shared by the C-ABI entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "kennett_hip.h"

namespace kh {

// Per-thread message of the last failed call (defined in kh_runtime.hip).
void set_error(const char* fmt, ...);
void clear_error();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Reports a launch failure, if any, for the kernel just enqueued.
inline int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return KH_ERR_LAUNCH;
  }
  return KH_OK;
}

// Validates the layer table dimensions shared by every entry point.
inline int check_model(const char* what, const double* layers, int nl) {
  if (nl < 2) {
    set_error("%s: need at least 2 layers (water + half-space), got %d", what, nl);
    return KH_ERR_INVALID_ARGUMENT;
  }
  if (nl > KH_MAX_LAYERS) {
    set_error("%s: %d layers exceeds KH_MAX_LAYERS=%d", what, nl, KH_MAX_LAYERS);
    return KH_ERR_UNSUPPORTED;
  }
  if (!layers) {
    set_error("%s: layers is NULL", what);
    return KH_ERR_INVALID_ARGUMENT;
  }
  return KH_OK;
}

}  // namespace kh

#define KH_REQUIRE(cond, code, ...)  \
  do {                               \
    if (!(cond)) {                   \
      ::kh::set_error(__VA_ARGS__);  \
      return (code);                 \
    }                                \
  } while (0)

#define KH_TRY(expr)                 \
  do {                               \
    const int kh_rc_ = (expr);       \
    if (kh_rc_ != KH_OK) return kh_rc_; \
  } while (0)
