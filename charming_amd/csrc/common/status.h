// Host-side status plumbing shared by the charming_amd libraries: a
// per-thread last-error message and HIP call checking that turns failures
// into negative return codes instead of exit(-1) (src/util/util.h:6-12).
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "chm_status.h"

namespace chm {

inline char* last_error_buf() {
  static thread_local char buf[1024] = {0};
  return buf;
}

inline int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(last_error_buf(), 1024, fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace chm

#define CHM_TRY_HIP(call)                                                   \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return ::chm::fail(CHM_E_HIP, "%s failed: %s (%s:%d)", #call,         \
                         hipGetErrorString(e_), __FILE__, __LINE__);        \
  } while (0)

#define CHM_TRY(expr)      \
  do {                     \
    int rc_ = (expr);      \
    if (rc_ < 0) return rc_; \
  } while (0)
