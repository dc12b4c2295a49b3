// Host-side helpers shared by the C-ABI translation units: the thread-local
// error message, status plumbing and launch checking.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/flagsparse_hip.h"

namespace fsp {

// Defined once in plan.hip.
void set_error(const char* fmt, ...);

inline int fail(int status, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  set_error("%s", buf);
  return status;
}

// Converts the sticky launch error of the last kernel into a status.
inline int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    return fail(FSP_ERR_LAUNCH, "%s: %s", what, hipGetErrorString(e));
  }
  return FSP_OK;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

// Largest grid.x used by one launch; bigger work is covered by a grid-stride
// loop over `blockIdx.x + k * gridDim.x`.
constexpr int64_t kMaxGridX = 1 << 30;

}  // namespace fsp

// Dispatch helpers: bind a runtime enum to a type alias and run the body.
#define FSP_DISPATCH_INDEX(code, NAME, ...)                                  \
  do {                                                                       \
    if ((code) == FSP_I32) {                                                 \
      using NAME = int32_t;                                                  \
      __VA_ARGS__;                                                           \
    } else if ((code) == FSP_I64) {                                          \
      using NAME = int64_t;                                                  \
      __VA_ARGS__;                                                           \
    } else {                                                                 \
      return ::fsp::fail(FSP_ERR_DTYPE, "unsupported index type %d", (int)(code)); \
    }                                                                        \
  } while (0)

#define FSP_DISPATCH_VALUE(code, NAME, ...)                                  \
  do {                                                                       \
    if ((code) == FSP_F32) {                                                 \
      using NAME = float;                                                    \
      __VA_ARGS__;                                                           \
    } else if ((code) == FSP_F64) {                                          \
      using NAME = double;                                                   \
      __VA_ARGS__;                                                           \
    } else if ((code) == FSP_F16) {                                          \
      using NAME = _Float16;                                                 \
      __VA_ARGS__;                                                           \
    } else if ((code) == FSP_BF16) {                                         \
      using NAME = __bf16;                                                   \
      __VA_ARGS__;                                                           \
    } else {                                                                 \
      return ::fsp::fail(FSP_ERR_DTYPE, "unsupported value type %d", (int)(code)); \
    }                                                                        \
  } while (0)
