// Shared helpers for the sslab HIP library (gfx950 only).
//
// Every entry point of the C ABI returns an int status (0 = success) and records a
// human-readable message retrievable with sslab_last_error().  Kernels never allocate;
// all device memory (operands, factors, workspaces) is owned by the caller.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

namespace sslab {

// Status codes shared with include/sslab_hip.h.
enum Status : int {
  kOk = 0,
  kBadArgument = 1,
  kHipError = 2,
  kWorkspaceTooSmall = 3,
  kSingular = 4,
  kNotSupported = 5,
};

void set_error(const char* fmt, ...);
const char* get_error();

#define SSLAB_HIP_CHECK(expr)                                                       \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      ::sslab::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,        \
                         hipGetErrorString(_e));                                    \
      return ::sslab::kHipError;                                                    \
    }                                                                               \
  } while (0)

#define SSLAB_REQUIRE(cond, ...)                                                    \
  do {                                                                              \
    if (!(cond)) {                                                                  \
      ::sslab::set_error(__VA_ARGS__);                                              \
      return ::sslab::kBadArgument;                                                 \
    }                                                                               \
  } while (0)

#define SSLAB_TRY(expr)                                                             \
  do {                                                                              \
    int _s = (expr);                                                                \
    if (_s != 0) return _s;                                                         \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace sslab
