// Shared helpers of libdnx_fdtd: error reporting and launch utilities.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/dnx_fdtd.h"

namespace dnx {

// Thread-local message behind dnx_last_error().
void set_error(const char* fmt, ...);
const char* last_error();

// Process-wide kernel-selection option (DNX_OPT_*): the environment is read
// once, at the first query; dnx_set_option changes it afterwards.
int option(int id);

#define DNX_CHECK_HIP(expr)                                                      \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      ::dnx::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,       \
                       hipGetErrorString(_e));                                   \
      return DNX_EHIP;                                                           \
    }                                                                            \
  } while (0)

#define DNX_REQUIRE(cond, ...)            \
  do {                                    \
    if (!(cond)) {                        \
      ::dnx::set_error(__VA_ARGS__);      \
      return DNX_EINVAL;                  \
    }                                     \
  } while (0)

inline int64_t box_cells(const dnx_box& b) {
  return int64_t(b.z1 - b.z0) * (b.y1 - b.y0) * (b.x1 - b.x0);
}

inline bool box_ok(const dnx_box& b, int nz, int ny, int nx) {
  return b.z0 >= 0 && b.z1 <= nz && b.z0 < b.z1 && b.y0 >= 0 && b.y1 <= ny && b.y0 < b.y1 &&
         b.x0 >= 0 && b.x1 <= nx && b.x0 < b.x1;
}

inline int popcount3(int m) { return (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1); }

inline unsigned cdiv(int64_t a, int64_t b) { return unsigned((a + b - 1) / b); }

// Box geometry passed by value to kernels.
struct Box {
  int z0, y0, x0, nz, ny, nx;
  __host__ __device__ int64_t cells() const { return int64_t(nz) * ny * nx; }
};

inline Box make_box(const dnx_box& b) {
  return Box{b.z0, b.y0, b.x0, b.z1 - b.z0, b.y1 - b.y0, b.x1 - b.x0};
}

// Optional per-launch timing of the hot kernels (dnx_timing_*): when enabled,
// timed_begin records a start event on the launch stream and timed_end a stop
// event filed under `name`.  Disabled: both are no-ops.
struct TimedLaunch {
  hipEvent_t start = nullptr;
};
TimedLaunch timed_begin(const char* name, hipStream_t s);
void timed_end(const char* name, hipStream_t s, TimedLaunch& t);

// Validate sources / monitors against a grid (2D callers pass nz = 1).
int check_sources(const dnx_source* s, int n, int nz, int ny, int nx, int max_comp);
int check_monitors(const dnx_monitor* m, int n, int nz, int ny, int nx, int comp_bits,
                   bool need_grad);

}  // namespace dnx
