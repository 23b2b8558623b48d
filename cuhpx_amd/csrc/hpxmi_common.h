// Shared helpers for the hpxmi HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace hpxmi {

// Thread-local last-error text, returned by hpx_last_error().
void set_error(const std::string& msg);
const char* last_error();

}  // namespace hpxmi

#define HPX_FAIL(code, ...)                                    \
  do {                                                         \
    char _buf[512];                                            \
    std::snprintf(_buf, sizeof(_buf), __VA_ARGS__);            \
    hpxmi::set_error(_buf);                                    \
    return (code);                                             \
  } while (0)

#define HPX_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess) {                                                  \
      HPX_FAIL(HPX_ERR_HIP, "%s failed: %s (%s:%d)", #call,                  \
               hipGetErrorString(_e), __FILE__, __LINE__);                   \
    }                                                                        \
  } while (0)

#define HPX_LAUNCH_CHECK()                                                   \
  do {                                                                       \
    hipError_t _e = hipGetLastError();                                       \
    if (_e != hipSuccess) {                                                  \
      HPX_FAIL(HPX_ERR_HIP, "kernel launch failed: %s (%s:%d)",              \
               hipGetErrorString(_e), __FILE__, __LINE__);                   \
    }                                                                        \
  } while (0)

// Work-group index remapped so that consecutive logical work items land on
// the same XCD (MI355X dispatches work-groups round-robin over 8 XCDs, each
// with its own L2).  Valid when n % 8 == 0; identity otherwise.
__device__ __forceinline__ int xcd_local_index(int bid, int n) {
  if ((n & 7) != 0) return bid;
  return (bid & 7) * (n >> 3) + (bid >> 3);
}
// Grid size for n work items launched with xcd_local_index(blockIdx.x,
// gridDim.x): n rounded up to a multiple of 8, so that the remapping always
// applies; work-groups whose item is >= n exit at once.
inline unsigned xcd_grid(int64_t n) { return (unsigned)((n + 7) & ~int64_t(7)); }
