// internal.hpp -- shared host-side helpers of libvclhip (error reporting, argument checks).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "vclhip.h"

namespace vclhip {

// Thread-local text of the last failure; returned by vclhip_last_error().
void set_error(const char* fmt, ...);
const char* last_error();

// Converts a HIP error into the ABI's status (positive hipError_t) and records its text.
inline int hip_status(hipError_t e, const char* where)
{
  if (e == hipSuccess) return VCLHIP_SUCCESS;
  set_error("%s: %s", where, hipGetErrorString(e));
  return static_cast<int>(e);
}

inline int invalid(const char* what)
{
  set_error("invalid argument: %s", what);
  return VCLHIP_ERROR_INVALID_VALUE;
}

// Checks the launch that was just enqueued (VIENNACL_CUDA_LAST_ERROR_CHECK, cuda/common.hpp:165-175).
inline int check_launch(const char* kernel)
{
  return hip_status(hipGetLastError(), kernel);
}

#define VCLHIP_TRY(expr)              \
  do {                                \
    int vclhip_status_ = (expr);      \
    if (vclhip_status_ != 0)          \
      return vclhip_status_;          \
  } while (0)

#define VCLHIP_HIP(expr) VCLHIP_TRY(::vclhip::hip_status((expr), #expr))

// Number of compute units of the current device (cached per device).
int device_cu_count();

// Measurement hook (vclhip_time_next_launch): events that the next SpMV or streaming kernel
// launched by this thread records at the kernel's own start and end.  Taking them clears them.
struct LaunchTiming {
  hipEvent_t start = nullptr;
  hipEvent_t stop = nullptr;
};
LaunchTiming take_launch_timing();

// Launches `kernel`, through hipExtLaunchKernelGGL when a timing request is pending.
template <typename F, typename... Args>
inline void launch_timed(F kernel, dim3 grid, dim3 block, hipStream_t s, Args... args)
{
  const LaunchTiming t = take_launch_timing();
  if (t.start || t.stop)
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, t.start, t.stop, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

}  // namespace vclhip
