// internal.hpp -- shared host-side helpers of libvclhip that touch the HIP runtime (status
// conversion, launch checks, the measurement hook); error reporting is in errors.hpp.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "errors.hpp"

namespace vclhip {

// Converts a HIP error into the ABI's status (positive hipError_t) and records its text.
inline int hip_status(hipError_t e, const char* where)
{
  if (e == hipSuccess) return VCLHIP_SUCCESS;
  set_error("%s: %s", where, hipGetErrorString(e));
  return static_cast<int>(e);
}

// Checks the launch that was just enqueued (VIENNACL_CUDA_LAST_ERROR_CHECK, cuda/common.hpp:165-175).
inline int check_launch(const char* kernel)
{
  return hip_status(hipGetLastError(), kernel);
}

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
