// runtime.hip -- the HIP memory domain of libvclhip: device selection, allocation, copies,
// streams and events.
#include <cstring>
#include <mutex>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace vclhip {

static thread_local LaunchTiming g_timing;

LaunchTiming take_launch_timing()
{
  const LaunchTiming t = g_timing;
  g_timing = LaunchTiming();
  return t;
}

int device_cu_count()
{
  static std::mutex mu;
  static std::vector<int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  std::lock_guard<std::mutex> lock(mu);
  if (static_cast<int>(cache.size()) <= dev) cache.resize(dev + 1, 0);
  if (cache[dev] == 0) {
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) cu = 256;
    cache[dev] = cu;
  }
  return cache[dev];
}

__global__ __launch_bounds__(kBlock) void fill_bytes_kernel(uint8_t* __restrict__ p, size_t bytes)
{
  // 16-byte stores over the aligned body, byte stores for the head and tail
  const size_t tid = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
  const uintptr_t addr = reinterpret_cast<uintptr_t>(p);
  const size_t head = ((16 - (addr & 15)) & 15) < bytes ? ((16 - (addr & 15)) & 15) : bytes;
  const size_t body = (bytes - head) / 16;
  for (size_t i = tid; i < head; i += stride) p[i] = 0;
  uint4* q = reinterpret_cast<uint4*>(p + head);
  for (size_t i = tid; i < body; i += stride) q[i] = make_uint4(0, 0, 0, 0);
  for (size_t i = head + body * 16 + tid; i < bytes; i += stride) p[i] = 0;
}

}  // namespace vclhip

extern "C" {

const char* vclhip_version(void) { return "vclhip 0.1 (gfx950)"; }

const char* vclhip_status_string(int status)
{
  switch (status) {
    case VCLHIP_SUCCESS: return "success";
    case VCLHIP_ERROR_INVALID_VALUE: return "invalid value";
    case VCLHIP_ERROR_ZERO_DIAGONAL: return "zero on diagonal";
    case VCLHIP_ERROR_NOT_SUPPORTED: return "not supported";
    case VCLHIP_ERROR_NO_DEVICE: return "no HIP device";
    case VCLHIP_ERROR_TOO_LARGE: return "size exceeds the documented maximum";
    default: return status > 0 ? hipGetErrorString(static_cast<hipError_t>(status)) : "unknown status";
  }
}

int vclhip_device_count(int* count)
{
  if (!count) return vclhip::invalid("count is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e == hipErrorNoDevice || (e == hipSuccess && c == 0)) {
    *count = 0;
    return VCLHIP_SUCCESS;
  }
  VCLHIP_HIP(e);
  *count = c;
  return VCLHIP_SUCCESS;
}

int vclhip_set_device(int device)
{
  VCLHIP_HIP(hipSetDevice(device));
  return VCLHIP_SUCCESS;
}

int vclhip_get_device(int* device)
{
  if (!device) return vclhip::invalid("device is NULL");
  VCLHIP_HIP(hipGetDevice(device));
  return VCLHIP_SUCCESS;
}

int vclhip_device_info(int device, char* name, size_t name_len, size_t* total_mem, int* num_cu, int* l2_bytes,
                       int* llc_bytes)
{
  hipDeviceProp_t prop;
  VCLHIP_HIP(hipGetDeviceProperties(&prop, device));
  if (name && name_len) {
    std::strncpy(name, prop.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (total_mem) *total_mem = prop.totalGlobalMem;
  if (num_cu) *num_cu = prop.multiProcessorCount;
  if (l2_bytes) *l2_bytes = prop.l2CacheSize;
  if (llc_bytes) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeL2CacheSize, device) != hipSuccess) v = 0;
    *llc_bytes = v;
  }
  return VCLHIP_SUCCESS;
}

int vclhip_malloc(void** ptr, size_t bytes)
{
  if (!ptr) return vclhip::invalid("ptr is NULL");
  *ptr = nullptr;
  if (bytes == 0) return VCLHIP_SUCCESS;
  VCLHIP_HIP(hipMalloc(ptr, bytes));
  return VCLHIP_SUCCESS;
}

int vclhip_free(void* ptr)
{
  if (!ptr) return VCLHIP_SUCCESS;
  VCLHIP_HIP(hipFree(ptr));
  return VCLHIP_SUCCESS;
}

int vclhip_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream)
{
  if (bytes == 0) return VCLHIP_SUCCESS;
  if (!dst || !src) return vclhip::invalid("null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  VCLHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  VCLHIP_HIP(hipStreamSynchronize(s));
  return VCLHIP_SUCCESS;
}

int vclhip_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream)
{
  if (bytes == 0) return VCLHIP_SUCCESS;
  if (!dst || !src) return vclhip::invalid("null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  VCLHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  VCLHIP_HIP(hipStreamSynchronize(s));
  return VCLHIP_SUCCESS;
}

int vclhip_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream)
{
  if (bytes == 0) return VCLHIP_SUCCESS;
  if (!dst || !src) return vclhip::invalid("null pointer");
  VCLHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
  return VCLHIP_SUCCESS;
}

int vclhip_fill_zero(void* dst, size_t bytes, void* stream)
{
  if (bytes == 0) return VCLHIP_SUCCESS;
  if (!dst) return vclhip::invalid("null pointer");
  size_t work = bytes / 16 + 1;
  size_t grid = (work + vclhip::kBlock - 1) / vclhip::kBlock;
  if (grid > 4096) grid = 4096;
  hipLaunchKernelGGL(vclhip::fill_bytes_kernel, dim3(static_cast<unsigned>(grid)), dim3(vclhip::kBlock), 0,
                     static_cast<hipStream_t>(stream), static_cast<uint8_t*>(dst), bytes);
  return vclhip::check_launch("fill_bytes_kernel");
}

int vclhip_stream_create(void** stream)
{
  if (!stream) return vclhip::invalid("stream is NULL");
  hipStream_t s;
  VCLHIP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return VCLHIP_SUCCESS;
}

int vclhip_stream_destroy(void* stream)
{
  if (!stream) return VCLHIP_SUCCESS;
  VCLHIP_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return VCLHIP_SUCCESS;
}

int vclhip_stream_synchronize(void* stream)
{
  VCLHIP_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return VCLHIP_SUCCESS;
}

int vclhip_device_synchronize(void)
{
  VCLHIP_HIP(hipDeviceSynchronize());
  return VCLHIP_SUCCESS;
}

int vclhip_time_next_launch(void* start_event, void* stop_event)
{
  vclhip::g_timing.start = static_cast<hipEvent_t>(start_event);
  vclhip::g_timing.stop = static_cast<hipEvent_t>(stop_event);
  return VCLHIP_SUCCESS;
}

int vclhip_event_create(void** event)
{
  if (!event) return vclhip::invalid("event is NULL");
  hipEvent_t e;
  VCLHIP_HIP(hipEventCreate(&e));
  *event = e;
  return VCLHIP_SUCCESS;
}

int vclhip_event_destroy(void* event)
{
  if (!event) return VCLHIP_SUCCESS;
  VCLHIP_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return VCLHIP_SUCCESS;
}

int vclhip_event_record(void* event, void* stream)
{
  if (!event) return vclhip::invalid("event is NULL");
  VCLHIP_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
  return VCLHIP_SUCCESS;
}

int vclhip_event_elapsed_ms(void* start, void* stop, float* ms)
{
  if (!start || !stop || !ms) return vclhip::invalid("null argument");
  VCLHIP_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  VCLHIP_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
  return VCLHIP_SUCCESS;
}

}  // extern "C"
