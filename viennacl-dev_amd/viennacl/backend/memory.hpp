#ifndef VIENNACL_BACKEND_MEMORY_HPP
#define VIENNACL_BACKEND_MEMORY_HPP
// memory.hpp -- memory_create / copy / write / read / finish for the HIP and host domains
// (reference backend/memory.hpp:54-292, :367-464).  Device traffic goes through the C ABI.

#include <cstring>

#include "viennacl/backend/mem_handle.hpp"
#include "viennacl/context.hpp"

namespace viennacl
{
namespace backend
{

/** Blocks until the HIP work of the calling thread's stream has finished (memory.hpp:54-62). */
inline void finish()
{
  hip::check(vclhip_stream_synchronize(hip::current_stream()), "finish");
}

/** Allocates `bytes` in ctx's domain, optionally filled from host memory (memory.hpp:87-119). */
inline void memory_create(mem_handle& handle, vcl_size_t bytes, viennacl::context const& ctx,
                          const void* host_ptr = nullptr)
{
  if (bytes == 0) {
    handle = mem_handle();
    handle.switch_active_handle_id(ctx.memory_type());
    return;
  }
  if (ctx.memory_type() == MAIN_MEMORY) {
    handle.reset_ram(bytes);
    if (host_ptr) std::memcpy(handle.ram_handle(), host_ptr, bytes);
    return;
  }
  if (ctx.memory_type() != HIP_MEMORY) throw memory_exception("unknown memory domain");
  void* p = nullptr;
  hip::check(vclhip_malloc(&p, bytes), "memory_create");
  handle.reset_hip(p, bytes);
  if (host_ptr)
    hip::check(vclhip_memcpy_h2d(p, host_ptr, bytes, hip::current_stream()), "memory_create (write)");
  else
    hip::check(vclhip_fill_zero(p, bytes, hip::current_stream()), "memory_create (zero)");
}

/** dst[dst_offset..] = src[src_offset..] for `bytes` bytes (memory.hpp:140-171). */
inline void memory_copy(mem_handle const& src, mem_handle& dst, vcl_size_t src_offset, vcl_size_t dst_offset,
                        vcl_size_t bytes)
{
  if (bytes == 0) return;
  if (src.get_active_handle_id() == HIP_MEMORY && dst.get_active_handle_id() == HIP_MEMORY) {
    hip::check(vclhip_memcpy_d2d(static_cast<char*>(dst.hip_handle()) + dst_offset,
                                 static_cast<const char*>(src.hip_handle()) + src_offset, bytes, hip::current_stream()),
               "memory_copy");
  } else if (src.get_active_handle_id() == MAIN_MEMORY && dst.get_active_handle_id() == MAIN_MEMORY) {
    std::memmove(dst.ram_handle() + dst_offset, src.ram_handle() + src_offset, bytes);
  } else {
    throw memory_exception("memory_copy between different domains");
  }
}

/** Host -> buffer (memory.hpp:220-249). */
inline void memory_write(mem_handle& dst, vcl_size_t offset, vcl_size_t bytes, const void* ptr, bool /*async*/ = false)
{
  if (bytes == 0) return;
  if (dst.get_active_handle_id() == HIP_MEMORY)
    hip::check(vclhip_memcpy_h2d(static_cast<char*>(dst.hip_handle()) + offset, ptr, bytes, hip::current_stream()),
               "memory_write");
  else if (dst.get_active_handle_id() == MAIN_MEMORY)
    std::memcpy(dst.ram_handle() + offset, ptr, bytes);
  else
    throw memory_exception("memory_write to an uninitialized handle");
}

/** Buffer -> host (memory.hpp:261-292). */
inline void memory_read(mem_handle const& src, vcl_size_t offset, vcl_size_t bytes, void* ptr, bool /*async*/ = false)
{
  if (bytes == 0) return;
  if (src.get_active_handle_id() == HIP_MEMORY)
    hip::check(vclhip_memcpy_d2h(ptr, static_cast<const char*>(src.hip_handle()) + offset, bytes, hip::current_stream()),
               "memory_read");
  else if (src.get_active_handle_id() == MAIN_MEMORY)
    std::memcpy(ptr, src.ram_handle() + offset, bytes);
  else
    throw memory_exception("memory_read from an uninitialized handle");
}

/** Moves a buffer to another domain (memory.hpp:367-464). */
inline void switch_memory_context(mem_handle& handle, viennacl::context new_ctx)
{
  if (handle.get_active_handle_id() == new_ctx.memory_type()) return;
  if (handle.get_active_handle_id() == MEMORY_NOT_INITIALIZED || handle.raw_size() == 0) {
    handle.switch_active_handle_id(new_ctx.memory_type());
    return;
  }
  mem_handle moved;
  if (new_ctx.memory_type() == MAIN_MEMORY) {
    moved.reset_ram(handle.raw_size());
    memory_read(handle, 0, handle.raw_size(), moved.ram_handle());
  } else {
    memory_create(moved, handle.raw_size(), new_ctx, handle.ram_handle());
  }
  handle.swap(moved);
}

}  // namespace backend
}  // namespace viennacl

#endif
