#ifndef VIENNACL_BACKEND_MEM_HANDLE_HPP
#define VIENNACL_BACKEND_MEM_HANDLE_HPP
// mem_handle.hpp -- a buffer in one memory domain (reference backend/mem_handle.hpp:89-245):
// a reference-counted host buffer (MAIN_MEMORY) or device buffer (HIP_MEMORY).  A device buffer
// can also wrap memory owned by the caller (the `.inc()` idiom of compressed_matrix.hpp:751-758
// and vector.hpp:271): the handle then never frees it.

#include <cstring>
#include <memory>

#include "viennacl/forwards.h"

namespace viennacl
{
namespace backend
{
namespace hip
{
/** Stream used by the HIP domain on the calling thread (null stream by default). */
inline void*& current_stream()
{
  static thread_local void* s = nullptr;
  return s;
}
inline void set_stream(void* s) { current_stream() = s; }
}  // namespace hip

/** Default domain of new objects: HIP_MEMORY (the reference picks CUDA > OpenCL > host, mem_handle.hpp:51-66). */
inline memory_types& default_memory_type_ref()
{
  static memory_types t = HIP_MEMORY;
  return t;
}
inline memory_types default_memory_type() { return default_memory_type_ref(); }
inline void default_memory_type(memory_types t) { default_memory_type_ref() = t; }

class mem_handle
{
public:
  mem_handle() : active_(MEMORY_NOT_INITIALIZED), size_(0) {}

  memory_types get_active_handle_id() const { return active_; }
  void switch_active_handle_id(memory_types t)
  {
    if (t != active_) {
      if (active_ == MEMORY_NOT_INITIALIZED || size_ == 0)
        active_ = t;
      else
        throw memory_exception("switching a live buffer between domains needs switch_memory_context()");
    }
  }

  vcl_size_t raw_size() const { return size_; }

  char* ram_handle() const { return ram_.get(); }
  void* hip_handle() const { return hip_.get(); }

  template <typename T>
  T* hip_ptr() const { return static_cast<T*>(hip_.get()); }

  /** Takes ownership of a fresh device allocation. */
  void reset_hip(void* p, vcl_size_t bytes)
  {
    hip_ = std::shared_ptr<void>(p, [](void* q) { (void)vclhip_free(q); });
    ram_.reset();
    size_ = bytes;
    active_ = HIP_MEMORY;
  }
  /** Wraps device memory owned by the caller; it is never freed here. */
  void wrap_hip(void* p, vcl_size_t bytes)
  {
    hip_ = std::shared_ptr<void>(p, [](void*) {});
    ram_.reset();
    size_ = bytes;
    active_ = HIP_MEMORY;
  }
  void reset_ram(vcl_size_t bytes)
  {
    ram_ = std::shared_ptr<char>(new char[bytes > 0 ? bytes : 1], std::default_delete<char[]>());
    hip_.reset();
    size_ = bytes;
    active_ = MAIN_MEMORY;
  }

  bool operator==(mem_handle const& o) const
  {
    if (active_ != o.active_) return false;
    if (active_ == HIP_MEMORY) return hip_.get() == o.hip_.get();
    if (active_ == MAIN_MEMORY) return ram_.get() == o.ram_.get();
    return true;
  }
  bool operator!=(mem_handle const& o) const { return !(*this == o); }

  void swap(mem_handle& o)
  {
    std::swap(active_, o.active_);
    std::swap(size_, o.size_);
    hip_.swap(o.hip_);
    ram_.swap(o.ram_);
  }

private:
  memory_types active_;
  vcl_size_t size_;
  std::shared_ptr<void> hip_;
  std::shared_ptr<char> ram_;
};

}  // namespace backend
}  // namespace viennacl

#endif
