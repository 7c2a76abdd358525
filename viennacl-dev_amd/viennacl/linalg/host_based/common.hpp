#ifndef VIENNACL_LINALG_HOST_BASED_COMMON_HPP
#define VIENNACL_LINALG_HOST_BASED_COMMON_HPP
// common.hpp -- extract_raw_pointer (reference linalg/host_based/common.hpp:36-46): the host array of
// a MAIN_MEMORY buffer, which user code such as a matrix-free operator's host worker
// (examples/tutorial/matrix-free.cpp:156-160) reads and writes directly.  Here MAIN_MEMORY buffers
// are host arrays and every computation of the library runs in HIP_MEMORY, so the pointer of a
// HIP_MEMORY buffer is not a host array: asking for it throws.

#include "viennacl/backend/mem_handle.hpp"

namespace viennacl
{
namespace linalg
{
namespace host_based
{
namespace detail
{
inline char* main_memory_array(viennacl::backend::mem_handle const& h)
{
  if (h.get_active_handle_id() != viennacl::MAIN_MEMORY)
    throw viennacl::memory_exception("extract_raw_pointer: the buffer is not in MAIN_MEMORY");
  return h.ram_handle();
}

template <typename ResultT>
ResultT* extract_raw_pointer(viennacl::backend::mem_handle& h)
{
  return reinterpret_cast<ResultT*>(main_memory_array(h));
}

template <typename ResultT>
ResultT const* extract_raw_pointer(viennacl::backend::mem_handle const& h)
{
  return reinterpret_cast<ResultT const*>(main_memory_array(h));
}

template <typename ResultT, typename VectorT>
ResultT* extract_raw_pointer(VectorT& vec)
{
  return reinterpret_cast<ResultT*>(main_memory_array(vec.handle()));
}

template <typename ResultT, typename VectorT>
ResultT const* extract_raw_pointer(VectorT const& vec)
{
  return reinterpret_cast<ResultT const*>(main_memory_array(vec.handle()));
}
}  // namespace detail
}  // namespace host_based
}  // namespace linalg
}  // namespace viennacl

#endif
