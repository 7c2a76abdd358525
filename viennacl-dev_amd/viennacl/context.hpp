#ifndef VIENNACL_CONTEXT_HPP
#define VIENNACL_CONTEXT_HPP
// context.hpp -- the memory domain an object lives in (reference context.hpp:39-83).

#include "viennacl/backend/mem_handle.hpp"

namespace viennacl
{

class context
{
public:
  context() : mem_type_(backend::default_memory_type()) {}
  explicit context(memory_types mtype) : mem_type_(mtype == MEMORY_NOT_INITIALIZED ? backend::default_memory_type() : mtype) {}
  memory_types memory_type() const { return mem_type_; }
  bool operator==(context const& o) const { return mem_type_ == o.mem_type_; }
  bool operator!=(context const& o) const { return !(*this == o); }

private:
  memory_types mem_type_;
};

namespace traits
{
inline context context_of(backend::mem_handle const& h)
{
  return context(h.get_active_handle_id());
}
}  // namespace traits

}  // namespace viennacl

#endif
