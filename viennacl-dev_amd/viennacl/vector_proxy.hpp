#ifndef VIENNACL_VECTOR_PROXY_HPP
#define VIENNACL_VECTOR_PROXY_HPP
// vector_proxy.hpp -- vector_range and vector_slice: views on part of a vector that share its
// device buffer (reference vector_proxy.hpp:38-338).  A view is a vector_base with its own start
// and stride over the parent's buffer, so every vector operation, prod() and the copies take it
// as they take a vector; assigning to a view writes into the parent.

#include "viennacl/range.hpp"
#include "viennacl/slice.hpp"
#include "viennacl/vector.hpp"

namespace viennacl
{

/** The entries [r.start(), r.start() + r.size()) of a vector (reference vector_proxy.hpp:38-79). */
template <typename VectorType>
class vector_range : public vector_base<typename VectorType::value_type>
{
  typedef vector_base<typename VectorType::value_type> base_type;

public:
  typedef typename VectorType::value_type value_type;
  typedef range::size_type size_type;
  typedef range::difference_type difference_type;
  typedef backend::mem_handle handle_type;
  typedef value_type cpu_value_type;

  vector_range(VectorType const& v, range const& r)
    : base_type(const_cast<handle_type&>(v.handle()), r.size(), v.start() + v.stride() * r.start(), v.stride()) {}
  vector_range(vector_range const& v, range const& r)
    : base_type(const_cast<handle_type&>(v.handle()), r.size(), v.start() + v.stride() * r.start(), v.stride()) {}
  /** Copying a view gives another view of the same entries (vector_proxy.hpp:62-64). */
  vector_range(vector_range const& other)
    : base_type(const_cast<handle_type&>(other.handle()), other.size(), other.start(), other.stride()) {}

  using base_type::operator=;
  vector_range& operator=(vector_range const& o)
  {
    base_type::operator=(static_cast<base_type const&>(o));
    return *this;
  }
};

/** The entries s.start() + i * s.stride(), i < s.size(), of a vector (reference vector_proxy.hpp:195-258). */
template <typename VectorType>
class vector_slice : public vector_base<typename VectorType::value_type>
{
  typedef vector_base<typename VectorType::value_type> base_type;

public:
  typedef typename VectorType::value_type value_type;
  typedef slice::size_type size_type;
  typedef slice::difference_type difference_type;
  typedef backend::mem_handle handle_type;
  typedef value_type cpu_value_type;

  vector_slice(VectorType const& v, slice const& s)
    : base_type(const_cast<handle_type&>(v.handle()), s.size(), v.start() + v.stride() * s.start(), v.stride() * s.stride()) {}
  vector_slice(vector_slice const& v, slice const& s)
    : base_type(const_cast<handle_type&>(v.handle()), s.size(), v.start() + v.stride() * s.start(), v.stride() * s.stride()) {}
  /** A range of a slice is a slice (vector_proxy.hpp:323-328). */
  vector_slice(vector_slice const& v, range const& r)
    : base_type(const_cast<handle_type&>(v.handle()), r.size(), v.start() + v.stride() * r.start(), v.stride()) {}
  /** A slice of a range is a slice (vector_proxy.hpp:330-335). */
  vector_slice(vector_range<VectorType> const& v, slice const& s)
    : base_type(const_cast<handle_type&>(v.handle()), s.size(), v.start() + v.stride() * s.start(), v.stride() * s.stride()) {}
  vector_slice(vector_slice const& other)
    : base_type(const_cast<handle_type&>(other.handle()), other.size(), other.start(), other.stride()) {}

  using base_type::operator=;
  vector_slice& operator=(vector_slice const& o)
  {
    base_type::operator=(static_cast<base_type const&>(o));
    return *this;
  }
};

// ---- project() (vector_proxy.hpp:168-180, :307-336)
template <typename VectorType>
vector_range<VectorType> project(VectorType const& vec, range const& r)
{
  return vector_range<VectorType>(vec, r);
}
template <typename VectorType>
vector_range<VectorType> project(vector_range<VectorType> const& vec, range const& r)
{
  assert(r.start() + r.size() <= vec.size() && bool("range exceeds the vector range"));
  return vector_range<VectorType>(vec, r);
}
template <typename VectorType>
vector_slice<VectorType> project(VectorType const& vec, slice const& s)
{
  return vector_slice<VectorType>(vec, s);
}
template <typename VectorType>
vector_slice<VectorType> project(vector_slice<VectorType> const& vec, slice const& s)
{
  return vector_slice<VectorType>(vec, s);
}
template <typename VectorType>
vector_slice<VectorType> project(vector_slice<VectorType> const& vec, range const& r)
{
  return vector_slice<VectorType>(vec, r);
}
template <typename VectorType>
vector_slice<VectorType> project(vector_range<VectorType> const& vec, slice const& s)
{
  return vector_slice<VectorType>(vec, s);
}

}  // namespace viennacl

#endif
