#ifndef VIENNACL_SLICE_HPP
#define VIENNACL_SLICE_HPP
// slice.hpp -- the strided index set {start + i * stride, i < size} of a vector_slice (reference
// slice.hpp:39-75).

#include "viennacl/forwards.h"

namespace viennacl
{

template <typename SizeT = vcl_size_t, typename DistanceT = vcl_ptrdiff_t>
class basic_slice
{
public:
  typedef SizeT size_type;
  typedef DistanceT difference_type;
  typedef size_type value_type;
  typedef value_type const_reference;
  typedef const_reference reference;

  basic_slice() : start_(0), stride_(1), size_(0) {}
  basic_slice(size_type start_index, size_type stride_arg, size_type size_arg)
    : start_(start_index), stride_(stride_arg), size_(size_arg) {}

  size_type start() const { return start_; }
  size_type stride() const { return stride_; }
  size_type size() const { return size_; }
  const_reference operator()(size_type i) const { return start_ + i * stride_; }
  const_reference operator[](size_type i) const { return start_ + i * stride_; }
  bool operator==(basic_slice const& o) const { return start_ == o.start_ && stride_ == o.stride_ && size_ == o.size_; }
  bool operator!=(basic_slice const& o) const { return !(*this == o); }

private:
  size_type start_;
  size_type stride_;
  size_type size_;
};

typedef basic_slice<> slice;

}  // namespace viennacl

#endif
