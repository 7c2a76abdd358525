#ifndef VIENNACL_RANGE_HPP
#define VIENNACL_RANGE_HPP
// range.hpp -- the index range [start, stop) of a vector_range (reference range.hpp:39-74).

#include <cassert>

#include "viennacl/forwards.h"

namespace viennacl
{

template <typename SizeT = vcl_size_t, typename DistanceT = vcl_ptrdiff_t>
class basic_range
{
public:
  typedef SizeT size_type;
  typedef DistanceT difference_type;
  typedef size_type value_type;
  typedef value_type const_reference;
  typedef const_reference reference;

  basic_range() : start_(0), size_(0) {}
  basic_range(size_type start_index, size_type stop_index) : start_(start_index), size_(stop_index - start_index)
  {
    assert(start_index <= stop_index);
  }

  size_type start() const { return start_; }
  size_type size() const { return size_; }
  const_reference operator()(size_type i) const { return start_ + i; }
  const_reference operator[](size_type i) const { return start_ + i; }
  bool operator==(basic_range const& o) const { return start_ == o.start_ && size_ == o.size_; }
  bool operator!=(basic_range const& o) const { return !(*this == o); }

private:
  size_type start_;
  size_type size_;
};

typedef basic_range<> range;

}  // namespace viennacl

#endif
