#ifndef VIENNACL_VECTOR_HPP
#define VIENNACL_VECTOR_HPP
// vector.hpp -- dense vectors of the drop-in API (reference detail/vector_def.hpp:104-349 and
// vector.hpp:251-3320): vector_base / vector with start, stride and 128-padding, host <-> device
// copies, and the expressions the Krylov path uses (x = a*y + b*z, x += a*y, x = prod(A, y),
// element_div, inner_prod, norm_2).  Every operation on HIP_MEMORY runs through libvclhip.

#include <algorithm>
#include <cmath>
#include <iterator>
#include <memory>
#include <type_traits>
#include <vector>

#include "viennacl/backend/memory.hpp"
#include "viennacl/context.hpp"
#include "viennacl/forwards.h"
#include "viennacl/linalg/host_based/common.hpp"

namespace viennacl
{

// ------------------------------------------------------------------ expression tags
struct op_add {};
struct op_sub {};
struct op_mult {};
struct op_div {};
struct op_prod {};
struct op_element_div {};

/** A lazily evaluated vector expression (reference forwards.h vector_expression). */
template <typename LHS, typename RHS, typename OP>
class vector_expression
{
  typedef typename std::conditional<std::is_arithmetic<LHS>::value, LHS, LHS const&>::type lhs_store;
  typedef typename std::conditional<std::is_arithmetic<RHS>::value, RHS, RHS const&>::type rhs_store;

public:
  vector_expression(LHS const& l, RHS const& r) : lhs_(l), rhs_(r) {}
  lhs_store lhs() const { return lhs_; }
  rhs_store rhs() const { return rhs_; }

private:
  lhs_store lhs_;
  rhs_store rhs_;
};

namespace detail
{
template <typename T> struct is_vector_like : std::false_type {};
template <typename T> struct is_vector_like<vector_base<T> > : std::true_type {};
template <typename T, unsigned int A> struct is_vector_like<vector<T, A> > : std::true_type {};
template <typename L, typename R, typename O> struct is_vector_like<vector_expression<L, R, O> > : std::true_type {};
template <typename V> struct is_vector_like<vector_range<V> > : std::true_type {};
template <typename V> struct is_vector_like<vector_slice<V> > : std::true_type {};

inline void* stream() { return backend::hip::current_stream(); }

template <typename T> struct abi;
template <> struct abi<double>
{
  static int fill(uint32_t n, double* v, uint32_t s, uint32_t i, double a) { return vclhip_fill_f64(n, v, s, i, a, stream()); }
  static int dot(uint32_t n, const double* x, uint32_t xs, uint32_t xi, const double* y, uint32_t ys, uint32_t yi,
                 double* r) { return vclhip_dot_f64(n, x, xs, xi, y, ys, yi, r, stream()); }
  static int norm2(uint32_t n, const double* x, uint32_t s, uint32_t i, double* r) { return vclhip_norm2_f64(n, x, s, i, r, stream()); }
  static int ediv(uint32_t n, const double* x, const double* y, double* z) { return vclhip_element_div_f64(n, x, y, z, stream()); }
  static int asum(uint32_t n, const double* x, uint32_t s, uint32_t i, double* r) { return vclhip_asum_f64(n, x, s, i, r, stream()); }
  static int iamax(uint32_t n, const double* x, uint32_t s, uint32_t i, uint32_t* r) { return vclhip_iamax_f64(n, x, s, i, r, stream()); }
  static int swap(uint32_t n, double* x, uint32_t xs, uint32_t xi, double* y, uint32_t ys, uint32_t yi)
  { return vclhip_swap_f64(n, x, xs, xi, y, ys, yi, stream()); }
  static int rot(uint32_t n, double* x, uint32_t xs, uint32_t xi, double* y, uint32_t ys, uint32_t yi, double c, double d)
  { return vclhip_rot_f64(n, x, xs, xi, y, ys, yi, c, d, stream()); }
};
template <> struct abi<float>
{
  static int fill(uint32_t n, float* v, uint32_t s, uint32_t i, float a) { return vclhip_fill_f32(n, v, s, i, a, stream()); }
  static int dot(uint32_t n, const float* x, uint32_t xs, uint32_t xi, const float* y, uint32_t ys, uint32_t yi,
                 float* r) { return vclhip_dot_f32(n, x, xs, xi, y, ys, yi, r, stream()); }
  static int norm2(uint32_t n, const float* x, uint32_t s, uint32_t i, float* r) { return vclhip_norm2_f32(n, x, s, i, r, stream()); }
  static int ediv(uint32_t n, const float* x, const float* y, float* z) { return vclhip_element_div_f32(n, x, y, z, stream()); }
  static int asum(uint32_t n, const float* x, uint32_t s, uint32_t i, float* r) { return vclhip_asum_f32(n, x, s, i, r, stream()); }
  static int iamax(uint32_t n, const float* x, uint32_t s, uint32_t i, uint32_t* r) { return vclhip_iamax_f32(n, x, s, i, r, stream()); }
  static int swap(uint32_t n, float* x, uint32_t xs, uint32_t xi, float* y, uint32_t ys, uint32_t yi)
  { return vclhip_swap_f32(n, x, xs, xi, y, ys, yi, stream()); }
  static int rot(uint32_t n, float* x, uint32_t xs, uint32_t xi, float* y, uint32_t ys, uint32_t yi, float c, float d)
  { return vclhip_rot_f32(n, x, xs, xi, y, ys, yi, c, d, stream()); }
};
}  // namespace detail

// ------------------------------------------------------------------ entry access and iterators

/** Read/write access to one entry of a device buffer through host round trips (reference
 *  entry_proxy, scalar.hpp:41-196): `offset` counts elements from the start of the buffer. */
template <typename NumericT>
class entry_proxy
{
public:
  entry_proxy(vcl_size_t offset, backend::mem_handle& h) : offset_(offset), h_(&h) {}
  operator NumericT() const
  {
    NumericT v;
    backend::memory_read(*h_, sizeof(NumericT) * offset_, sizeof(NumericT), &v);
    return v;
  }
  entry_proxy& operator=(NumericT v)
  {
    backend::memory_write(*h_, sizeof(NumericT) * offset_, sizeof(NumericT), &v);
    return *this;
  }
  entry_proxy& operator=(entry_proxy const& o) { return *this = NumericT(o); }
  entry_proxy& operator+=(NumericT v) { return *this = NumericT(*this) + v; }
  entry_proxy& operator-=(NumericT v) { return *this = NumericT(*this) - v; }
  entry_proxy& operator*=(NumericT v) { return *this = NumericT(*this) * v; }
  entry_proxy& operator/=(NumericT v) { return *this = NumericT(*this) / v; }
  vcl_size_t index() const { return offset_; }
  backend::mem_handle const& handle() const { return *h_; }

private:
  vcl_size_t offset_;
  backend::mem_handle* h_;
};

/** Read-only iterator over the entries of a vector or vector proxy (reference
 *  const_vector_iterator, vector.hpp:108-201): entry i sits at start + i * stride of the buffer;
 *  dereferencing reads it back to the host. */
template <class NumericT, unsigned int AlignmentV>
class const_vector_iterator
{
  typedef const_vector_iterator<NumericT, AlignmentV> self_type;

public:
  typedef NumericT value_type;
  typedef vcl_size_t size_type;
  typedef vcl_ptrdiff_t difference_type;
  typedef backend::mem_handle handle_type;

  const_vector_iterator(vector_base<NumericT> const& vec, size_type index, size_type start = 0, size_type stride = 1)
    : elements_(&vec.handle()), index_(index), start_(start), stride_(stride) {}
  const_vector_iterator(handle_type const& elements, size_type index, size_type start = 0, size_type stride = 1)
    : elements_(&elements), index_(index), start_(start), stride_(stride) {}

  NumericT operator*() const
  {
    NumericT v;
    backend::memory_read(*elements_, sizeof(NumericT) * offset(), sizeof(NumericT), &v);
    return v;
  }
  self_type& operator++() { ++index_; return *this; }
  self_type operator++(int) { self_type t = *this; ++index_; return t; }
  bool operator==(self_type const& o) const { return index_ == o.index_; }
  bool operator!=(self_type const& o) const { return index_ != o.index_; }
  difference_type operator-(self_type const& o) const { return difference_type(index_) - difference_type(o.index_); }
  self_type operator+(difference_type d) const { return self_type(*elements_, size_type(difference_type(index_) + d), start_, stride_); }

  size_type index() const { return index_; }
  size_type offset() const { return start_ + index_ * stride_; }
  size_type stride() const { return stride_; }
  handle_type const& handle() const { return *elements_; }

protected:
  handle_type const* elements_;
  size_type index_;
  size_type start_;
  size_type stride_;
};

/** Writable iterator (reference vector_iterator, vector.hpp:205-249): dereferencing gives an
 *  entry_proxy. */
template <class NumericT, unsigned int AlignmentV>
class vector_iterator : public const_vector_iterator<NumericT, AlignmentV>
{
  typedef const_vector_iterator<NumericT, AlignmentV> base_type;
  typedef vector_iterator<NumericT, AlignmentV> self_type;

public:
  typedef typename base_type::handle_type handle_type;
  typedef typename base_type::size_type size_type;
  typedef typename base_type::difference_type difference_type;

  vector_iterator(handle_type& elements, size_type index, size_type start = 0, size_type stride = 1)
    : base_type(elements, index, start, stride) {}
  vector_iterator(vector_base<NumericT>& vec, size_type index, size_type start = 0, size_type stride = 1)
    : base_type(vec.handle(), index, start, stride) {}

  entry_proxy<NumericT> operator*() const { return entry_proxy<NumericT>(this->offset(), handle()); }
  self_type& operator++() { ++this->index_; return *this; }
  self_type operator++(int) { self_type t = *this; ++this->index_; return t; }
  difference_type operator-(self_type const& o) const { return difference_type(this->index_) - difference_type(o.index_); }
  self_type operator+(difference_type d) const
  {
    return self_type(handle(), size_type(difference_type(this->index_) + d), this->start_, this->stride_);
  }
  handle_type& handle() const { return const_cast<handle_type&>(*this->elements_); }
};

// ------------------------------------------------------------------ vector_base

template <typename NumericT>
class vector_base
{
  static_assert(std::is_same<NumericT, double>::value || std::is_same<NumericT, float>::value,
                "the HIP backend provides double and float vectors");

public:
  typedef NumericT value_type;
  typedef NumericT cpu_value_type;
  typedef vcl_size_t size_type;
  typedef vcl_ptrdiff_t difference_type;

  typedef const_vector_iterator<NumericT, 1> const_iterator;
  typedef vector_iterator<NumericT, 1> iterator;

  vector_base() : size_(0), start_(0), stride_(1), internal_size_(0) {}

  explicit vector_base(size_type n, viennacl::context ctx = viennacl::context())
    : size_(n), start_(0), stride_(1), internal_size_(tools::align_to_multiple<size_type>(n, dense_padding_size))
  {
    if (n > 0) backend::memory_create(elements_, sizeof(NumericT) * internal_size_, ctx);
  }

  /** Wraps caller-owned device memory (vector.hpp:271): the handle never frees it. */
  vector_base(NumericT* ptr_to_mem, memory_types mem_type, size_type n, size_type start = 0, size_type stride = 1)
    : size_(n), start_(start), stride_(stride), internal_size_(n)
  {
    if (mem_type != HIP_MEMORY) throw memory_exception("only HIP_MEMORY buffers can be wrapped");
    elements_.wrap_hip(ptr_to_mem, sizeof(NumericT) * (start + (n ? (n - 1) * stride + 1 : 0)));
  }

  /** A view on an existing handle (used e.g. on a matrix's value array, mixed_precision_cg.hpp:134-135). */
  vector_base(backend::mem_handle& h, size_type n, size_type start, size_type stride)
    : size_(n), start_(start), stride_(stride), internal_size_(n), elements_(h) {}

  vector_base(vector_base const& other)
    : size_(other.size_), start_(0), stride_(1), internal_size_(tools::align_to_multiple<size_type>(other.size_, dense_padding_size))
  {
    if (size_ > 0) {
      backend::memory_create(elements_, sizeof(NumericT) * internal_size_, viennacl::traits::context_of(other.elements_));
      assign_linear(other, NumericT(1));
    }
  }

  template <typename L, typename R, typename O>
  vector_base(vector_expression<L, R, O> const& e) : size_(0), start_(0), stride_(1), internal_size_(0)
  {
    *this = e;
  }

  vector_base& operator=(vector_base const& other)
  {
    if (this == &other) return *this;
    if (size_ == 0) init(other.size_, viennacl::traits::context_of(other.elements_));
    check_size(other.size_);
    if (size_) assign_linear(other, NumericT(1));
    return *this;
  }

  template <typename L, typename R, typename O>
  vector_base& operator=(vector_expression<L, R, O> const& e);
  template <typename L, typename R, typename O>
  vector_base& operator+=(vector_expression<L, R, O> const& e);
  template <typename L, typename R, typename O>
  vector_base& operator-=(vector_expression<L, R, O> const& e);
  vector_base& operator+=(vector_base const& v);
  vector_base& operator-=(vector_base const& v);
  vector_base& operator*=(NumericT a);
  vector_base& operator/=(NumericT a);

  size_type size() const { return size_; }
  size_type internal_size() const { return internal_size_; }
  size_type start() const { return start_; }
  size_type stride() const { return stride_; }
  bool empty() const { return size_ == 0; }
  bool contiguous() const { return start_ == 0 && stride_ == 1; }

  backend::mem_handle& handle() { return elements_; }
  backend::mem_handle const& handle() const { return elements_; }
  NumericT* device_ptr() const { return elements_.template hip_ptr<NumericT>(); }

  const_iterator begin() const { return const_iterator(elements_, 0, start_, stride_); }
  const_iterator end() const { return const_iterator(elements_, size_, start_, stride_); }
  iterator begin() { return iterator(elements_, 0, start_, stride_); }
  iterator end() { return iterator(elements_, size_, start_, stride_); }

  entry_proxy<NumericT> operator[](size_type i) { return entry_proxy<NumericT>(start_ + i * stride_, elements_); }
  entry_proxy<NumericT> operator()(size_type i) { return entry_proxy<NumericT>(start_ + i * stride_, elements_); }
  NumericT operator[](size_type i) const { return get_entry(i); }
  NumericT operator()(size_type i) const { return get_entry(i); }

  /** Sets all entries (padding included) to zero (vector.hpp clear()). */
  void clear()
  {
    if (size_ == 0) return;
    require_hip();
    if (contiguous())
      backend::hip::check(vclhip_fill_zero(device_ptr(), sizeof(NumericT) * internal_size_, detail::stream()), "clear");
    else
      backend::hip::check(detail::abi<NumericT>::fill(uint32_t(size_), device_ptr(), uint32_t(start_), uint32_t(stride_), NumericT(0)), "clear");
  }

  void resize(size_type n, bool preserve = true)
  {
    if (n == size_) return;
    if (!contiguous()) throw memory_exception("resize of a strided view");
    vector_base tmp(n, viennacl::traits::context_of(elements_));
    if (preserve && size_ > 0) {
      size_type m = n < size_ ? n : size_;
      backend::memory_copy(elements_, tmp.elements_, 0, 0, sizeof(NumericT) * m);
    }
    swap(tmp);
  }

  void swap(vector_base& o)
  {
    std::swap(size_, o.size_);
    std::swap(start_, o.start_);
    std::swap(stride_, o.stride_);
    std::swap(internal_size_, o.internal_size_);
    elements_.swap(o.elements_);
  }

  void switch_memory_context(viennacl::context ctx) { backend::switch_memory_context(elements_, ctx); }

  // -- helpers used by the operation layer
  void init(size_type n, viennacl::context ctx)
  {
    size_ = n;
    start_ = 0;
    stride_ = 1;
    internal_size_ = tools::align_to_multiple<size_type>(n, dense_padding_size);
    if (n > 0) backend::memory_create(elements_, sizeof(NumericT) * internal_size_, ctx);
  }
  void check_size(size_type n) const
  {
    if (n != size_) throw std::invalid_argument("ViennaCL (HIP): vector size mismatch");
  }
  void require_hip() const
  {
    if (elements_.get_active_handle_id() != HIP_MEMORY)
      throw memory_exception("this build computes in HIP_MEMORY only; call switch_memory_context(HIP_MEMORY)");
  }
  /** this = v * a (elementwise, strided). */
  void assign_linear(vector_base const& v, NumericT a);

  NumericT get_entry(size_type i) const
  {
    NumericT val;
    backend::memory_read(elements_, sizeof(NumericT) * (start_ + i * stride_), sizeof(NumericT), &val);
    return val;
  }
  void set_entry(size_type i, NumericT val)
  {
    backend::memory_write(elements_, sizeof(NumericT) * (start_ + i * stride_), sizeof(NumericT), &val);
  }

private:
  size_type size_;
  size_type start_;
  size_type stride_;
  size_type internal_size_;
  backend::mem_handle elements_;
};

template <typename NumericT, unsigned int AlignmentV>
class vector : public vector_base<NumericT>
{
  typedef vector_base<NumericT> base_type;

public:
  typedef typename base_type::size_type size_type;
  vector() : base_type() {}
  explicit vector(size_type n, viennacl::context ctx = viennacl::context()) : base_type(n, ctx) {}
  vector(NumericT* ptr, memory_types mem_type, size_type n, size_type start = 0, size_type stride = 1)
    : base_type(ptr, mem_type, n, start, stride) {}
  vector(vector const& o) : base_type(o) {}
  vector(base_type const& o) : base_type(o) {}
  template <typename L, typename R, typename O>
  vector(vector_expression<L, R, O> const& e) : base_type()
  {
    base_type::operator=(e);
  }
  vector& operator=(vector const& o) { base_type::operator=(o); return *this; }
  vector& operator=(base_type const& o) { base_type::operator=(o); return *this; }
  template <typename L, typename R, typename O>
  vector& operator=(vector_expression<L, R, O> const& e) { base_type::operator=(e); return *this; }
};

/** Vector of zeros (reference zero_vector). */
template <typename NumericT>
vector<NumericT> zero_vector(vcl_size_t n, viennacl::context ctx = viennacl::context())
{
  return vector<NumericT>(n, ctx);  // new vectors are zero-filled
}

/** Vector with every entry equal to `val` (reference scalar_vector). */
template <typename NumericT>
vector<NumericT> scalar_vector(vcl_size_t n, NumericT val, viennacl::context ctx = viennacl::context())
{
  vector<NumericT> v(n, ctx);
  if (n) backend::hip::check(detail::abi<NumericT>::fill(uint32_t(n), v.device_ptr(), 0, 1, val), "scalar_vector");
  return v;
}

// ------------------------------------------------------------------ expression builders
template <typename T, typename S>
typename std::enable_if<std::is_arithmetic<S>::value, vector_expression<vector_base<T>, T, op_mult> >::type
operator*(S a, vector_base<T> const& v)
{
  return vector_expression<vector_base<T>, T, op_mult>(v, T(a));
}
template <typename T, typename S>
typename std::enable_if<std::is_arithmetic<S>::value, vector_expression<vector_base<T>, T, op_mult> >::type
operator*(vector_base<T> const& v, S a)
{
  return vector_expression<vector_base<T>, T, op_mult>(v, T(a));
}
template <typename T, typename S>
typename std::enable_if<std::is_arithmetic<S>::value, vector_expression<vector_base<T>, T, op_div> >::type
operator/(vector_base<T> const& v, S a)
{
  return vector_expression<vector_base<T>, T, op_div>(v, T(a));
}
template <typename L, typename R, typename O, typename S>
typename std::enable_if<std::is_arithmetic<S>::value, vector_expression<vector_expression<L, R, O>, S, op_mult> >::type
operator*(S a, vector_expression<L, R, O> const& e)
{
  return vector_expression<vector_expression<L, R, O>, S, op_mult>(e, a);
}

template <typename A, typename B>
typename std::enable_if<detail::is_vector_like<A>::value && detail::is_vector_like<B>::value,
                        vector_expression<A, B, op_add> >::type
operator+(A const& a, B const& b)
{
  return vector_expression<A, B, op_add>(a, b);
}
template <typename A, typename B>
typename std::enable_if<detail::is_vector_like<A>::value && detail::is_vector_like<B>::value,
                        vector_expression<A, B, op_sub> >::type
operator-(A const& a, B const& b)
{
  return vector_expression<A, B, op_sub>(a, b);
}

namespace linalg
{
/** element_div(x, y) (linalg/vector_operations.hpp:254-285 with op_div). */
template <typename T>
vector_expression<vector_base<T>, vector_base<T>, op_element_div> element_div(vector_base<T> const& x,
                                                                              vector_base<T> const& y)
{
  return vector_expression<vector_base<T>, vector_base<T>, op_element_div>(x, y);
}
}  // namespace linalg

// ------------------------------------------------------------------ evaluation
namespace detail
{
/** A vector expression flattened into sum_k coef_k * v_k when it is linear. */
template <typename T>
struct linear_form
{
  std::vector<std::pair<T, vector_base<T> const*> > terms;
  std::vector<std::shared_ptr<vector_base<T> > > temps;  // keeps evaluated sub-expressions alive
};

template <typename T> void collect(linear_form<T>& f, vector_base<T> const& v, T c) { f.terms.push_back(std::make_pair(c, &v)); }
template <typename T, unsigned A> void collect(linear_form<T>& f, vector<T, A> const& v, T c) { f.terms.push_back(std::make_pair(c, static_cast<vector_base<T> const*>(&v))); }
template <typename T, typename L, typename R, typename O> void collect(linear_form<T>& f, vector_expression<L, R, O> const& e, T c);

template <typename T, typename L, typename S>
typename std::enable_if<std::is_arithmetic<S>::value>::type
collect(linear_form<T>& f, vector_expression<L, S, op_mult> const& e, T c) { collect(f, e.lhs(), T(c * T(e.rhs()))); }
template <typename T, typename L, typename S>
typename std::enable_if<std::is_arithmetic<S>::value>::type
collect(linear_form<T>& f, vector_expression<L, S, op_div> const& e, T c) { collect(f, e.lhs(), T(c / T(e.rhs()))); }
template <typename T, typename L, typename R>
void collect(linear_form<T>& f, vector_expression<L, R, op_add> const& e, T c) { collect(f, e.lhs(), c); collect(f, e.rhs(), c); }
template <typename T, typename L, typename R>
void collect(linear_form<T>& f, vector_expression<L, R, op_sub> const& e, T c) { collect(f, e.lhs(), c); collect(f, e.rhs(), T(-c)); }
/** Non-linear pieces (products, element-wise division) are evaluated into temporaries. */
template <typename T, typename L, typename R, typename O>
void collect(linear_form<T>& f, vector_expression<L, R, O> const& e, T c)
{
  std::shared_ptr<vector_base<T> > tmp(new vector_base<T>());
  vcl_eval(*tmp, e);  // found by argument-dependent lookup, including the sparse-product overloads
  f.temps.push_back(tmp);
  f.terms.push_back(std::make_pair(c, static_cast<vector_base<T> const*>(tmp.get())));
}

template <typename T>
int avbv(vector_base<T>& y, vector_base<T> const& a, T ca, vector_base<T> const& b, T cb);

template <>
inline int avbv<double>(vector_base<double>& y, vector_base<double> const& a, double ca, vector_base<double> const& b,
                        double cb)
{
  return vclhip_avbv_f64(uint32_t(y.size()), y.device_ptr(), uint32_t(y.start()), uint32_t(y.stride()), a.device_ptr(),
                         uint32_t(a.start()), uint32_t(a.stride()), ca, b.device_ptr(), uint32_t(b.start()),
                         uint32_t(b.stride()), cb, stream());
}

template <>
inline int avbv<float>(vector_base<float>& y, vector_base<float> const& a, float ca, vector_base<float> const& b, float cb)
{
  if (y.contiguous() && a.contiguous() && b.contiguous())
    return vclhip_axpby_f32(uint32_t(y.size()), ca, a.device_ptr(), cb, b.device_ptr(), y.device_ptr(), stream());
  return vclhip_avbv_f32(uint32_t(y.size()), y.device_ptr(), uint32_t(y.start()), uint32_t(y.stride()), a.device_ptr(),
                         uint32_t(a.start()), uint32_t(a.stride()), ca, b.device_ptr(), uint32_t(b.start()),
                         uint32_t(b.stride()), cb, stream());
}

/** y = sum of the terms; y may appear among them. */
template <typename T>
void assign_form(vector_base<T>& y, linear_form<T> const& f)
{
  y.require_hip();
  const size_t m = f.terms.size();
  for (size_t k = 0; k < m; ++k) f.terms[k].second->check_size(y.size());
  if (y.size() == 0) return;
  // a later term aliasing y would read an overwritten y: evaluate through a temporary then
  for (size_t k = 2; k < m; ++k)
    if (f.terms[k].second->handle() == y.handle()) {
      vector_base<T> tmp(y.size());
      assign_form(tmp, f);
      y.assign_linear(tmp, T(1));
      return;
    }
  if (m == 1) {
    y.assign_linear(*f.terms[0].second, f.terms[0].first);
    return;
  }
  backend::hip::check(avbv(y, *f.terms[0].second, f.terms[0].first, *f.terms[1].second, f.terms[1].first), "avbv");
  for (size_t k = 2; k < m; ++k)
    backend::hip::check(avbv(y, y, T(1), *f.terms[k].second, f.terms[k].first), "avbv");
}

template <typename T, typename L, typename R, typename O>
void evaluate_linear(vector_base<T>& y, vector_expression<L, R, O> const& e)
{
  linear_form<T> f;
  collect(f, e, T(1));
  if (y.size() == 0) y.init(f.terms.front().second->size(), viennacl::context(HIP_MEMORY));
  assign_form(y, f);
}

}  // namespace detail

// Evaluation hooks, called unqualified so that argument-dependent lookup also finds the overloads
// for sparse products declared with the matrix types (linalg/sparse_matrix_operations.hpp).

/** y = element_div(x, z) */
template <typename T>
void vcl_eval(vector_base<T>& y, vector_expression<vector_base<T>, vector_base<T>, op_element_div> const& e)
{
  if (y.size() == 0) y.init(e.lhs().size(), viennacl::context(HIP_MEMORY));
  y.check_size(e.lhs().size());
  e.rhs().check_size(e.lhs().size());
  if (!y.contiguous() || !e.lhs().contiguous() || !e.rhs().contiguous())
    throw memory_exception("element_div on strided vectors is not provided by the HIP backend");
  backend::hip::check(viennacl::detail::abi<T>::ediv(uint32_t(y.size()), e.lhs().device_ptr(), e.rhs().device_ptr(), y.device_ptr()),
                      "element_div");
}

/** y = linear combination of vectors */
template <typename T, typename L, typename R, typename O>
typename std::enable_if<!std::is_same<O, op_prod>::value && !std::is_same<O, op_element_div>::value>::type
vcl_eval(vector_base<T>& y, vector_expression<L, R, O> const& e)
{
  detail::evaluate_linear(y, e);
}

/** y += sign * e for any expression (sparse products have cheaper overloads) */
template <typename T, typename L, typename R, typename O>
typename std::enable_if<!std::is_same<O, op_prod>::value>::type
vcl_inplace(vector_base<T>& y, vector_expression<L, R, O> const& e, T sign)
{
  detail::linear_form<T> f;
  detail::collect(f, y, T(1));
  detail::collect(f, e, sign);
  detail::assign_form(y, f);
}

template <typename T>
void vector_base<T>::assign_linear(vector_base<T> const& v, T a)
{
  require_hip();
  v.check_size(size_);
  if (size_ == 0) return;
  if (v.handle() == handle() && v.start() == start() && v.stride() == stride() && a == T(1)) return;
  if (contiguous() && v.contiguous() && a == T(1)) {
    backend::memory_copy(v.handle(), handle(), 0, 0, sizeof(T) * size_);
    return;
  }
  backend::hip::check(detail::avbv(*this, v, a, v, T(0)), "assign");
}

template <typename T>
template <typename L, typename R, typename O>
vector_base<T>& vector_base<T>::operator=(vector_expression<L, R, O> const& e)
{
  vcl_eval(*this, e);
  return *this;
}

template <typename T>
template <typename L, typename R, typename O>
vector_base<T>& vector_base<T>::operator+=(vector_expression<L, R, O> const& e)
{
  vcl_inplace(*this, e, T(1));
  return *this;
}

template <typename T>
template <typename L, typename R, typename O>
vector_base<T>& vector_base<T>::operator-=(vector_expression<L, R, O> const& e)
{
  vcl_inplace(*this, e, T(-1));
  return *this;
}

template <typename T>
vector_base<T>& vector_base<T>::operator+=(vector_base<T> const& v)
{
  require_hip();
  v.check_size(size_);
  if (size_) backend::hip::check(detail::avbv(*this, *this, T(1), v, T(1)), "x += y");
  return *this;
}

template <typename T>
vector_base<T>& vector_base<T>::operator-=(vector_base<T> const& v)
{
  require_hip();
  v.check_size(size_);
  if (size_) backend::hip::check(detail::avbv(*this, *this, T(1), v, T(-1)), "x -= y");
  return *this;
}

template <typename T>
vector_base<T>& vector_base<T>::operator*=(T a)
{
  assign_linear(*this, a);
  return *this;
}

template <typename T>
vector_base<T>& vector_base<T>::operator/=(T a)
{
  assign_linear(*this, T(1) / a);
  return *this;
}

// ------------------------------------------------------------------ reductions
namespace linalg
{
/** inner_prod(x, y) (linalg/vector_operations.hpp:565-590), result on the host. */
template <typename T>
T inner_prod(vector_base<T> const& x, vector_base<T> const& y)
{
  x.require_hip();
  y.check_size(x.size());
  T r = 0;
  backend::hip::check(viennacl::detail::abi<T>::dot(uint32_t(x.size()), x.device_ptr(), uint32_t(x.start()), uint32_t(x.stride()),
                                          y.device_ptr(), uint32_t(y.start()), uint32_t(y.stride()), &r),
                      "inner_prod");
  return r;
}

/** norm_2(x) (linalg/vector_operations.hpp:812-840). */
template <typename T>
T norm_2(vector_base<T> const& x)
{
  x.require_hip();
  T r = 0;
  backend::hip::check(viennacl::detail::abi<T>::norm2(uint32_t(x.size()), x.device_ptr(), uint32_t(x.start()), uint32_t(x.stride()), &r),
                      "norm_2");
  return r;
}

template <typename L, typename R, typename O>
auto norm_2(vector_expression<L, R, O> const& e) -> decltype(norm_2(vector_base<double>()))
{
  vector_base<double> tmp;
  tmp = e;
  return norm_2(tmp);
}

/** norm_1(x) = sum |x_i| (linalg/norm_1.hpp:76-86; host_based/vector_operations.hpp:708-733). */
template <typename T>
T norm_1(vector_base<T> const& x)
{
  x.require_hip();
  T r = 0;
  backend::hip::check(viennacl::detail::abi<T>::asum(uint32_t(x.size()), x.device_ptr(), uint32_t(x.start()),
                                                     uint32_t(x.stride()), &r),
                      "norm_1");
  return r;
}

/** index_norm_inf(x): index of the first largest |x_i| (linalg/vector_operations.hpp:952-967;
 *  host_based/vector_operations.hpp:891-950, which returns the vector's start offset when no entry
 *  is above zero). */
template <typename T>
vcl_size_t index_norm_inf(vector_base<T> const& x)
{
  x.require_hip();
  uint32_t r = 0;
  backend::hip::check(viennacl::detail::abi<T>::iamax(uint32_t(x.size()), x.device_ptr(), uint32_t(x.start()),
                                                      uint32_t(x.stride()), &r),
                      "index_norm_inf");
  return r;
}

/** norm_inf(x) = max |x_i| (linalg/norm_inf.hpp:78-88); 0 for an empty or all-zero vector. */
template <typename T>
T norm_inf(vector_base<T> const& x)
{
  const vcl_size_t i = index_norm_inf(x);
  if (i >= x.size()) return T(0);
  return std::fabs(x.get_entry(i));
}

/** (x, y) <- (alpha x + beta y, alpha y - beta x) (plane_rotation, linalg/vector_operations.hpp:1259-1280;
 *  host_based/vector_operations.hpp:1080-1110). */
template <typename T>
void plane_rotation(vector_base<T>& x, vector_base<T>& y, T alpha, T beta)
{
  x.require_hip();
  y.check_size(x.size());
  backend::hip::check(viennacl::detail::abi<T>::rot(uint32_t(x.size()), x.device_ptr(), uint32_t(x.start()),
                                                    uint32_t(x.stride()), y.device_ptr(), uint32_t(y.start()),
                                                    uint32_t(y.stride()), alpha, beta),
                      "plane_rotation");
}
}  // namespace linalg

/** swap(x, y): exchanges the entries of two vectors of one size (vector.hpp:1648-1660, vector_swap). */
template <typename T>
void swap(vector_base<T>& x, vector_base<T>& y)
{
  x.require_hip();
  y.check_size(x.size());
  backend::hip::check(viennacl::detail::abi<T>::swap(uint32_t(x.size()), x.device_ptr(), uint32_t(x.start()),
                                                     uint32_t(x.stride()), y.device_ptr(), uint32_t(y.start()),
                                                     uint32_t(y.stride())),
                      "swap");
}

// ------------------------------------------------------------------ host <-> device copies
/** copy(host vector -> vector) (vector.hpp:1509-1527): any host container with begin()/end(); an empty
 *  target is sized, a view is written in place. */
template <typename HostVectorT, typename T>
typename std::enable_if<!detail::is_vector_like<HostVectorT>::value>::type
copy(HostVectorT const& cpu, vector_base<T>& gpu)
{
  const vcl_size_t n = vcl_size_t(std::distance(cpu.begin(), cpu.end()));
  if (gpu.size() == 0) gpu.init(n, viennacl::context(HIP_MEMORY));
  gpu.check_size(n);
  copy(cpu.begin(), cpu.end(), gpu.begin());
}

/** copy(vector -> host vector) (vector.hpp:1358-1366): the host container must hold size() entries;
 *  a std::vector is resized. */
template <typename T, typename CpuT>
void copy(vector_base<T> const& gpu, std::vector<CpuT>& cpu)
{
  cpu.resize(gpu.size());
  copy(gpu.begin(), gpu.end(), cpu.begin());
}

/** fast_copy(gpu_begin, gpu_end, cpu_begin) (vector.hpp:1233-1258): device range -> host memory that is
 *  one linear piece (&*cpu_begin onwards); a strided range is read with its gaps and unpacked. */
template <typename NumericT, unsigned int AlignmentV, typename CpuIt>
void fast_copy(const_vector_iterator<NumericT, AlignmentV> const& gpu_begin,
               const_vector_iterator<NumericT, AlignmentV> const& gpu_end, CpuIt cpu_begin)
{
  if (gpu_end - gpu_begin <= 0) return;
  const vcl_size_t count = vcl_size_t(gpu_end - gpu_begin);
  NumericT* out = &(*cpu_begin);
  if (gpu_begin.stride() == 1) {
    backend::memory_read(gpu_begin.handle(), sizeof(NumericT) * gpu_begin.offset(), sizeof(NumericT) * count, out);
    return;
  }
  std::vector<NumericT> span((count - 1) * gpu_begin.stride() + 1);
  backend::memory_read(gpu_begin.handle(), sizeof(NumericT) * gpu_begin.offset(), sizeof(NumericT) * span.size(), span.data());
  for (vcl_size_t i = 0; i < count; ++i) out[i] = span[i * gpu_begin.stride()];
}

/** fast_copy(cpu_begin, cpu_end, gpu_begin) (vector.hpp:1399-1426): linear host memory -> device range; a
 *  strided range is read, patched and written back so that the gaps keep their values. */
template <typename CpuIt, typename NumericT, unsigned int AlignmentV>
void fast_copy(CpuIt const& cpu_begin, CpuIt const& cpu_end, vector_iterator<NumericT, AlignmentV> gpu_begin)
{
  if (cpu_end - cpu_begin <= 0) return;
  const vcl_size_t count = vcl_size_t(cpu_end - cpu_begin);
  const NumericT* in = &(*cpu_begin);
  if (gpu_begin.stride() == 1) {
    backend::memory_write(gpu_begin.handle(), sizeof(NumericT) * gpu_begin.offset(), sizeof(NumericT) * count, in);
    return;
  }
  std::vector<NumericT> span((count - 1) * gpu_begin.stride() + 1);
  backend::memory_read(gpu_begin.handle(), sizeof(NumericT) * gpu_begin.offset(), sizeof(NumericT) * span.size(), span.data());
  for (vcl_size_t i = 0; i < count; ++i) span[i * gpu_begin.stride()] = in[i];
  backend::memory_write(gpu_begin.handle(), sizeof(NumericT) * gpu_begin.offset(), sizeof(NumericT) * span.size(), span.data());
}

template <typename NumericT>
void fast_copy(vector_base<NumericT> const& gpu, std::vector<NumericT>& cpu)
{
  cpu.resize(gpu.size());
  if (gpu.size()) fast_copy(gpu.begin(), gpu.end(), cpu.begin());
}

template <typename NumericT>
void fast_copy(std::vector<NumericT> const& cpu, vector_base<NumericT>& gpu)
{
  if (gpu.size() == 0) gpu.init(cpu.size(), viennacl::context(HIP_MEMORY));
  gpu.check_size(cpu.size());
  if (cpu.size()) fast_copy(cpu.begin(), cpu.end(), gpu.begin());
}

/** copy(gpu_begin, gpu_end, cpu_begin) (vector.hpp:1321-1356): device range -> any host output iterator,
 *  converting to its value type. */
template <typename NumericT, unsigned int AlignmentV, typename CpuIt>
void copy(const_vector_iterator<NumericT, AlignmentV> const& gpu_begin,
          const_vector_iterator<NumericT, AlignmentV> const& gpu_end, CpuIt cpu_begin)
{
  if (gpu_end - gpu_begin <= 0) return;
  std::vector<NumericT> tmp(static_cast<vcl_size_t>(gpu_end - gpu_begin));
  fast_copy(gpu_begin, gpu_end, tmp.begin());
  std::copy(tmp.begin(), tmp.end(), cpu_begin);
}

template <typename NumericT, unsigned int AlignmentV, typename CpuIt>
void copy(vector_iterator<NumericT, AlignmentV> const& gpu_begin, vector_iterator<NumericT, AlignmentV> const& gpu_end,
          CpuIt cpu_begin)
{
  copy(static_cast<const_vector_iterator<NumericT, AlignmentV> const&>(gpu_begin),
       static_cast<const_vector_iterator<NumericT, AlignmentV> const&>(gpu_end), cpu_begin);
}

/** copy(cpu_begin, cpu_end, gpu_begin) (vector.hpp:1487-1500): any host input range -> device range. */
template <typename CpuIt, typename NumericT, unsigned int AlignmentV>
void copy(CpuIt const& cpu_begin, CpuIt const& cpu_end, vector_iterator<NumericT, AlignmentV> gpu_begin)
{
  const vcl_ptrdiff_t count = static_cast<vcl_ptrdiff_t>(std::distance(cpu_begin, cpu_end));
  if (count <= 0) return;
  std::vector<NumericT> tmp(static_cast<vcl_size_t>(count));
  std::copy(cpu_begin, cpu_end, tmp.begin());
  fast_copy(tmp.begin(), tmp.end(), gpu_begin);
}

/** Device range -> device range (vector.hpp:1565-1600): an entry-wise copy on the device. */
template <typename NumericT, unsigned int AlignmentV>
void copy(const_vector_iterator<NumericT, AlignmentV> const& src_begin,
          const_vector_iterator<NumericT, AlignmentV> const& src_end, vector_iterator<NumericT, AlignmentV> dst_begin)
{
  if (src_end - src_begin <= 0) return;
  const vcl_size_t count = vcl_size_t(src_end - src_begin);
  vector_base<NumericT> src(const_cast<backend::mem_handle&>(src_begin.handle()), count, src_begin.offset(), src_begin.stride());
  vector_base<NumericT> dst(dst_begin.handle(), count, dst_begin.offset(), dst_begin.stride());
  dst.assign_linear(src, NumericT(1));
}

template <typename NumericT, unsigned int AlignmentV>
void copy(vector_iterator<NumericT, AlignmentV> const& src_begin, vector_iterator<NumericT, AlignmentV> const& src_end,
          vector_iterator<NumericT, AlignmentV> dst_begin)
{
  copy(static_cast<const_vector_iterator<NumericT, AlignmentV> const&>(src_begin),
       static_cast<const_vector_iterator<NumericT, AlignmentV> const&>(src_end), dst_begin);
}

/** Precision conversion y = x between float and double vectors (vector.hpp:410-427 -> convert). */
inline void convert(vector_base<float>& dst, vector_base<double> const& src)
{
  if (dst.size() == 0) dst.init(src.size(), viennacl::context(HIP_MEMORY));
  dst.check_size(src.size());
  if (!dst.contiguous() || !src.contiguous()) throw memory_exception("convert needs contiguous vectors");
  backend::hip::check(vclhip_convert_f64_to_f32(uint32_t(src.size()), src.device_ptr(), dst.device_ptr(), detail::stream()),
                      "convert");
}

inline void convert(vector_base<double>& dst, vector_base<float> const& src)
{
  if (dst.size() == 0) dst.init(src.size(), viennacl::context(HIP_MEMORY));
  dst.check_size(src.size());
  if (!dst.contiguous() || !src.contiguous()) throw memory_exception("convert needs contiguous vectors");
  backend::hip::check(vclhip_convert_f32_to_f64(uint32_t(src.size()), src.device_ptr(), dst.device_ptr(), detail::stream()),
                      "convert");
}

namespace traits
{
template <typename T> vcl_size_t size(vector_base<T> const& v) { return v.size(); }
template <typename T> viennacl::context context(vector_base<T> const& v) { return context_of(v.handle()); }
template <typename T> void clear(vector_base<T>& v) { v.clear(); }
/** The memory domain of an object's buffer (traits/handle.hpp:155-165): HIP_MEMORY for everything
 *  that computes in this build. */
template <typename T> memory_types active_handle_id(T const& obj) { return obj.handle().get_active_handle_id(); }
}  // namespace traits

}  // namespace viennacl

#endif
