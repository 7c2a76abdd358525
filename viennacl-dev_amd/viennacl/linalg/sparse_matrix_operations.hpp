#ifndef VIENNACL_LINALG_SPARSE_MATRIX_OPERATIONS_HPP
#define VIENNACL_LINALG_SPARSE_MATRIX_OPERATIONS_HPP
// sparse_matrix_operations.hpp -- prod_impl and row_info for the sparse formats (CSR, SELL-C, COO,
// ELL, HYB), the operations behind y = prod(A, x), y += prod(A, x), y -= prod(A, x)
// (reference linalg/sparse_matrix_operations.hpp:48-121 and the op_executors of
// compressed_matrix.hpp:1231-1318 / sliced_ell_matrix.hpp:286-375).  HIP_MEMORY only.

#include <map>
#include <type_traits>
#include <utility>
#include <vector>

#include "viennacl/forwards.h"
#include "viennacl/vector.hpp"

namespace viennacl
{
namespace linalg
{
namespace hip
{
inline int spmv(vclhip_csr const& d, const double* x, uint32_t xs, uint32_t xi, double a, double* y, uint32_t ys,
                uint32_t yi, double b)
{
  return vclhip_csr_spmv_f64(&d, x, xs, xi, a, y, ys, yi, b, backend::hip::current_stream());
}
inline int spmv(vclhip_csr const& d, const float* x, uint32_t xs, uint32_t xi, float a, float* y, uint32_t ys,
                uint32_t yi, float b)
{
  return vclhip_csr_spmv_f32(&d, x, xs, xi, a, y, ys, yi, b, backend::hip::current_stream());
}
inline int spmv(vclhip_sell const& d, const double* x, uint32_t xs, uint32_t xi, double a, double* y, uint32_t ys,
                uint32_t yi, double b)
{
  return vclhip_sell_spmv_f64(&d, x, xs, xi, a, y, ys, yi, b, backend::hip::current_stream());
}
inline int spmv(vclhip_sell const& d, const float* x, uint32_t xs, uint32_t xi, float a, float* y, uint32_t ys,
                uint32_t yi, float b)
{
  return vclhip_sell_spmv_f32(&d, x, xs, xi, a, y, ys, yi, b, backend::hip::current_stream());
}
#define VCLHIP_FORMAT_SPMV(DESC, FN)                                                                          \
  inline int spmv(DESC const& d, const double* x, uint32_t xs, uint32_t xi, double a, double* y, uint32_t ys,  \
                  uint32_t yi, double b)                                                                       \
  {                                                                                                            \
    return FN##_f64(&d, x, xs, xi, a, y, ys, yi, b, backend::hip::current_stream());                            \
  }                                                                                                            \
  inline int spmv(DESC const& d, const float* x, uint32_t xs, uint32_t xi, float a, float* y, uint32_t ys,     \
                  uint32_t yi, float b)                                                                        \
  {                                                                                                            \
    return FN##_f32(&d, x, xs, xi, a, y, ys, yi, b, backend::hip::current_stream());                            \
  }
VCLHIP_FORMAT_SPMV(vclhip_coo, vclhip_coo_spmv)
VCLHIP_FORMAT_SPMV(vclhip_ell, vclhip_ell_spmv)
VCLHIP_FORMAT_SPMV(vclhip_hyb, vclhip_hyb_spmv)
#undef VCLHIP_FORMAT_SPMV

/** The format-generic descriptor of a matrix (vclhip_matrix) for the solvers and the diagonal. */
inline int format_of(vclhip_csr const&) { return VCLHIP_FORMAT_CSR; }
inline int format_of(vclhip_sell const&) { return VCLHIP_FORMAT_SELL; }
inline int format_of(vclhip_coo const&) { return VCLHIP_FORMAT_COO; }
inline int format_of(vclhip_ell const&) { return VCLHIP_FORMAT_ELL; }
inline int format_of(vclhip_hyb const&) { return VCLHIP_FORMAT_HYB; }
template <typename DescT>
vclhip_matrix generic(DescT const& d)
{
  vclhip_matrix m;
  m.format = format_of(d);
  m.desc = &d;
  return m;
}
}  // namespace hip

/** y = alpha * A * x + beta * y (linalg/sparse_matrix_operations.hpp:90-121).  x and y must not alias. */
template <typename MatrixT, typename T>
void prod_impl(MatrixT const& A, vector_base<T> const& x, T alpha, vector_base<T>& y, T beta)
{
  if (x.size() != A.size2() || y.size() != A.size1())
    throw std::invalid_argument("ViennaCL (HIP): prod_impl size mismatch");
  x.require_hip();
  y.require_hip();
  if (A.size1() == 0) return;
  backend::hip::check(hip::spmv(A.descriptor(), x.device_ptr(), uint32_t(x.start()), uint32_t(x.stride()), alpha,
                                y.device_ptr(), uint32_t(y.start()), uint32_t(y.stride()), beta),
                      "prod_impl");
}

/** prod(A, x) as a lazy expression (linalg/prod.hpp:350-360). */
template <typename MatrixT, typename T>
vector_expression<MatrixT, vector_base<T>, op_prod> prod(MatrixT const& A, vector_base<T> const& x)
{
  return vector_expression<MatrixT, vector_base<T>, op_prod>(A, x);
}

namespace detail
{
/** Per-row information (linalg/sparse_matrix_operations.hpp:48-74): norms or the diagonal. */
template <typename T, unsigned int A>
void row_info(compressed_matrix<T, A> const& mat, vector_base<T>& vec, row_info_types info)
{
  if (vec.size() == 0) vec.init(mat.size1(), viennacl::context(HIP_MEMORY));
  if (!vec.contiguous()) throw memory_exception("row_info needs a contiguous vector");
  vclhip_csr d = mat.descriptor();
  if (std::is_same<T, double>::value)
    backend::hip::check(vclhip_csr_row_info_f64(&d, reinterpret_cast<double*>(vec.device_ptr()), int(info),
                                                backend::hip::current_stream()), "row_info");
  else
    backend::hip::check(vclhip_csr_row_info_f32(&d, reinterpret_cast<float*>(vec.device_ptr()), int(info),
                                                backend::hip::current_stream()), "row_info");
}

/** Diagonal of a COO, ELL or HYB matrix (row_info with SPARSE_ROW_DIAGONAL; COO:
 *  host_based/sparse_matrix_operations.hpp:1156-1205).  Other row_info kinds exist for CSR only. */
template <typename MatrixT, typename T>
typename std::enable_if<!std::is_same<MatrixT, compressed_matrix<T, 1> >::value>::type
row_info(MatrixT const& mat, vector_base<T>& vec, row_info_types info)
{
  if (info != SPARSE_ROW_DIAGONAL) throw std::invalid_argument("ViennaCL (HIP): row_info: only the diagonal for this format");
  if (vec.size() == 0) vec.init(mat.size1(), viennacl::context(HIP_MEMORY));
  if (!vec.contiguous()) throw memory_exception("row_info needs a contiguous vector");
  auto d = mat.descriptor();
  vclhip_matrix m = hip::generic(d);
  if (std::is_same<T, double>::value)
    backend::hip::check(vclhip_matrix_diag_f64(&m, reinterpret_cast<double*>(vec.device_ptr()),
                                               backend::hip::current_stream()), "row_info");
  else
    backend::hip::check(vclhip_matrix_diag_f32(&m, reinterpret_cast<float*>(vec.device_ptr()),
                                               backend::hip::current_stream()), "row_info");
}
}  // namespace detail
}  // namespace linalg

namespace linalg
{
namespace detail
{
/** Whether MatrixT is one of the sparse formats (it hands a device descriptor to the C ABI) or a
 *  user-provided operator, which only has apply(x, y) and size1() (examples/tutorial/matrix-free.cpp:58-82). */
template <typename MatrixT, typename = void>
struct has_descriptor : std::false_type {};
template <typename MatrixT>
struct has_descriptor<MatrixT, decltype(void(std::declval<MatrixT const&>().descriptor()))> : std::true_type {};

/** y = A x for a sparse matrix (y, x distinct). */
template <typename MatrixT, typename T>
typename std::enable_if<has_descriptor<MatrixT>::value>::type
apply_operator(MatrixT const& A, vector_base<T> const& x, vector_base<T>& y)
{
  linalg::prod_impl(A, x, T(1), y, T(0));
}

/** y = A x for a user operator: its apply() (the generic op_executor of vector.hpp:3315-3320). */
template <typename MatrixT, typename T>
typename std::enable_if<!has_descriptor<MatrixT>::value>::type
apply_operator(MatrixT const& A, vector_base<T> const& x, vector_base<T>& y)
{
  A.apply(x, y);
}

/** y += sign * A x for a sparse matrix: one fused product (beta = 1). */
template <typename MatrixT, typename T>
typename std::enable_if<has_descriptor<MatrixT>::value>::type
accumulate_operator(MatrixT const& A, vector_base<T> const& x, T sign, vector_base<T>& y)
{
  linalg::prod_impl(A, x, sign, y, T(1));
}

/** y += sign * A x for a user operator, through a temporary. */
template <typename MatrixT, typename T>
typename std::enable_if<!has_descriptor<MatrixT>::value>::type
accumulate_operator(MatrixT const& A, vector_base<T> const& x, T sign, vector_base<T>& y)
{
  vector_base<T> tmp(A.size1());
  A.apply(x, tmp);
  y.require_hip();
  backend::hip::check(viennacl::detail::avbv(y, y, T(1), tmp, sign), "y += A x");
}
}  // namespace detail
}  // namespace linalg

/** y = prod(A, x) for a sparse matrix or a user operator; aliasing of y and x goes through a
 *  temporary (compressed_matrix.hpp:1231-1246). */
template <typename T, typename MatrixT>
void vcl_eval(vector_base<T>& y, vector_expression<MatrixT, vector_base<T>, op_prod> const& e)
{
  MatrixT const& A = e.lhs();
  vector_base<T> const& x = e.rhs();
  if (y.size() == 0) y.init(A.size1(), viennacl::context(HIP_MEMORY));
  if (y.handle() == x.handle()) {
    vector_base<T> tmp(A.size1());
    linalg::detail::apply_operator(A, x, tmp);
    y.assign_linear(tmp, T(1));
    return;
  }
  linalg::detail::apply_operator(A, x, y);
}

/** y += sign * prod(A, x) (compressed_matrix.hpp:1248-1318); a user operator's product goes
 *  through a temporary. */
template <typename T, typename MatrixT>
void vcl_inplace(vector_base<T>& y, vector_expression<MatrixT, vector_base<T>, op_prod> const& e, T sign)
{
  MatrixT const& A = e.lhs();
  vector_base<T> const& x = e.rhs();
  if (y.handle() == x.handle()) {
    vector_base<T> tmp(A.size1());
    linalg::detail::apply_operator(A, x, tmp);
    y.require_hip();
    backend::hip::check(detail::avbv(y, y, T(1), tmp, sign), "y += A x");
    return;
  }
  linalg::detail::accumulate_operator(A, x, sign, y);
}

}  // namespace viennacl

#endif
