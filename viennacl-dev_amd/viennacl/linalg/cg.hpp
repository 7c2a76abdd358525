#ifndef VIENNACL_LINALG_CG_HPP
#define VIENNACL_LINALG_CG_HPP
// cg.hpp -- conjugate gradients (reference linalg/cg.hpp).  For compressed_matrix and
// sliced_ell_matrix with no preconditioner or Jacobi, solve() runs the device-resident pipelined CG
// of libvclhip (no per-iteration host round trip); other preconditioners run the classic
// preconditioned CG of cg.hpp:257-322 on the HIP vector operations.

#include <cmath>
#include <map>
#include <vector>

#include "viennacl/linalg/detail/solver_glue.hpp"
#include "viennacl/linalg/jacobi_precond.hpp"

namespace viennacl
{
namespace linalg
{

/** Solver parameters and results (cg.hpp:48-87). */
class cg_tag
{
public:
  cg_tag(double tol = 1e-8, unsigned int max_iterations = 300)
    : tol_(tol), abs_tol_(0), iterations_(max_iterations), iters_taken_(0), last_error_(0) {}
  double tolerance() const { return tol_; }
  double abs_tolerance() const { return abs_tol_; }
  void abs_tolerance(double new_tol) { if (new_tol >= 0) abs_tol_ = new_tol; }
  unsigned int max_iterations() const { return iterations_; }
  unsigned int iters() const { return iters_taken_; }
  void iters(unsigned int i) const { iters_taken_ = i; }
  double error() const { return last_error_; }
  void error(double e) const { last_error_ = e; }

private:
  double tol_;
  double abs_tol_;
  unsigned int iterations_;
  mutable unsigned int iters_taken_;
  mutable double last_error_;
};

namespace detail
{

/** Device-resident CG for the sparse types with no preconditioner or Jacobi. */
template <typename MatrixT, typename T>
vector<T> cg_device(MatrixT const& A, vector_base<T> const& rhs, cg_tag const& tag, int precond,
                    bool (*monitor)(vector<T> const&, T, void*) = nullptr, void* monitor_data = nullptr)
{
  vector<T> b = contiguous_copy(rhs);
  vector<T> x(rhs.size());
  vclhip_solver_params P = default_params();
  P.dia = dia_of(A);
  P.tol = tag.tolerance();
  P.abs_tol = tag.abs_tolerance();
  P.max_iters = tag.max_iterations();
  monitor_bridge<vector<T> > bridge = {monitor, monitor_data, rhs.size()};
  if (monitor) {
    P.monitor = &monitor_bridge<vector<T> >::call;
    P.monitor_user = &bridge;
  }
  vclhip_solver_result R;
  if (rhs.size()) {
    backend::hip::check(cg_solve(A.descriptor(), b.device_ptr(), x.device_ptr(), precond, P, R), "cg solve");
    tag.iters(R.iters);
    tag.error(R.error);
  }
  return x;
}

/** Classic preconditioned CG with a user preconditioner (cg.hpp:257-322, Saad Alg. 9.1). */
template <typename MatrixT, typename VectorT, typename PreconditionerT>
VectorT cg_generic(MatrixT const& matrix, VectorT const& rhs, cg_tag const& tag, PreconditionerT const& precond,
                   bool (*monitor)(VectorT const&, typename VectorT::value_type, void*) = nullptr,
                   void* monitor_data = nullptr)
{
  typedef typename VectorT::value_type T;
  VectorT result(rhs.size());
  VectorT residual = rhs;
  VectorT tmp(rhs.size());
  VectorT z = rhs;
  precond.apply(z);
  VectorT p = z;
  T ip_rr = inner_prod(residual, z);
  T new_ip_rr = 0;
  T norm_rhs_squared = ip_rr;
  if (std::fabs(norm_rhs_squared) <= tag.abs_tolerance() * tag.abs_tolerance()) return result;
  for (unsigned int i = 0; i < tag.max_iterations(); ++i) {
    tag.iters(i + 1);
    tmp = prod(matrix, p);
    T alpha = ip_rr / inner_prod(tmp, p);
    result += alpha * p;
    residual -= alpha * tmp;
    z = residual;
    precond.apply(z);
    new_ip_rr = inner_prod(residual, z);
    T ratio = new_ip_rr / norm_rhs_squared;
    if (monitor && monitor(result, std::sqrt(std::fabs(ratio)), monitor_data)) break;
    if (std::fabs(ratio) < tag.tolerance() * tag.tolerance() || std::fabs(new_ip_rr) < tag.abs_tolerance() * tag.abs_tolerance())
      break;
    T beta = new_ip_rr / ip_rr;
    ip_rr = new_ip_rr;
    p = z + beta * p;
  }
  tag.error(std::sqrt(std::fabs(new_ip_rr / norm_rhs_squared)));
  return result;
}

template <typename MatrixT, typename T, typename PreconditionerT>
typename std::enable_if<is_hip_sparse<MatrixT>::value, vector<T> >::type
solve_impl(MatrixT const& A, vector_base<T> const& rhs, cg_tag const& tag, PreconditionerT const& precond,
           bool (*monitor)(vector<T> const&, T, void*) = nullptr, void* monitor_data = nullptr)
{
  if (std::is_same<PreconditionerT, no_precond>::value)
    return cg_device(A, rhs, tag, VCLHIP_PRECOND_NONE, monitor, monitor_data);
  if (is_jacobi<PreconditionerT>::value)
    return cg_device(A, rhs, tag, VCLHIP_PRECOND_JACOBI, monitor, monitor_data);
  vector<T> b(rhs);
  return cg_generic(A, b, tag, precond, monitor, monitor_data);
}
/** A user operator (apply(x, y) and size1(); examples/tutorial/matrix-free.cpp:58-82): the classic
 *  recurrences on the HIP vector operations, each product through apply(). */
template <typename MatrixT, typename T, typename PreconditionerT>
typename std::enable_if<!is_hip_sparse<MatrixT>::value, vector<T> >::type
solve_impl(MatrixT const& A, vector_base<T> const& rhs, cg_tag const& tag, PreconditionerT const& precond,
           bool (*monitor)(vector<T> const&, T, void*) = nullptr, void* monitor_data = nullptr)
{
  vector<T> b(rhs);
  return cg_generic(A, b, tag, precond, monitor, monitor_data);
}
}  // namespace detail

/** solve(A, b, cg_tag, precond) (cg.hpp:338-342). */
template <typename MatrixT, typename VectorT, typename PreconditionerT>
vector<typename VectorT::value_type> solve(MatrixT const& matrix, VectorT const& rhs, cg_tag const& tag,
                                           PreconditionerT const& precond)
{
  return detail::solve_impl(matrix, static_cast<vector_base<typename VectorT::value_type> const&>(rhs), tag, precond);
}

/** solve(A, b, cg_tag) (cg.hpp:371-375). */
template <typename MatrixT, typename VectorT>
vector<typename VectorT::value_type> solve(MatrixT const& matrix, VectorT const& rhs, cg_tag const& tag)
{
  return solve(matrix, rhs, tag, no_precond());
}

/** STL convenience overload (cg.hpp:349-363). */
template <typename IndexT, typename NumericT, typename PreconditionerT>
std::vector<NumericT> solve(std::vector<std::map<IndexT, NumericT> > const& A, std::vector<NumericT> const& rhs,
                            cg_tag const& tag, PreconditionerT const& precond)
{
  compressed_matrix<NumericT> vcl_A;
  copy(A, vcl_A);
  vector<NumericT> vcl_rhs(rhs.size());
  copy(rhs, vcl_rhs);
  vector<NumericT> vcl_result;
  if (detail::is_jacobi<PreconditionerT>::value || std::is_same<PreconditionerT, no_precond>::value)
    vcl_result = detail::cg_device(vcl_A, vcl_rhs, tag,
                                   std::is_same<PreconditionerT, no_precond>::value ? VCLHIP_PRECOND_NONE
                                                                                     : VCLHIP_PRECOND_JACOBI);
  else
    vcl_result = detail::cg_generic(vcl_A, vcl_rhs, tag, precond);
  std::vector<NumericT> result(vcl_result.size());
  copy(vcl_result, result);
  return result;
}

/** Solver object with an initial guess and a per-iteration monitor (cg.hpp:379-434). */
template <typename VectorT>
class cg_solver
{
public:
  typedef typename VectorT::value_type numeric_type;

  cg_solver(cg_tag const& tag) : tag_(tag), monitor_callback_(nullptr), user_data_(nullptr) {}

  template <typename MatrixT, typename PreconditionerT>
  VectorT operator()(MatrixT const& A, VectorT const& b, PreconditionerT const& precond) const
  {
    if (init_guess_.size() > 0) {
      VectorT mod_rhs = prod(A, init_guess_);
      mod_rhs = b - mod_rhs;
      VectorT y = detail::solve_impl(A, mod_rhs, tag_, precond, monitor_callback_, user_data_);
      return init_guess_ + y;
    }
    return detail::solve_impl(A, b, tag_, precond, monitor_callback_, user_data_);
  }

  template <typename MatrixT>
  VectorT operator()(MatrixT const& A, VectorT const& b) const
  {
    return operator()(A, b, no_precond());
  }

  void set_initial_guess(VectorT const& x) { init_guess_ = x; }
  void set_monitor(bool (*monitor_fun)(VectorT const&, numeric_type, void*), void* user_data)
  {
    monitor_callback_ = monitor_fun;
    user_data_ = user_data;
  }
  cg_tag const& tag() const { return tag_; }

private:
  cg_tag tag_;
  VectorT init_guess_;
  bool (*monitor_callback_)(VectorT const&, numeric_type, void*);
  void* user_data_;
};

}  // namespace linalg
}  // namespace viennacl

#endif
