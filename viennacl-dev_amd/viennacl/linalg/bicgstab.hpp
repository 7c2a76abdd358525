#ifndef VIENNACL_LINALG_BICGSTAB_HPP
#define VIENNACL_LINALG_BICGSTAB_HPP
// bicgstab.hpp -- stabilized bi-conjugate gradients (reference linalg/bicgstab.hpp).  For
// compressed_matrix and sliced_ell_matrix with no preconditioner or Jacobi, solve() runs the
// device-resident BiCGStab of libvclhip; other preconditioners run the left-preconditioned
// BiCGStab with restarts of bicgstab.hpp:398-489 on the HIP vector operations.

#include <cmath>
#include <map>
#include <vector>

#include "viennacl/linalg/detail/solver_glue.hpp"
#include "viennacl/linalg/jacobi_precond.hpp"

namespace viennacl
{
namespace linalg
{

/** Solver parameters and results (bicgstab.hpp:47-87). */
class bicgstab_tag
{
public:
  bicgstab_tag(double tol = 1e-8, vcl_size_t max_iters = 400, vcl_size_t max_iters_before_restart = 200)
    : tol_(tol), abs_tol_(0), iterations_(max_iters), iterations_before_restart_(max_iters_before_restart),
      iters_taken_(0), last_error_(0) {}
  double tolerance() const { return tol_; }
  double abs_tolerance() const { return abs_tol_; }
  void abs_tolerance(double new_tol) { if (new_tol >= 0) abs_tol_ = new_tol; }
  vcl_size_t max_iterations() const { return iterations_; }
  vcl_size_t max_iterations_before_restart() const { return iterations_before_restart_; }
  vcl_size_t iters() const { return iters_taken_; }
  void iters(vcl_size_t i) const { iters_taken_ = i; }
  double error() const { return last_error_; }
  void error(double e) const { last_error_ = e; }

private:
  double tol_;
  double abs_tol_;
  vcl_size_t iterations_;
  vcl_size_t iterations_before_restart_;
  mutable vcl_size_t iters_taken_;
  mutable double last_error_;
};

namespace detail
{

template <typename MatrixT, typename T>
vector<T> bicgstab_device(MatrixT const& A, vector_base<T> const& rhs, bicgstab_tag const& tag, int precond,
                          bool (*monitor)(vector<T> const&, T, void*) = nullptr, void* monitor_data = nullptr)
{
  vector<T> b = contiguous_copy(rhs);
  vector<T> x(rhs.size());
  vclhip_solver_params P = default_params();
  P.dia = dia_of(A);
  P.tol = tag.tolerance();
  P.abs_tol = tag.abs_tolerance();
  P.max_iters = static_cast<uint32_t>(tag.max_iterations());
  P.max_iters_before_restart = static_cast<uint32_t>(tag.max_iterations_before_restart());
  monitor_bridge<vector<T> > bridge = {monitor, monitor_data, rhs.size()};
  if (monitor) {
    P.monitor = &monitor_bridge<vector<T> >::call;
    P.monitor_user = &bridge;
  }
  vclhip_solver_result R;
  if (rhs.size()) {
    backend::hip::check(bicgstab_solve(A.descriptor(), b.device_ptr(), x.device_ptr(), precond, P, R),
                        "bicgstab solve");
    tag.iters(R.iters);
    tag.error(R.error);
  }
  return x;
}

/** Left-preconditioned BiCGStab with restarts for a user preconditioner (bicgstab.hpp:398-489).
 *  The restart on a degenerate direction or after max_iterations_before_restart iterations and
 *  the mixed stopping norm ||M^-1 r|| / ||b|| follow the reference. */
template <typename MatrixT, typename VectorT, typename PreconditionerT>
VectorT bicgstab_generic(MatrixT const& matrix, VectorT const& rhs, bicgstab_tag const& tag,
                         PreconditionerT const& precond,
                         bool (*monitor)(VectorT const&, typename VectorT::value_type, void*) = nullptr,
                         void* monitor_data = nullptr)
{
  typedef typename VectorT::value_type T;
  VectorT result(rhs.size());
  VectorT residual = rhs, r0star = rhs, tmp0 = rhs, tmp1 = rhs, s = rhs, p = rhs;
  T norm_rhs = norm_2(rhs);
  T ip_rr0star = norm_rhs, residual_norm = norm_rhs;
  if (norm_rhs <= tag.abs_tolerance()) return result;

  bool restart = true;
  vcl_size_t last_restart = 0;
  for (vcl_size_t i = 0; i < tag.max_iterations(); ++i) {
    if (restart) {
      residual = prod(matrix, result);
      residual = rhs - residual;
      precond.apply(residual);
      p = residual;
      r0star = residual;
      ip_rr0star = norm_2(residual);
      ip_rr0star *= ip_rr0star;
      restart = false;
      last_restart = i;
    }
    tag.iters(i + 1);
    tmp0 = prod(matrix, p);
    precond.apply(tmp0);
    T alpha = ip_rr0star / inner_prod(tmp0, r0star);
    s = residual - alpha * tmp0;
    tmp1 = prod(matrix, s);
    precond.apply(tmp1);
    T n1 = norm_2(tmp1);
    T omega = inner_prod(tmp1, s) / (n1 * n1);
    result += alpha * p + omega * s;
    residual = s - omega * tmp1;
    residual_norm = norm_2(residual);
    if (monitor && monitor(result, std::fabs(residual_norm / norm_rhs), monitor_data)) break;
    if (residual_norm / norm_rhs < tag.tolerance() || residual_norm < tag.abs_tolerance()) break;
    T new_ip = inner_prod(residual, r0star);
    T beta = new_ip / ip_rr0star * alpha / omega;
    ip_rr0star = new_ip;
    if (ip_rr0star == T(0) || omega == T(0) || i - last_restart > tag.max_iterations_before_restart()) restart = true;
    p -= omega * tmp0;
    p = residual + beta * p;
  }
  tag.error(residual_norm / norm_rhs);
  return result;
}

template <typename MatrixT, typename T, typename PreconditionerT>
typename std::enable_if<is_hip_sparse<MatrixT>::value, vector<T> >::type
solve_impl(MatrixT const& A, vector_base<T> const& rhs, bicgstab_tag const& tag, PreconditionerT const& precond,
           bool (*monitor)(vector<T> const&, T, void*) = nullptr, void* monitor_data = nullptr)
{
  if (std::is_same<PreconditionerT, no_precond>::value)
    return bicgstab_device(A, rhs, tag, VCLHIP_PRECOND_NONE, monitor, monitor_data);
  if (is_jacobi<PreconditionerT>::value)
    return bicgstab_device(A, rhs, tag, VCLHIP_PRECOND_JACOBI, monitor, monitor_data);
  vector<T> b(rhs);
  return bicgstab_generic(A, b, tag, precond, monitor, monitor_data);
}
/** A user operator (apply(x, y) and size1(); examples/tutorial/matrix-free.cpp:58-82): the classic
 *  recurrences on the HIP vector operations, each product through apply(). */
template <typename MatrixT, typename T, typename PreconditionerT>
typename std::enable_if<!is_hip_sparse<MatrixT>::value, vector<T> >::type
solve_impl(MatrixT const& A, vector_base<T> const& rhs, bicgstab_tag const& tag, PreconditionerT const& precond,
           bool (*monitor)(vector<T> const&, T, void*) = nullptr, void* monitor_data = nullptr)
{
  vector<T> b(rhs);
  return bicgstab_generic(A, b, tag, precond, monitor, monitor_data);
}
}  // namespace detail

/** solve(A, b, bicgstab_tag, precond) (bicgstab.hpp:493-500). */
template <typename MatrixT, typename VectorT, typename PreconditionerT>
vector<typename VectorT::value_type> solve(MatrixT const& matrix, VectorT const& rhs, bicgstab_tag const& tag,
                                           PreconditionerT const& precond)
{
  return detail::solve_impl(matrix, static_cast<vector_base<typename VectorT::value_type> const&>(rhs), tag, precond);
}

/** solve(A, b, bicgstab_tag) (bicgstab.hpp:525-529). */
template <typename MatrixT, typename VectorT>
vector<typename VectorT::value_type> solve(MatrixT const& matrix, VectorT const& rhs, bicgstab_tag const& tag)
{
  return solve(matrix, rhs, tag, no_precond());
}

/** STL convenience overload (bicgstab.hpp:504-520). */
template <typename IndexT, typename NumericT, typename PreconditionerT>
std::vector<NumericT> solve(std::vector<std::map<IndexT, NumericT> > const& A, std::vector<NumericT> const& rhs,
                            bicgstab_tag const& tag, PreconditionerT const& precond)
{
  compressed_matrix<NumericT> vcl_A;
  copy(A, vcl_A);
  vector<NumericT> vcl_rhs(rhs.size());
  copy(rhs, vcl_rhs);
  vector<NumericT> vcl_result;
  if (detail::is_jacobi<PreconditionerT>::value || std::is_same<PreconditionerT, no_precond>::value)
    vcl_result = detail::bicgstab_device(vcl_A, vcl_rhs, tag,
                                         std::is_same<PreconditionerT, no_precond>::value ? VCLHIP_PRECOND_NONE
                                                                                           : VCLHIP_PRECOND_JACOBI);
  else
    vcl_result = detail::bicgstab_generic(vcl_A, vcl_rhs, tag, precond);
  std::vector<NumericT> result(vcl_result.size());
  copy(vcl_result, result);
  return result;
}

/** Solver object with an initial guess and a monitor (bicgstab.hpp:533-593). */
template <typename VectorT>
class bicgstab_solver
{
public:
  typedef typename VectorT::value_type numeric_type;

  bicgstab_solver(bicgstab_tag const& tag) : tag_(tag), monitor_callback_(nullptr), user_data_(nullptr) {}

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
  bicgstab_tag const& tag() const { return tag_; }

private:
  bicgstab_tag tag_;
  VectorT init_guess_;
  bool (*monitor_callback_)(VectorT const&, numeric_type, void*);
  void* user_data_;
};

}  // namespace linalg
}  // namespace viennacl

#endif
