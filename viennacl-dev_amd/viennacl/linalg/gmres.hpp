#ifndef VIENNACL_LINALG_GMRES_HPP
#define VIENNACL_LINALG_GMRES_HPP
// gmres.hpp -- restarted GMRES (reference linalg/gmres.hpp).  For compressed_matrix and
// sliced_ell_matrix without a preconditioner, solve() runs the device-resident pipelined GMRES of
// libvclhip (gmres.hpp:182-370: Walker's simpler GMRES with classical Gram-Schmidt); with a
// preconditioner it runs the Householder GMRES of gmres.hpp:436-626 on the HIP vector operations.

#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "viennacl/linalg/detail/solver_glue.hpp"
#include "viennacl/linalg/jacobi_precond.hpp"

namespace viennacl
{
namespace linalg
{

/** Solver parameters and results (gmres.hpp:47-100). */
class gmres_tag
{
public:
  gmres_tag(double tol = 1e-10, unsigned int max_iterations = 300, unsigned int krylov_dim = 20)
    : tol_(tol), abs_tol_(0), iterations_(max_iterations), krylov_dim_(krylov_dim), iters_taken_(0),
      last_error_(0) {}
  double tolerance() const { return tol_; }
  double abs_tolerance() const { return abs_tol_; }
  void abs_tolerance(double new_tol) { if (new_tol >= 0) abs_tol_ = new_tol; }
  unsigned int max_iterations() const { return iterations_; }
  unsigned int krylov_dim() const { return krylov_dim_; }
  /** Cycles after the first: max_iterations / krylov_dim, one less when that divides evenly. */
  unsigned int max_restarts() const
  {
    unsigned int ret = iterations_ / krylov_dim_;
    if (ret > 0 && ret * krylov_dim_ == iterations_) return ret - 1;
    return ret;
  }
  unsigned int iters() const { return iters_taken_; }
  void iters(unsigned int i) const { iters_taken_ = i; }
  double error() const { return last_error_; }
  void error(double e) const { last_error_ = e; }

private:
  double tol_;
  double abs_tol_;
  unsigned int iterations_;
  unsigned int krylov_dim_;
  mutable unsigned int iters_taken_;
  mutable double last_error_;
};

namespace detail
{

inline int gmres_solve(vclhip_csr const& A, const double* b, double* x, vclhip_solver_params const& P,
                       vclhip_solver_result& R)
{ return vclhip_gmres_solve_csr_f64(&A, b, x, &P, &R, backend::hip::current_stream()); }
inline int gmres_solve(vclhip_csr const& A, const float* b, float* x, vclhip_solver_params const& P,
                       vclhip_solver_result& R)
{ return vclhip_gmres_solve_csr_f32(&A, b, x, &P, &R, backend::hip::current_stream()); }
inline int gmres_solve(vclhip_sell const& A, const double* b, double* x, vclhip_solver_params const& P,
                       vclhip_solver_result& R)
{ return vclhip_gmres_solve_sell_f64(&A, b, x, &P, &R, backend::hip::current_stream()); }
inline int gmres_solve(vclhip_sell const& A, const float* b, float* x, vclhip_solver_params const& P,
                       vclhip_solver_result& R)
{ return vclhip_gmres_solve_sell_f32(&A, b, x, &P, &R, backend::hip::current_stream()); }

template <typename DescT>
inline int gmres_solve(DescT const& A, const double* b, double* x, vclhip_solver_params const& P,
                       vclhip_solver_result& R)
{
  vclhip_matrix m = hip::generic(A);
  return vclhip_gmres_solve_f64(&m, b, x, &P, &R, backend::hip::current_stream());
}
template <typename DescT>
inline int gmres_solve(DescT const& A, const float* b, float* x, vclhip_solver_params const& P,
                       vclhip_solver_result& R)
{
  vclhip_matrix m = hip::generic(A);
  return vclhip_gmres_solve_f32(&m, b, x, &P, &R, backend::hip::current_stream());
}

/** Device-resident pipelined GMRES (gmres.hpp:182-370); the monitor is called once per cycle. */
template <typename MatrixT, typename T>
vector<T> gmres_device(MatrixT const& A, vector_base<T> const& rhs, gmres_tag const& tag,
                       bool (*monitor)(vector<T> const&, T, void*) = nullptr, void* monitor_data = nullptr)
{
  vector<T> b = contiguous_copy(rhs);
  vector<T> x(rhs.size());
  vclhip_solver_params P = default_params();
  P.dia = dia_of(A);
  P.tol = tag.tolerance();
  P.abs_tol = tag.abs_tolerance();
  P.max_iters = tag.max_iterations();
  P.krylov_dim = tag.krylov_dim();
  monitor_bridge<vector<T> > bridge = {monitor, monitor_data, rhs.size()};
  if (monitor) {
    P.monitor = &monitor_bridge<vector<T> >::call;
    P.monitor_user = &bridge;
  }
  vclhip_solver_result R;
  tag.iters(0);
  if (rhs.size()) {
    backend::hip::check(gmres_solve(A.descriptor(), b.device_ptr(), x.device_ptr(), P, R), "gmres solve");
    tag.iters(R.iters);
    tag.error(R.error);
  }
  return x;
}

// ---- Householder pieces of the preconditioned GMRES (gmres.hpp:110-166)

/** h = the Householder vector that zeroes the entries of `in` below j (gmres.hpp:130-158). */
template <typename T>
void gmres_householder_vector(vector_base<T> const& in, vector_base<T>& h, T& beta, T& mu, vcl_size_t j)
{
  const T in_j = in(j);
  const vcl_size_t n = h.size();
  if (j + 1 < n)  // h[j+1 ..) = in[j+1 ..); the entries up to j are zero
    backend::memory_copy(in.handle(), h.handle(), sizeof(T) * (j + 1), sizeof(T) * (j + 1), sizeof(T) * (n - j - 1));
  T sigma = norm_2(h);
  sigma *= sigma;
  if (sigma <= 0) {
    beta = 0;
    mu = in_j;
    return;
  }
  mu = std::sqrt(sigma + in_j * in_j);
  const T h0 = (in_j <= 0) ? (in_j - mu) : (-sigma / (in_j + mu));
  beta = T(2) * h0 * h0 / (sigma + h0 * h0);
  h /= h0;
  h[j] = T(1);
}

/** x = (I - beta h h^T) x  (gmres.hpp:161-166) */
template <typename T>
void gmres_reflect(vector_base<T>& x, vector_base<T> const& h, T beta)
{
  const T hx = inner_prod(h, x);
  x -= (beta * hx) * h;
}

/** Restarted GMRES with a preconditioner, Householder orthogonalization (gmres.hpp:436-626). */
template <typename MatrixT, typename VectorT, typename PreconditionerT>
VectorT gmres_generic(MatrixT const& matrix, VectorT const& rhs, gmres_tag const& tag,
                      PreconditionerT const& precond,
                      bool (*monitor)(VectorT const&, typename VectorT::value_type, void*) = nullptr,
                      void* monitor_data = nullptr)
{
  typedef typename VectorT::value_type T;
  const vcl_size_t n = rhs.size();
  VectorT result(n);
  const vcl_size_t K = std::min<vcl_size_t>(tag.krylov_dim(), n);
  VectorT res = rhs, vk = rhs, vtmp = rhs;
  std::vector<std::vector<T> > R(K, std::vector<T>(tag.krylov_dim()));
  std::vector<T> proj(K);
  std::vector<VectorT> hh(K, rhs);
  std::vector<T> betas(K);
  const T norm_rhs = norm_2(rhs);
  if (norm_rhs <= tag.abs_tolerance()) return result;
  tag.iters(0);
  for (unsigned int it = 0; it <= tag.max_restarts(); ++it) {
    res = prod(matrix, result);
    res = rhs - res;
    precond.apply(res);
    const T rho_0 = norm_2(res);
    if (rho_0 / norm_rhs < tag.tolerance() || rho_0 < tag.abs_tolerance()) {
      tag.error(rho_0 / norm_rhs);
      return result;
    }
    res /= rho_0;
    T rho = T(1);
    vcl_size_t k = 0;
    for (k = 0; k < K; ++k) {
      tag.iters(tag.iters() + 1);
      std::fill(R[k].begin(), R[k].end(), T(0));
      hh[k].clear();
      if (k == 0) {
        vk = prod(matrix, res);
        precond.apply(vk);
      } else {
        vk.clear();
        vk[k - 1] = T(1);
        for (int i = static_cast<int>(k) - 1; i > -1; --i) gmres_reflect(vk, hh[vcl_size_t(i)], betas[vcl_size_t(i)]);
        vtmp = prod(matrix, vk);
        precond.apply(vtmp);
        vk = vtmp;
        for (vcl_size_t i = 0; i < k; ++i) gmres_reflect(vk, hh[i], betas[i]);
      }
      T rho_kk = 0;
      gmres_householder_vector(vk, hh[k], betas[k], rho_kk, k);
      if (k > 0) fast_copy(vk.begin(), vk.begin() + static_cast<typename VectorT::difference_type>(k), &R[k][0]);
      R[k][k] = rho_kk;
      gmres_reflect(res, hh[k], betas[k]);
      T rk = res(k);
      if (rk > rho) rk = rho;  // machine precision reached
      if (rk < -rho) rk = -rho;
      res[k] = rk;
      proj[k] = rk;
      rho *= std::sin(std::acos(proj[k] / rho));
      if (std::fabs(rho * rho_0 / norm_rhs) < tag.tolerance()) {
        tag.error(std::fabs(rho * rho_0 / norm_rhs));
        ++k;
        break;
      }
    }
    for (int i2 = static_cast<int>(k) - 1; i2 > -1; --i2) {  // R is stored transposed
      const vcl_size_t i = static_cast<vcl_size_t>(i2);
      for (vcl_size_t j = i + 1; j < k; ++j) proj[i] -= R[j][i] * proj[j];
      proj[i] /= R[i][i];
    }
    res *= proj[0];
    if (k > 0)
      for (vcl_size_t i = 0; i + 1 < k; ++i) res[i] = res(i) + proj[i + 1];
    for (int i = static_cast<int>(k) - 1; i >= 0; --i) gmres_reflect(res, hh[vcl_size_t(i)], betas[vcl_size_t(i)]);
    res *= rho_0;
    result += res;
    tag.error(std::fabs(rho * rho_0 / norm_rhs));
    if (monitor && monitor(result, std::fabs(rho * rho_0 / norm_rhs), monitor_data)) break;
    if (tag.error() < tag.tolerance()) return result;
  }
  return result;
}

template <typename MatrixT, typename T, typename PreconditionerT>
typename std::enable_if<is_hip_sparse<MatrixT>::value, vector<T> >::type
solve_impl(MatrixT const& A, vector_base<T> const& rhs, gmres_tag const& tag, PreconditionerT const& precond,
           bool (*monitor)(vector<T> const&, T, void*) = nullptr, void* monitor_data = nullptr)
{
  if (std::is_same<PreconditionerT, no_precond>::value) return gmres_device(A, rhs, tag, monitor, monitor_data);
  vector<T> b(rhs);
  return gmres_generic(A, b, tag, precond, monitor, monitor_data);
}
/** A user operator (apply(x, y) and size1(); examples/tutorial/matrix-free.cpp:58-82): the classic
 *  recurrences on the HIP vector operations, each product through apply(). */
template <typename MatrixT, typename T, typename PreconditionerT>
typename std::enable_if<!is_hip_sparse<MatrixT>::value, vector<T> >::type
solve_impl(MatrixT const& A, vector_base<T> const& rhs, gmres_tag const& tag, PreconditionerT const& precond,
           bool (*monitor)(vector<T> const&, T, void*) = nullptr, void* monitor_data = nullptr)
{
  vector<T> b(rhs);
  return gmres_generic(A, b, tag, precond, monitor, monitor_data);
}
}  // namespace detail

/** solve(A, b, gmres_tag, precond) (gmres.hpp:630-634). */
template <typename MatrixT, typename VectorT, typename PreconditionerT>
vector<typename VectorT::value_type> solve(MatrixT const& matrix, VectorT const& rhs, gmres_tag const& tag,
                                           PreconditionerT const& precond)
{
  return detail::solve_impl(matrix, static_cast<vector_base<typename VectorT::value_type> const&>(rhs), tag, precond);
}

/** solve(A, b, gmres_tag) (gmres.hpp:664-668). */
template <typename MatrixT, typename VectorT>
vector<typename VectorT::value_type> solve(MatrixT const& matrix, VectorT const& rhs, gmres_tag const& tag)
{
  return solve(matrix, rhs, tag, no_precond());
}

/** STL convenience overload (gmres.hpp:641-655). */
template <typename IndexT, typename NumericT, typename PreconditionerT>
std::vector<NumericT> solve(std::vector<std::map<IndexT, NumericT> > const& A, std::vector<NumericT> const& rhs,
                            gmres_tag const& tag, PreconditionerT const& precond)
{
  compressed_matrix<NumericT> vcl_A;
  copy(A, vcl_A);
  vector<NumericT> vcl_rhs(rhs.size());
  copy(rhs, vcl_rhs);
  vector<NumericT> vcl_result = solve(vcl_A, vcl_rhs, tag, precond);
  std::vector<NumericT> result(vcl_result.size());
  copy(vcl_result, result);
  return result;
}

/** Solver object with an initial guess and a per-cycle monitor (gmres.hpp:672-733). */
template <typename VectorT>
class gmres_solver
{
public:
  typedef typename VectorT::value_type numeric_type;

  gmres_solver(gmres_tag const& tag) : tag_(tag), monitor_callback_(nullptr), user_data_(nullptr) {}

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
  gmres_tag const& tag() const { return tag_; }

private:
  gmres_tag tag_;
  VectorT init_guess_;
  bool (*monitor_callback_)(VectorT const&, numeric_type, void*);
  void* user_data_;
};

}  // namespace linalg
}  // namespace viennacl

#endif
