#ifndef VIENNACL_LINALG_MIXED_PRECISION_CG_HPP
#define VIENNACL_LINALG_MIXED_PRECISION_CG_HPP
// mixed_precision_cg.hpp -- mixed-precision CG (reference linalg/mixed_precision_cg.hpp).  The inner
// iteration runs in FP32 on the device (SpMV, vector updates, Jacobi scaling; inner products
// accumulated in FP64); FP64 is used for x += x32, r = b - A x and the stopping norm.  By default
// the reference's schedule is kept (restart of the Krylov space whenever the FP32 <r,r> has
// fallen by inner_tolerance, mixed_precision_cg.hpp:161).  reliable_updates(true) keeps the search
// direction across the FP64 corrections instead, which needs about as many iterations as FP64 CG.

#include "viennacl/linalg/detail/solver_glue.hpp"
#include "viennacl/linalg/jacobi_precond.hpp"

namespace viennacl
{
namespace linalg
{

/** Solver parameters and results (mixed_precision_cg.hpp:47-86). */
class mixed_precision_cg_tag
{
public:
  mixed_precision_cg_tag(double tol = 1e-8, unsigned int max_iterations = 300, float inner_tol = 1e-2f)
    : tol_(tol), iterations_(max_iterations), inner_tol_(inner_tol), reliable_(false), iters_taken_(0),
      last_error_(0), corrections_(0) {}
  double tolerance() const { return tol_; }
  float inner_tolerance() const { return inner_tol_; }
  unsigned int max_iterations() const { return iterations_; }
  unsigned int iters() const { return iters_taken_; }
  void iters(unsigned int i) const { iters_taken_ = i; }
  double error() const { return last_error_; }
  void error(double e) const { last_error_ = e; }

  /** Keeps the search direction across FP64 corrections (no reference counterpart). */
  bool reliable_updates() const { return reliable_; }
  void reliable_updates(bool on) { reliable_ = on; }
  /** Number of FP64 corrections of the last solve. */
  unsigned int corrections() const { return corrections_; }
  void corrections(unsigned int c) const { corrections_ = c; }

private:
  double tol_;
  unsigned int iterations_;
  float inner_tol_;
  bool reliable_;
  mutable unsigned int iters_taken_;
  mutable double last_error_;
  mutable unsigned int corrections_;
};

namespace detail
{
inline int mixed_solve(vclhip_csr const& A, const double* b, double* x, int pc, vclhip_solver_params const& P,
                       vclhip_solver_result& R)
{ return vclhip_mixed_cg_solve_csr(&A, nullptr, b, x, pc, &P, &R, backend::hip::current_stream()); }
inline int mixed_solve(vclhip_sell const& A, const double* b, double* x, int pc, vclhip_solver_params const& P,
                       vclhip_solver_result& R)
{ return vclhip_mixed_cg_solve_sell(&A, nullptr, b, x, pc, &P, &R, backend::hip::current_stream()); }

template <typename MatrixT>
vector<double> mixed_cg_device(MatrixT const& A, vector_base<double> const& rhs, mixed_precision_cg_tag const& tag,
                               int precond)
{
  static_assert(std::is_same<typename MatrixT::value_type, double>::value,
                "mixed_precision_cg solves FP64 systems (mixed_precision_cg.hpp:98-100)");
  vector<double> b = contiguous_copy(rhs);
  vector<double> x(rhs.size());
  vclhip_solver_params P = default_params();
  P.tol = tag.tolerance();
  P.max_iters = tag.max_iterations();
  P.inner_tol = tag.inner_tolerance();
  P.mixed_mode = tag.reliable_updates() ? 1 : 0;
  vclhip_solver_result R;
  if (rhs.size()) {
    backend::hip::check(mixed_solve(A.descriptor(), b.device_ptr(), x.device_ptr(), precond, P, R),
                        "mixed_precision_cg solve");
    tag.iters(R.iters);
    tag.error(R.error);
    tag.corrections(R.restarts);
  }
  return x;
}
}  // namespace detail

/** solve(A, b, mixed_precision_cg_tag) (mixed_precision_cg.hpp:95-187). */
template <typename MatrixT, typename VectorT>
typename std::enable_if<detail::is_hip_sparse<MatrixT>::value, vector<double> >::type
solve(MatrixT const& matrix, VectorT const& rhs, mixed_precision_cg_tag const& tag)
{
  return detail::mixed_cg_device(matrix, static_cast<vector_base<double> const&>(rhs), tag, VCLHIP_PRECOND_NONE);
}

/** solve(A, b, mixed_precision_cg_tag, no_precond) (mixed_precision_cg.hpp:189-193). */
template <typename MatrixT, typename VectorT>
typename std::enable_if<detail::is_hip_sparse<MatrixT>::value, vector<double> >::type
solve(MatrixT const& matrix, VectorT const& rhs, mixed_precision_cg_tag const& tag, no_precond)
{
  return solve(matrix, rhs, tag);
}

/** Jacobi-preconditioned mixed CG (FP32 scaling inside; needs reliable_updates). */
template <typename MatrixT, typename VectorT, typename M2>
typename std::enable_if<detail::is_hip_sparse<MatrixT>::value, vector<double> >::type
solve(MatrixT const& matrix, VectorT const& rhs, mixed_precision_cg_tag const& tag, jacobi_precond<M2> const&)
{
  if (!tag.reliable_updates())
    throw std::invalid_argument("ViennaCL (HIP): the reference schedule of mixed_precision_cg has no preconditioner; "
                                "set reliable_updates(true)");
  return detail::mixed_cg_device(matrix, static_cast<vector_base<double> const&>(rhs), tag, VCLHIP_PRECOND_JACOBI);
}

}  // namespace linalg
}  // namespace viennacl

#endif
