#ifndef VIENNACL_LINALG_DETAIL_SOLVER_GLUE_HPP
#define VIENNACL_LINALG_DETAIL_SOLVER_GLUE_HPP
// solver_glue.hpp -- shared plumbing between the solver tags / solve() overloads and the C ABI's
// device-resident solver drivers: parameter packing, the monitor trampoline, and dispatch by value
// type and matrix format.

#include "viennacl/compressed_matrix.hpp"
#include "viennacl/coordinate_matrix.hpp"
#include "viennacl/ell_matrix.hpp"
#include "viennacl/hyb_matrix.hpp"
#include "viennacl/sliced_ell_matrix.hpp"
#include "viennacl/vector.hpp"

namespace viennacl
{
namespace linalg
{
namespace detail
{

/** Adapts the reference's monitor signature bool(VectorT const&, T, void*) (cg.hpp:413-424) to the C ABI. */
template <typename VectorT>
struct monitor_bridge
{
  typedef typename VectorT::value_type T;
  bool (*fn)(VectorT const&, T, void*);
  void* user;
  vcl_size_t n;

  static int call(const void* x_dev, double rel, void* self)
  {
    monitor_bridge* b = static_cast<monitor_bridge*>(self);
    VectorT x(const_cast<T*>(static_cast<const T*>(x_dev)), HIP_MEMORY, b->n);
    return b->fn(x, T(rel), b->user) ? 1 : 0;
  }
};

inline vclhip_solver_params default_params()
{
  vclhip_solver_params p;
  vclhip_solver_params_default(&p);
  return p;
}

/** The DIA copy a CSR matrix keeps for its solves (compressed_matrix::dia_cache); none for the
 *  other formats. */
template <typename MatrixT>
vclhip_dia const* dia_of(MatrixT const&) { return nullptr; }
template <typename T, unsigned A>
vclhip_dia const* dia_of(compressed_matrix<T, A> const& m) { return m.dia_cache(); }

// ---- dispatch on value type and format
inline int cg_solve(vclhip_csr const& A, const double* b, double* x, int pc, vclhip_solver_params const& P,
                    vclhip_solver_result& R)
{ return vclhip_cg_solve_csr_f64(&A, b, x, pc, &P, &R, backend::hip::current_stream()); }
inline int cg_solve(vclhip_csr const& A, const float* b, float* x, int pc, vclhip_solver_params const& P,
                    vclhip_solver_result& R)
{ return vclhip_cg_solve_csr_f32(&A, b, x, pc, &P, &R, backend::hip::current_stream()); }
inline int cg_solve(vclhip_sell const& A, const double* b, double* x, int pc, vclhip_solver_params const& P,
                    vclhip_solver_result& R)
{ return vclhip_cg_solve_sell_f64(&A, b, x, pc, &P, &R, backend::hip::current_stream()); }
inline int cg_solve(vclhip_sell const& A, const float* b, float* x, int pc, vclhip_solver_params const& P,
                    vclhip_solver_result& R)
{ return vclhip_cg_solve_sell_f32(&A, b, x, pc, &P, &R, backend::hip::current_stream()); }

inline int bicgstab_solve(vclhip_csr const& A, const double* b, double* x, int pc, vclhip_solver_params const& P,
                          vclhip_solver_result& R)
{ return vclhip_bicgstab_solve_csr_f64(&A, b, x, pc, &P, &R, backend::hip::current_stream()); }
inline int bicgstab_solve(vclhip_csr const& A, const float* b, float* x, int pc, vclhip_solver_params const& P,
                          vclhip_solver_result& R)
{ return vclhip_bicgstab_solve_csr_f32(&A, b, x, pc, &P, &R, backend::hip::current_stream()); }
inline int bicgstab_solve(vclhip_sell const& A, const double* b, double* x, int pc, vclhip_solver_params const& P,
                          vclhip_solver_result& R)
{ return vclhip_bicgstab_solve_sell_f64(&A, b, x, pc, &P, &R, backend::hip::current_stream()); }
inline int bicgstab_solve(vclhip_sell const& A, const float* b, float* x, int pc, vclhip_solver_params const& P,
                          vclhip_solver_result& R)
{ return vclhip_bicgstab_solve_sell_f32(&A, b, x, pc, &P, &R, backend::hip::current_stream()); }

// COO / ELL / HYB go through the format-generic drivers
template <typename DescT>
inline int cg_solve(DescT const& A, const double* b, double* x, int pc, vclhip_solver_params const& P,
                    vclhip_solver_result& R)
{
  vclhip_matrix m = hip::generic(A);
  return vclhip_cg_solve_f64(&m, b, x, pc, &P, &R, backend::hip::current_stream());
}
template <typename DescT>
inline int cg_solve(DescT const& A, const float* b, float* x, int pc, vclhip_solver_params const& P,
                    vclhip_solver_result& R)
{
  vclhip_matrix m = hip::generic(A);
  return vclhip_cg_solve_f32(&m, b, x, pc, &P, &R, backend::hip::current_stream());
}
template <typename DescT>
inline int bicgstab_solve(DescT const& A, const double* b, double* x, int pc, vclhip_solver_params const& P,
                          vclhip_solver_result& R)
{
  vclhip_matrix m = hip::generic(A);
  return vclhip_bicgstab_solve_f64(&m, b, x, pc, &P, &R, backend::hip::current_stream());
}
template <typename DescT>
inline int bicgstab_solve(DescT const& A, const float* b, float* x, int pc, vclhip_solver_params const& P,
                          vclhip_solver_result& R)
{
  vclhip_matrix m = hip::generic(A);
  return vclhip_bicgstab_solve_f32(&m, b, x, pc, &P, &R, backend::hip::current_stream());
}

/** Contiguous copy of b (the drivers take unit-stride vectors). */
template <typename T>
vector<T> contiguous_copy(vector_base<T> const& b)
{
  vector<T> c(b.size());
  c.assign_linear(b, T(1));
  return c;
}

template <typename T> struct is_hip_sparse : std::false_type {};
template <typename T, unsigned A> struct is_hip_sparse<compressed_matrix<T, A> > : std::true_type {};
template <typename T, typename I> struct is_hip_sparse<sliced_ell_matrix<T, I> > : std::true_type {};
template <typename T, unsigned A> struct is_hip_sparse<coordinate_matrix<T, A> > : std::true_type {};
template <typename T, unsigned A> struct is_hip_sparse<ell_matrix<T, A> > : std::true_type {};
template <typename T, unsigned A> struct is_hip_sparse<hyb_matrix<T, A> > : std::true_type {};

}  // namespace detail
}  // namespace linalg
}  // namespace viennacl

#endif
