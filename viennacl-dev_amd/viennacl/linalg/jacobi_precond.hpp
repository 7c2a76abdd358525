#ifndef VIENNACL_LINALG_JACOBI_PRECOND_HPP
#define VIENNACL_LINALG_JACOBI_PRECOND_HPP
// jacobi_precond.hpp -- Jacobi preconditioner (reference jacobi_precond.hpp:41-130).  For the
// CSR type the diagonal is extracted on the device with row_info(SPARSE_ROW_DIAGONAL); apply()
// is element_div.  The solvers recognize this class and fuse the scaling into their kernels.

#include "viennacl/compressed_matrix.hpp"
#include "viennacl/vector.hpp"

namespace viennacl
{
namespace linalg
{

class jacobi_tag {};

template <typename MatrixT>
class jacobi_precond
{
  typedef typename MatrixT::value_type NumericType;

public:
  jacobi_precond(MatrixT const& mat, jacobi_tag const&) : diag_A_(mat.size1()) { init(mat); }

  void init(MatrixT const& mat)
  {
    detail::row_info(mat, diag_A_, detail::SPARSE_ROW_DIAGONAL);
  }

  /** vec = vec ./ diag(A) (jacobi_precond.hpp:121-126). */
  template <typename VectorT>
  void apply(VectorT& vec) const
  {
    if (vec.size() != diag_A_.size()) throw std::invalid_argument("ViennaCL (HIP): Jacobi size mismatch");
    vec = element_div(vec, diag_A_);
  }

  vector<NumericType> const& diagonal() const { return diag_A_; }

private:
  vector<NumericType> diag_A_;
};

namespace detail
{
template <typename P> struct is_jacobi : std::false_type {};
template <typename M> struct is_jacobi<jacobi_precond<M> > : std::true_type {};
}  // namespace detail

}  // namespace linalg
}  // namespace viennacl

#endif
