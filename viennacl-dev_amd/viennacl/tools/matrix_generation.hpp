#ifndef VIENNACL_TOOLS_MATRIX_GENERATION_HPP
#define VIENNACL_TOOLS_MATRIX_GENERATION_HPP
// matrix_generation.hpp -- tools::generate_fdm_laplace (reference tools/matrix_generation.hpp:47-86):
// the 2-D 5-point finite-difference Laplacian, row = i + j * points_x, diagonal 4, neighbours -1.
// The CSR arrays are assembled on the host by libvclhip and uploaded once.

#include <vector>

#include "viennacl/compressed_matrix.hpp"

namespace viennacl
{
namespace tools
{

template <typename NumericT, unsigned int AlignmentV>
void generate_fdm_laplace(viennacl::compressed_matrix<NumericT, AlignmentV>& A, vcl_size_t points_x,
                          vcl_size_t points_y)
{
  vcl_size_t n = points_x * points_y;
  std::vector<uint32_t> rp(n + 1), col(5 * n + 1);
  std::vector<double> val(5 * n + 1);
  uint32_t nnz = 0;
  viennacl::backend::hip::check(vclhip_gen_fdm_laplace_2d(uint32_t(points_x), uint32_t(points_y), rp.data(), col.data(),
                                                          val.data(), &nnz), "generate_fdm_laplace");
  std::vector<NumericT> v(val.begin(), val.begin() + nnz);
  A.set(rp.data(), col.data(), v.data(), n, n, nnz);
}

}  // namespace tools
}  // namespace viennacl

#endif
