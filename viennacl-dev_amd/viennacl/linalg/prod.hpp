#ifndef VIENNACL_LINALG_PROD_HPP
#define VIENNACL_LINALG_PROD_HPP
// prod.hpp -- prod() (reference linalg/prod.hpp).  For the device types, prod(A, x) is the lazy
// expression of linalg/sparse_matrix_operations.hpp (a sparse matrix or a user operator with
// apply()).  For STL containers it is the reference's serial host product, which its tests use as
// the ground truth (linalg/prod.hpp:97-131): each row summed in column order, result zeroed first.

#include <map>
#include <vector>

#include "viennacl/linalg/sparse_matrix_operations.hpp"
#include "viennacl/vector.hpp"

namespace viennacl
{
namespace linalg
{

/** Dense STL matrix times vector (linalg/prod.hpp:97-111). */
template <typename T, typename A1, typename A2, typename VectorT>
VectorT prod(std::vector<std::vector<T, A1>, A2> const& matrix, VectorT const& vec)
{
  VectorT result(matrix.size());
  for (typename std::vector<T, A1>::size_type i = 0; i < matrix.size(); ++i) {
    result[i] = 0;
    for (typename std::vector<T, A1>::size_type j = 0; j < matrix[i].size(); ++j) result[i] += matrix[i][j] * vec[j];
  }
  return result;
}

/** Sparse STL matrix (one std::map per row) times vector (linalg/prod.hpp:115-131). */
template <typename KeyT, typename DataT, typename CompareT, typename AllocMapT, typename AllocVecT, typename VectorT>
VectorT prod(std::vector<std::map<KeyT, DataT, CompareT, AllocMapT>, AllocVecT> const& matrix, VectorT const& vec)
{
  VectorT result(matrix.size());
  for (typename std::vector<std::map<KeyT, DataT, CompareT, AllocMapT>, AllocVecT>::size_type i = 0; i < matrix.size(); ++i) {
    result[i] = 0;
    for (typename std::map<KeyT, DataT, CompareT, AllocMapT>::const_iterator it = matrix[i].begin(); it != matrix[i].end(); ++it)
      result[i] += it->second * vec[it->first];
  }
  return result;
}

}  // namespace linalg
}  // namespace viennacl

#endif
