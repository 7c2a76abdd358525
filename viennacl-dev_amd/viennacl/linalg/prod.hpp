#ifndef VIENNACL_LINALG_PROD_HPP
#define VIENNACL_LINALG_PROD_HPP
// prod.hpp -- include point of the reference's linalg/prod.hpp; the operation is defined with the
// vector and sparse-matrix types of the HIP build.
#include "viennacl/vector.hpp"
#include "viennacl/linalg/sparse_matrix_operations.hpp"
#endif
