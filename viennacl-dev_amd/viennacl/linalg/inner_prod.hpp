#ifndef VIENNACL_LINALG_INNER_PROD_HPP
#define VIENNACL_LINALG_INNER_PROD_HPP
// inner_prod.hpp -- include point of the reference's linalg/inner_prod.hpp; the operation is defined with the
// vector and sparse-matrix types of the HIP build.
#include "viennacl/vector.hpp"
#include "viennacl/linalg/sparse_matrix_operations.hpp"
#endif
