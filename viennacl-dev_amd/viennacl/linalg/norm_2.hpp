#ifndef VIENNACL_LINALG_NORM_2_HPP
#define VIENNACL_LINALG_NORM_2_HPP
// norm_2.hpp -- include point of the reference's linalg/norm_2.hpp; the operation is defined with the
// vector and sparse-matrix types of the HIP build.
#include "viennacl/vector.hpp"
#include "viennacl/linalg/sparse_matrix_operations.hpp"
#endif
