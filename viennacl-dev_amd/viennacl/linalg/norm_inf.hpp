#ifndef VIENNACL_LINALG_NORM_INF_HPP
#define VIENNACL_LINALG_NORM_INF_HPP
// norm_inf.hpp -- include point of the reference's linalg/norm_inf.hpp; norm_inf and index_norm_inf are defined with the
// vector type of the HIP build.
#include "viennacl/vector.hpp"
#endif
