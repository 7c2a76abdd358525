#ifndef VIENNACL_LINALG_NORM_1_HPP
#define VIENNACL_LINALG_NORM_1_HPP
// norm_1.hpp -- include point of the reference's linalg/norm_1.hpp; norm_1 are defined with the
// vector type of the HIP build.
#include "viennacl/vector.hpp"
#endif
