#ifndef VIENNACL_SCALAR_HPP
#define VIENNACL_SCALAR_HPP
// scalar.hpp -- include point of the reference's viennacl/scalar.hpp, which the tutorials include
// first.  On this path inner products and norms return host values (T), so no device scalar type
// is defined; scalar_vector is in vector.hpp.
#include "viennacl/forwards.h"
#include "viennacl/vector.hpp"
#endif
