#ifndef VIENNACL_FORWARDS_H
#define VIENNACL_FORWARDS_H
// forwards.h -- basic types, memory domains, tags and exceptions of the drop-in API.
// Counterpart of viennacl/forwards.h (reference forwards.h:369-375 memory_types, :596-625
// exceptions, :861-867 row_info_types) for the HIP build: one compute domain, HIP_MEMORY.

#include <cstddef>
#include <stdexcept>
#include <string>

#include "vclhip.h"

namespace viennacl
{

typedef std::size_t vcl_size_t;
typedef std::ptrdiff_t vcl_ptrdiff_t;

/** Memory domains (reference forwards.h:369-375 with OPENCL_MEMORY/CUDA_MEMORY replaced by HIP_MEMORY).
 *  MAIN_MEMORY holds host copies only; every computation runs in HIP_MEMORY. */
enum memory_types
{
  MEMORY_NOT_INITIALIZED,
  MAIN_MEMORY,
  HIP_MEMORY
};

/** Thrown when an operation is requested in a memory domain that cannot perform it (forwards.h:596-607). */
class memory_exception : public std::exception
{
public:
  memory_exception() : message_("ViennaCL: Internal memory error") {}
  memory_exception(std::string message) : message_("ViennaCL: Internal memory error: " + message) {}
  virtual const char* what() const throw() { return message_.c_str(); }
  virtual ~memory_exception() throw() {}
private:
  std::string message_;
};

/** Thrown when the Jacobi preconditioner meets a zero on the diagonal (forwards.h:621-625). */
class zero_on_diagonal_exception : public std::runtime_error
{
public:
  zero_on_diagonal_exception(std::string const& what_arg) : std::runtime_error(what_arg) {}
};

/** A failure reported by the HIP backend (counterpart of cuda_exception, backend/cuda.hpp:38). */
class hip_exception : public std::runtime_error
{
public:
  hip_exception(std::string const& what_arg, int status) : std::runtime_error(what_arg), status_(status) {}
  int status() const { return status_; }
private:
  int status_;
};

namespace backend
{
namespace hip
{
/** Turns a nonzero status of the C ABI into the exception the reference's callers expect. */
inline void check(int status, const char* where)
{
  if (status == VCLHIP_SUCCESS)
    return;
  std::string msg = std::string("ViennaCL (HIP): ") + where + ": " + vclhip_last_error();
  if (status == VCLHIP_ERROR_ZERO_DIAGONAL)
    throw zero_on_diagonal_exception(msg);
  throw hip_exception(msg, status);
}
}  // namespace hip
}  // namespace backend

template <typename NumericT> class vector_base;
template <typename NumericT, unsigned int AlignmentV = 1> class vector;
template <class NumericT, unsigned int AlignmentV = 1> class const_vector_iterator;
template <class NumericT, unsigned int AlignmentV = 1> class vector_iterator;
template <typename VectorType> class vector_range;
template <typename VectorType> class vector_slice;
template <typename NumericT, unsigned int AlignmentV = 1> class compressed_matrix;
template <typename ScalarT, typename IndexT = unsigned int> class sliced_ell_matrix;
template <typename NumericT, unsigned int AlignmentV = 128> class coordinate_matrix;
template <typename NumericT, unsigned int AlignmentV = 1> class ell_matrix;
template <typename NumericT, unsigned int AlignmentV = 1> class hyb_matrix;

/** Vectors are padded to a multiple of this many entries (reference forwards.h:385). */
static const vcl_size_t dense_padding_size = 128;

namespace linalg
{
class no_precond
{
public:
  template <typename VectorT>
  void apply(VectorT&) const {}
};

namespace detail
{
/** Selector of linalg::detail::row_info (forwards.h:861-867). */
enum row_info_types
{
  SPARSE_ROW_NORM_INF = 0,
  SPARSE_ROW_NORM_1,
  SPARSE_ROW_NORM_2,
  SPARSE_ROW_DIAGONAL
};
}  // namespace detail
}  // namespace linalg

namespace tools
{
/** Rounds up to a multiple of `base` (tools/tools.hpp:133-137). */
template <typename T>
T align_to_multiple(T to_reach, T base)
{
  if (to_reach % base == 0)
    return to_reach;
  return ((to_reach / base) + 1) * base;
}
}  // namespace tools

}  // namespace viennacl

#endif
