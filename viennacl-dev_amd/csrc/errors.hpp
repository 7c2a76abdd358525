// errors.hpp -- error reporting of libvclhip, free of the HIP runtime so that the host-only parts
// (generators, plans, format conversions, MatrixMarket files) also build with a host compiler.
#pragma once

#include <cstdarg>
#include <cstdio>

#include "vclhip.h"

namespace vclhip {

// Thread-local text of the last failure; returned by vclhip_last_error().
void set_error(const char* fmt, ...);
const char* last_error();

inline int invalid(const char* what)
{
  set_error("invalid argument: %s", what);
  return VCLHIP_ERROR_INVALID_VALUE;
}

#define VCLHIP_TRY(expr)              \
  do {                                \
    int vclhip_status_ = (expr);      \
    if (vclhip_status_ != 0)          \
      return vclhip_status_;          \
  } while (0)

}  // namespace vclhip
