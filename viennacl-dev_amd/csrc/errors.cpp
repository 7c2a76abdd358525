// errors.cpp -- the thread-local text of the last failure (vclhip_last_error).
#include "errors.hpp"

namespace vclhip {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_last_error; }

}  // namespace vclhip

extern "C" const char* vclhip_last_error(void) { return vclhip::last_error(); }
