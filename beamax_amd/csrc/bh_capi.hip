// Error state and version of the C-ABI.
#include <stdarg.h>

#include "bh_internal.hpp"

namespace {
thread_local char g_err[1024] = {0};
}

namespace bh {
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
void clear_error() { g_err[0] = 0; }
}  // namespace bh

extern "C" const char* bh_last_error(void) { return g_err; }
extern "C" int bh_version(void) { return 1; }
