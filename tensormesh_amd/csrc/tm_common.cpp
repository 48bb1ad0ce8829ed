// Error reporting and version query of the C-ABI.
#include <cstdarg>
#include <cstdio>

#include "tm_internal.h"

namespace {
thread_local char g_last_error[512] = "";
}

namespace tmh {
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}
}  // namespace tmh

extern "C" int tm_abi_version(void) { return TM_ABI_VERSION; }

extern "C" const char* tm_last_error(void) { return g_last_error; }
