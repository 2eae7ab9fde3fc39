// Library-wide C-ABI entry points: version and per-thread error message.
#include <stdarg.h>
#include <stdio.h>

#include "smk_internal.h"

namespace {
thread_local char g_last_error[1024] = "";
}

namespace smk {
int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}
}  // namespace smk

extern "C" int smk_abi_version(void) { return SMK_ABI_VERSION; }
extern "C" const char* smk_last_error(void) { return g_last_error; }
