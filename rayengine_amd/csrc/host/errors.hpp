// Error reporting shared by every C-ABI entry point: a per-thread message for
// calls without a context, and a setter that returns the status it records.
#pragma once

#include <stdarg.h>
#include <stdio.h>

#include "../../../include/rayengine.h"

namespace re {

extern thread_local char g_last_error[512];

inline re_status fail(re_status st, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return st;
}

}  // namespace re
