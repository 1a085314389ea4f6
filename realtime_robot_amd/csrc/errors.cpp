// The thread's last error message (rtr_last_error), shared by every part of
// the library, the host-only parts included.
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/rtr.h"

namespace rtr {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

}  // namespace rtr

extern "C" const char* rtr_last_error(void) { return rtr::g_err.c_str(); }
