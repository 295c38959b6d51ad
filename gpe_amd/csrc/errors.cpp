// errors.cpp -- the library's error reporting: every failing entry point
// records its message here (gpe_last_error reads it back, per thread) and
// returns its GPE_E* code.  Plain C++ with no device code, so host-only
// builds of the parsers (the sanitizer tests) link it too.
#include <cstdarg>
#include <cstdio>
#include <string>

namespace gpe {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace gpe
