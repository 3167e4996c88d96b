// Library-wide state: version and per-thread error messages.
#include "kh_runtime.hpp"

namespace kh {

namespace {
thread_local char g_last_error[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

void clear_error() { g_last_error[0] = '\0'; }

}  // namespace kh

extern "C" const char* kh_version(void) { return "kennett_hip 0.1.0 (gfx950)"; }

extern "C" const char* kh_last_error(void) { return kh::g_last_error; }
