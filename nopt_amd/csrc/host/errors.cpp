#include "errors.hpp"

namespace nopt {
static thread_local std::string g_last_error;
int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
void clear_error() { g_last_error.clear(); }
}  // namespace nopt

extern "C" const char* nopt_last_error(void) { return nopt::g_last_error.c_str(); }
