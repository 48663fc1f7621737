// Last-error storage shared by every C entry point of libnoptjk.
#pragma once
#include <string>

namespace nopt {
int set_error(int code, const std::string& msg);  // returns code
void clear_error();
}  // namespace nopt
