// Thread-local last-error storage behind mrad_last_error().  Status codes and
// their Java exception classes are listed in include/mrad.h.
#pragma once
#include <string>

namespace mrad {

int set_error(int code, const std::string &message);
void clear_error();
const char *last_error();

}  // namespace mrad
