// errors.cpp -- the library's error state (host only, no HIP): internal failures
// are thrown as hcfd::Error and turned into a status code at the C boundary,
// with the message kept per thread for hcfd_last_error (include/hcfd.h).
#include <string>

#include "common.hpp"

namespace hcfd {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

[[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }

}  // namespace hcfd

extern "C" const char* hcfd_last_error(void) { return hcfd::g_last_error.c_str(); }
