// vh_error.h - the thread-local error text behind vh_last_error(), shared by the C ABI's
// translation units (defined in vh_capi.hip).
#pragma once
#include <string>

namespace vh {
// Stores msg as the calling thread's last error and returns code.
int set_last_error(int code, const std::string& msg);
}  // namespace vh
