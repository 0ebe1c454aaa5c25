// Shared helpers of the ccd_amd library (host side).
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

namespace ccd {

// Records the calling thread's last error message; returns `code`.
int fail(int code, const char *fmt, ...);
void clear_error();

}  // namespace ccd
