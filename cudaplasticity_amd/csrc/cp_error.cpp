// cp_error.cpp -- the calling thread's last error message of the C-ABI
// (cp_last_error) and the helpers that record it.  Host C++ only, shared by
// every translation unit of the library.
#include <cstdarg>
#include <cstdio>
#include <string>

#include "cp_abi.h"
#include "cp_internal.h"

namespace {
thread_local std::string g_err;
}

namespace cp {

int fail(int code, const char* msg)
{
    g_err = msg;
    return code;
}

int failf(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace cp

extern "C" const char* cp_last_error(void) { return g_err.c_str(); }
