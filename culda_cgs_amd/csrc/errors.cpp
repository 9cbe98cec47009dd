// errors.cpp -- thread-local last-error storage for the C ABI.
#include <cstring>

#include "common.h"

namespace culda {

static thread_local char g_last_error[1024];

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
    va_end(ap);
}

void clear_error() { g_last_error[0] = 0; }

void fail(int status, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Error(status, buf);
}

}  // namespace culda

extern "C" const char *culda_last_error(void) { return culda::g_last_error; }
extern "C" int culda_abi_version(void) { return CULDA_ABI_VERSION; }
