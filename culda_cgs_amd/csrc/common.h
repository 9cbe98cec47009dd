// common.h -- status codes and the thread-local last-error message shared by
// the host library.  The reference aborts the process on every failure
// (culda_argument.h:40-48 gpuAssert, doc.cpp:68-72 exit(0)); here every
// failure becomes a culda_status plus a message.
#pragma once
#include <cstdarg>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/culda.h"

namespace culda {

void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

// Exception carrying a status; converted at the C-ABI edge.
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] void fail(int status, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// Run f, turning exceptions into a status and the last-error message.
template <class F>
int guarded(F &&f) {
    try {
        clear_error();
        return f();
    } catch (const Error &e) {
        set_error("%s", e.what());
        return e.status;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return CULDA_ERR_MEMORY;
    } catch (const std::exception &e) {
        set_error("%s", e.what());
        return CULDA_ERR_ARG;
    }
}

}  // namespace culda
