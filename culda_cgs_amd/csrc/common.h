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

// A file written as path + ".tmp" and renamed over path once every write
// and the close have succeeded: a failed save (full disk, I/O error) is
// reported as CULDA_ERR_IO and never leaves a truncated file in place of a
// good one.
class OutFile {
public:
    OutFile(const std::string &path, const char *mode) : path_(path), tmp_(path + ".tmp") {
        f_ = std::fopen(tmp_.c_str(), mode);
        if (!f_) fail(CULDA_ERR_IO, "%s: open for writing failed", tmp_.c_str());
    }
    OutFile(const OutFile &) = delete;
    OutFile &operator=(const OutFile &) = delete;
    ~OutFile() {
        if (f_) {
            std::fclose(f_);
            std::remove(tmp_.c_str());
        }
    }
    void write(const void *p, size_t size, size_t n) {
        if (n && std::fwrite(p, size, n, f_) != n) fail(CULDA_ERR_IO, "%s: write failed", tmp_.c_str());
    }
    void printf(const char *fmt, ...) __attribute__((format(printf, 2, 3))) {
        va_list ap;
        va_start(ap, fmt);
        const int r = std::vfprintf(f_, fmt, ap);
        va_end(ap);
        if (r < 0) fail(CULDA_ERR_IO, "%s: write failed", tmp_.c_str());
    }
    void commit() {
        FILE *f = f_;
        f_ = nullptr;
        if (std::fclose(f) != 0) {
            std::remove(tmp_.c_str());
            fail(CULDA_ERR_IO, "%s: close failed", tmp_.c_str());
        }
        if (std::rename(tmp_.c_str(), path_.c_str()) != 0) {
            std::remove(tmp_.c_str());
            fail(CULDA_ERR_IO, "%s: rename to %s failed", tmp_.c_str(), path_.c_str());
        }
    }

private:
    std::string path_, tmp_;
    FILE *f_ = nullptr;
};

}  // namespace culda
