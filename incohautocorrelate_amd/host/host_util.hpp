// host_util.hpp -- small helpers of the host mirror: raw binary files and a
// device buffer owned by a context.
#pragma once

#include <cstddef>
#include <string>

#include "iac_hip.h"

class Settings;

namespace iac_host {

// A device of a Settings pool held for one scope: OCL_ReserveDevice at entry and
// OCL_FreeDevice on every exit, exceptions included (IAC/Settings.cpp:432-470).
// Throws std::runtime_error("<who>: no free device") when the pool is exhausted.
class Reserved {
public:
    Reserved(Settings &pool, const char *who);
    ~Reserved();
    Reserved(const Reserved &) = delete;
    Reserved &operator=(const Reserved &) = delete;
    int dev = -1;
    iac_ctx *ctx = nullptr;

private:
    Settings &pool_;
};

void read_raw(const std::string &path, void *dst, size_t bytes, size_t offset = 0);
void write_raw(const std::string &path, const void *src, size_t bytes);

// Device allocation on ctx, freed on destruction; errors exit like Settings::checkErr.
class DevBuf {
public:
    DevBuf() = default;
    DevBuf(iac_ctx *ctx, size_t bytes) { reset(ctx, bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    void reset(iac_ctx *ctx, size_t bytes);
    void release();
    void zero();
    void upload(const void *src, size_t bytes);
    void download(void *dst, size_t bytes) const;
    template <typename T>
    T *as() const { return static_cast<T *>(ptr_); }
    size_t size() const { return bytes_; }

private:
    iac_ctx *ctx_ = nullptr;
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
};

// exit(EXIT_FAILURE) with the library's message, as Settings::checkErr does
void check(int rc, const char *where);

}  // namespace iac_host
