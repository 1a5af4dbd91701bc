// host_util.hpp -- small helpers of the host mirror: raw binary files and a
// device buffer owned by a context.
#pragma once

#include <cstddef>
#include <string>

#include "iac_hip.h"

namespace iac_host {

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
