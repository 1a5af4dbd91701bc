// host_util.hpp -- small helpers of the host mirror: raw binary files and a
// device buffer owned by a context.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "iac_hip.h"
#include "iac_host.hpp"


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

// Frames of events [first, first + count) read on a reader thread (Settings::LoadFrame, so a
// FrameLoader runs on that thread, one call at a time) into two page-locked stacks of up to
// `stack` frames: while the caller uploads and processes one stack, the next is read.  The
// stacks come from a process-wide pool of page-locked blocks, kept for reuse until exit.
class FrameStream {
public:
    FrameStream(const Settings &S, const std::vector<Settings::HitEvent> &events, size_t first, size_t count,
                size_t npix, uint32_t stack);
    ~FrameStream();
    FrameStream(const FrameStream &) = delete;
    FrameStream &operator=(const FrameStream &) = delete;
    // The next stack in event order (nullptr after the last): nb frames of npix floats,
    // valid until release() or the next call; rethrows a reader error.
    const float *next(uint32_t &nb);
    // The current stack's memory may be refilled (e.g. after its upload).
    void release();
    double read_seconds() const;  // reader busy time (file reads)
    double wait_seconds() const;  // time next() waited for the reader
    bool pinned() const { return pinned_; }

private:
    struct Impl;
    Impl *impl_;
    bool pinned_ = false;
    size_t bytes_ = 0;
};

// Timing of the last evaluation's frame passes (for the driver's report).
struct IngestStats {
    uint64_t frames = 0;        // frames streamed (each pass counts)
    double read_seconds = 0;    // reader thread busy reading
    double wait_seconds = 0;    // device loop waiting for frames
    double loop_seconds = 0;    // wall time of the streamed passes
    double upload_seconds = 0;  // host-to-device copies from the pinned stacks
};
IngestStats &ingest_stats();

}  // namespace iac_host
