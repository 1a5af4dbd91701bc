// framestream.cpp -- FrameStream (host_util.hpp): frame stacks read ahead on a reader thread
// into page-locked memory, so file reads overlap the uploads and device passes of the
// previous stack (the reference reads each event's frame inside its frame loop,
// IAC/RunIAC.cpp:524-550, IAC/Detector.cpp:700-760).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>
#include <utility>

#include "host_util.hpp"

namespace iac_host {

namespace {

// Page-locked stacks are kept for reuse until the process ends: the first CPU write to fresh
// page-locked memory runs at a few GB/s (the pages are mapped then), reuse at memory speed, so
// every pass after the first reads at file speed.  A new block is touched once by up to 16
// threads (the first-touch rate scales with them).
struct PinnedPool {
    std::mutex m;
    std::vector<std::pair<size_t, float *>> idle;
};

PinnedPool &pool()
{
    static PinnedPool *p = new PinnedPool();  // never destroyed: its blocks live until exit
    return *p;
}

float *pool_take(size_t bytes)
{
    {
        std::lock_guard<std::mutex> g(pool().m);
        auto &v = pool().idle;
        size_t best = v.size();
        for (size_t i = 0; i < v.size(); i++)
            if (v[i].first >= bytes && (best == v.size() || v[i].first < v[best].first)) best = i;
        if (best < v.size()) {
            float *b = v[best].second;
            v.erase(v.begin() + (long)best);
            return b;
        }
    }
    void *p = nullptr;
    if (iac_pinned_alloc(bytes, &p) != IAC_OK) return nullptr;
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t chunk = (bytes + hw - 1) / hw;
    std::vector<std::thread> ts;
    for (unsigned i = 0; i < hw; i++) {
        const size_t a = std::min(bytes, i * chunk), e = std::min(bytes, a + chunk);
        if (e > a) ts.emplace_back([=] { std::memset(static_cast<char *>(p) + a, 0, e - a); });
    }
    for (auto &t : ts) t.join();
    return static_cast<float *>(p);
}

void pool_give(float *b, size_t bytes)
{
    std::lock_guard<std::mutex> g(pool().m);
    pool().idle.emplace_back(bytes, b);
}

}  // namespace

struct FrameStream::Impl {
    const Settings &S;
    const std::vector<Settings::HitEvent> &events;
    size_t first, count, npix;
    uint32_t stack;
    float *buf[2] = {nullptr, nullptr};
    std::vector<float> fallback[2];  // pageable stacks when page-locked memory is not available
    // slot state: 0 free, 1 being read, 2 full (nb frames), 3 held by the consumer
    int state[2] = {0, 0};
    uint32_t nb[2] = {0, 0};
    size_t next_stack = 0, nstacks = 0;  // reader: the next stack to read
    size_t take = 0;                     // consumer: the next stack to hand out
    int held = -1;
    bool stop = false;
    std::exception_ptr error;
    double read_s = 0, wait_s = 0;
    std::mutex m;
    std::condition_variable cv;
    std::thread reader;

    Impl(const Settings &s, const std::vector<Settings::HitEvent> &ev, size_t f, size_t c, size_t p, uint32_t b)
        : S(s), events(ev), first(f), count(c), npix(p), stack(b)
    {
        nstacks = (count + stack - 1) / stack;
    }

    void run()
    {
        for (size_t k = 0; k < nstacks; k++) {
            const int slot = (int)(k % 2);
            {
                std::unique_lock<std::mutex> g(m);
                cv.wait(g, [&] { return stop || state[slot] == 0; });
                if (stop) return;
                state[slot] = 1;
            }
            const auto t0 = std::chrono::steady_clock::now();
            const size_t f0 = first + k * stack;
            const uint32_t n = (uint32_t)std::min<size_t>(stack, first + count - f0);
            try {
                for (uint32_t f = 0; f < n; f++) S.LoadFrame(events[f0 + f], buf[slot] + (size_t)f * npix, npix);
            } catch (...) {
                std::lock_guard<std::mutex> g(m);
                error = std::current_exception();
                state[slot] = 2;
                nb[slot] = 0;
                cv.notify_all();
                return;
            }
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::lock_guard<std::mutex> g(m);
            read_s += dt;
            nb[slot] = n;
            state[slot] = 2;
            cv.notify_all();
        }
    }
};

FrameStream::FrameStream(const Settings &S, const std::vector<Settings::HitEvent> &events, size_t first,
                         size_t count, size_t npix, uint32_t stack)
{
    if (stack == 0 || npix == 0) throw std::runtime_error("FrameStream: empty stacks");
    if (first + count > events.size()) throw std::runtime_error("FrameStream: event range past the list");
    impl_ = new Impl(S, events, first, count, npix, stack);
    bytes_ = (size_t)stack * npix * sizeof(float);
    pinned_ = true;
    for (int i = 0; i < 2; i++) {
        float *p = pinned_ ? pool_take(bytes_) : nullptr;
        if (p) {
            impl_->buf[i] = p;
        } else {
            if (i == 1 && impl_->buf[0]) {  // keep both stacks of one kind
                pool_give(impl_->buf[0], bytes_);
                impl_->fallback[0].resize((size_t)stack * npix);
                impl_->buf[0] = impl_->fallback[0].data();
            }
            pinned_ = false;
            impl_->fallback[i].resize((size_t)stack * npix);
            impl_->buf[i] = impl_->fallback[i].data();
        }
    }
    impl_->reader = std::thread([this] { impl_->run(); });
}

FrameStream::~FrameStream()
{
    {
        std::lock_guard<std::mutex> g(impl_->m);
        impl_->stop = true;
    }
    impl_->cv.notify_all();
    if (impl_->reader.joinable()) impl_->reader.join();
    if (pinned_)
        for (float *b : impl_->buf) pool_give(b, bytes_);
    delete impl_;
}

void FrameStream::release()
{
    std::lock_guard<std::mutex> g(impl_->m);
    if (impl_->held >= 0) {
        impl_->state[impl_->held] = 0;
        impl_->held = -1;
        impl_->cv.notify_all();
    }
}

const float *FrameStream::next(uint32_t &nb)
{
    release();
    nb = 0;
    std::unique_lock<std::mutex> g(impl_->m);
    if (impl_->take >= impl_->nstacks) return nullptr;
    const int slot = (int)(impl_->take % 2);
    const auto t0 = std::chrono::steady_clock::now();
    impl_->cv.wait(g, [&] { return impl_->state[slot] == 2; });
    impl_->wait_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (impl_->error) std::rethrow_exception(impl_->error);
    impl_->state[slot] = 3;
    impl_->held = slot;
    impl_->take++;
    nb = impl_->nb[slot];
    return impl_->buf[slot];
}

double FrameStream::read_seconds() const
{
    std::lock_guard<std::mutex> g(impl_->m);
    return impl_->read_s;
}

double FrameStream::wait_seconds() const
{
    std::lock_guard<std::mutex> g(impl_->m);
    return impl_->wait_s;
}

IngestStats &ingest_stats()
{
    static IngestStats s;
    return s;
}

}  // namespace iac_host
