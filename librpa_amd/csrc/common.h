// Shared helpers for the MI355X (gfx950) implementation of LibRPA's space-time
// chi0 -> RPA correlation energy path: HIP error checks, device buffers and
// the event-based profiler.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace lra {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define LRA_THROW(msg)                                                                  \
    throw ::lra::Error(std::string(__FILE__) + ":" + std::to_string(__LINE__) + ": " + (msg))

#define HIP_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            LRA_THROW(std::string("HIP error ") + hipGetErrorString(e_) + " in " #expr); \
    } while (0)

// Owning device allocation.
class DevBuf {
  public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void alloc(size_t bytes) {
        release();
        if (bytes == 0) return;
        HIP_CHECK(hipMalloc(&p_, bytes));
        n_ = bytes;
    }
    // grow-only reallocation (contents not preserved)
    void reserve(size_t bytes) { if (bytes > n_) alloc(bytes); }
    void release() {
        if (p_) { (void)hipFree(p_); p_ = nullptr; n_ = 0; }
    }
    template <class T> T *as() const { return reinterpret_cast<T *>(p_); }
    void *get() const { return p_; }
    size_t bytes() const { return n_; }

  private:
    void *p_ = nullptr;
    size_t n_ = 0;
};

template <class T>
DevBuf upload(const std::vector<T> &v, hipStream_t s = 0) {
    DevBuf b(v.size() * sizeof(T));
    if (!v.empty()) HIP_CHECK(hipMemcpyAsync(b.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return b;
}

// Named kernel-region timers built on hipEvents recorded on the work stream.
class Profiler {
  public:
    struct Rec { double ms = 0; long calls = 0; double flops = 0; double bytes = 0; };
    void enable(bool on) { on_ = on; }
    bool enabled() const { return on_; }
    void begin(const std::string &name, hipStream_t s);
    void end(const std::string &name, hipStream_t s, double flops = 0, double bytes = 0);
    void flush();  // synchronises and folds pending event pairs into totals
    void clear() { flush(); recs_.clear(); }
    const std::map<std::string, Rec> &records() { flush(); return recs_; }
    std::string report();

  private:
    struct Pending { std::string name; hipEvent_t a, b; double flops, bytes; };
    bool on_ = true;
    std::map<std::string, hipEvent_t> open_;
    std::vector<Pending> pending_;
    std::map<std::string, Rec> recs_;
};

Profiler &profiler();

inline int ceil_div(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace lra
