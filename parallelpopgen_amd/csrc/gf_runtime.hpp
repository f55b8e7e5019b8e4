/*
 * gf_runtime.hpp -- host-side helpers shared by the library's translation
 * units: error reporting, device buffers, kernel timing.
 */
#ifndef GF_RUNTIME_HPP_
#define GF_RUNTIME_HPP_

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>
#include <thread>
#include <cstring>
#include <string>
#include <vector>
#include <chrono>
#include <algorithm>
#include "../../include/gofish_capi.h"

namespace gf {

void set_error(const std::string &msg);
const std::string &get_error();
int timing_stride();   /* 0: kernel timing off; s: every s-th launch timed */
int generations_per_launch(int dflt);

struct Status {
    int code = GF_OK;
    explicit operator bool() const { return code != GF_OK; }
};

#define GF_HIP(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::gf::set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " +   \
                            __FILE__ + ":" + std::to_string(__LINE__) + ": " #expr);       \
            return GF_ERR_DEVICE;                                                          \
        }                                                                                  \
    } while (0)

#define GF_TRY(expr)                     \
    do {                                 \
        int rc_ = (expr);                \
        if (rc_ != GF_OK) return rc_;    \
    } while (0)

/* owning device buffer */
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        if (count <= n && p) return GF_OK;
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            set_error(std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
            return GF_ERR_ALLOC;
        }
        n = count;
        return GF_OK;
    }
    void swap(DevBuf &o) { std::swap(p, o.p); std::swap(n, o.n); }
};

/* pinned host buffer */
template <typename T>
struct PinBuf {
    T *p = nullptr;
    size_t n = 0;
    PinBuf() = default;
    PinBuf(const PinBuf &) = delete;
    PinBuf &operator=(const PinBuf &) = delete;
    ~PinBuf() { if (p) (void)hipHostFree(p); }
    int alloc(size_t count, unsigned int flags = hipHostMallocDefault) {
        if (count <= n && p) return GF_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        if (count == 0) count = 1;
        hipError_t e = hipHostMalloc((void **)&p, count * sizeof(T), flags);
        if (e != hipSuccess) { p = nullptr; set_error("hipHostMalloc failed"); return GF_ERR_ALLOC; }
        n = count;
        return GF_OK;
    }
};

/* host memcpy split over a few threads (one core copies ~10 GB/s) */
inline void parallel_copy(char *dst, const char *src, size_t bytes) {
    const size_t min_part = (size_t)2 << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t parts = std::min<size_t>(std::min<size_t>(8, hw), std::max<size_t>(1, bytes / min_part));
    if (parts <= 1) { std::memcpy(dst, src, bytes); return; }
    const size_t step = (bytes + parts - 1) / parts;
    std::vector<std::thread> th;
    for (size_t k = 1; k < parts; k++) {
        const size_t off = k * step;
        if (off >= bytes) break;
        th.emplace_back([=] { std::memcpy(dst + off, src + off, std::min(step, bytes - off)); });
    }
    std::memcpy(dst, src, std::min(step, bytes));
    for (auto &t : th) t.join();
}

/* Device -> host copy through a process-wide pinned staging buffer (kept
 * between calls, grown on demand): a copy into pageable memory runs at a
 * fraction of the link's bandwidth.  Serialised by a mutex. */
inline int readback(void *dst, const void *src, size_t bytes, hipStream_t st) {
    static std::mutex mu;
    static PinBuf<char> stage;
    if (bytes == 0) return GF_OK;
    std::lock_guard<std::mutex> lock(mu);
    const size_t chunk = (size_t)64 << 20;
    if (stage.n < std::min(bytes, chunk)) GF_TRY(stage.alloc(std::min(bytes, chunk)));
    for (size_t off = 0; off < bytes; off += stage.n) {
        const size_t b = std::min(stage.n, bytes - off);
        GF_HIP(hipMemcpyAsync(stage.p, (const char *)src + off, b, hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        parallel_copy((char *)dst + off, stage.p, b);
    }
    return GF_OK;
}

/* HIP-event based kernel timer (on the library's stream): with stride s,
 * launches 0, s, 2s, ... are bracketed by a pair of timing events (each
 * pair adds a few microseconds to the stream's timeline); every launch is
 * counted. */
struct KernelTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    int stride = 0;            /* 0: off */
    int64_t launches = 0;
    hipStream_t stream = nullptr;
    void begin(hipStream_t s, hipEvent_t &a) {
        a = nullptr;
        const bool timed = stride > 0 && launches % stride == 0;
        launches++;
        if (!timed) return;
        (void)hipEventCreate(&a);
        (void)hipEventRecord(a, s);
        stream = s;
    }
    void end(hipEvent_t a) {
        if (!a) return;
        hipEvent_t b;
        (void)hipEventCreate(&b);
        (void)hipEventRecord(b, stream);
        ev.emplace_back(a, b);
    }
    /* call after the stream has been synchronised */
    double total_ms() {
        double t = 0;
        for (auto &e : ev) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) t += ms;
        }
        return t;
    }
    int64_t count() const { return launches; }
    int64_t timed() const { return (int64_t)ev.size(); }
    ~KernelTimer() {
        for (auto &e : ev) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    }
};

/* device scan of int32 counts into int64 exclusive offsets plus total */
int scan_counts(const int *d_in, int64_t n, int64_t *d_out, int64_t *d_total, hipStream_t stream,
                DevBuf<int64_t> &scratch);

int select_device(int device);
int num_cus();

} /* namespace gf */

#endif /* GF_RUNTIME_HPP_ */
