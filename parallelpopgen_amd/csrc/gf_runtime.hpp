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
#include <condition_variable>
#include <cstring>
#include <string>
#include <vector>
#include <chrono>
#include <algorithm>
#include <memory>
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

/* A library stream and its liveness.  Every entry point that issues device
 * work creates an OwnedStream and makes it current for its thread
 * (StreamScope); device buffers allocated meanwhile remember it, and on
 * release record an event on it instead of waiting for the device.  An
 * OwnedStream synchronises before it is destroyed and then marks its tag
 * dead, so a buffer that outlives its stream has no pending work on it
 * (every entry point also synchronises its stream before it returns). */
struct StreamTag {
    hipStream_t st = nullptr;
    bool alive = true;
};

inline std::shared_ptr<StreamTag> &current_stream_tag() {
    static thread_local std::shared_ptr<StreamTag> t;
    return t;
}

struct OwnedStream {
    hipStream_t st = nullptr;
    std::shared_ptr<StreamTag> tag;
    OwnedStream() = default;
    OwnedStream(const OwnedStream &) = delete;
    OwnedStream &operator=(const OwnedStream &) = delete;
    hipError_t create() {
        if (st) return hipSuccess;
        const hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) { st = nullptr; return e; }
        tag = std::make_shared<StreamTag>();
        tag->st = st;
        return hipSuccess;
    }
    ~OwnedStream() {
        if (!st) return;
        (void)hipStreamSynchronize(st);
        tag->alive = false;
        (void)hipStreamDestroy(st);
    }
    operator hipStream_t() const { return st; }
};

/* makes `s` the calling thread's current library stream for its lifetime */
struct StreamScope {
    std::shared_ptr<StreamTag> prev;
    explicit StreamScope(const OwnedStream &s) : prev(current_stream_tag()) { current_stream_tag() = s.tag; }
    ~StreamScope() { current_stream_tag() = prev; }
};

/* Per-device cache of released device blocks.  hipFree waits for the whole
 * device and took up to 80 ms for large blocks (a batched run releases
 * about 60 buffers); a released block is kept instead.  A block released
 * by a buffer whose stream is alive carries an event recorded on that
 * stream at release, and is handed out again only after that event
 * completes (no device-wide wait); a block whose stream has ended is free
 * at once; a block of unknown stream waits for the device at release.  An
 * allocation takes the smallest cached block of its size class up to twice
 * its size, preferring blocks whose event has completed.  At most a third
 * of the device's memory (96 GiB on an MI355X; at least kKeepMin) is kept
 * per device, the oldest blocks freed first: a workload whose buffers
 * vary from call to call across size classes (batched runs of 8,192
 * simulations hold about 30 GiB each) keeps one block per class it uses
 * instead of freeing and allocating multi-GiB blocks on every call.
 * Blocks above kMaxBlock are freed at once; a failed hipMalloc empties the
 * device's cache and tries again; gf_release_device_cache empties it on
 * request.  The cache lives until the process ends. */
class DevCache {
public:
    static constexpr size_t kKeepMin = (size_t)32 << 30;
    static constexpr size_t kMaxBlock = (size_t)16 << 30;

    static DevCache &get() {
        static DevCache *c = new DevCache();   /* never destroyed: no teardown-order hazard */
        return *c;
    }

    /* size class: powers of two from 256 B below 1 MiB, 2 MiB steps up to
     * 64 MiB, then four classes per power of two (64, 80, 96, 112, 128,
     * 160 ... MiB): a workload whose large buffers vary by some percent
     * from call to call (batched runs of different draws) asks for the
     * same few classes again, instead of leaving the cache full of blocks
     * that are each a little too small */
    static size_t size_class(size_t bytes) {
        const size_t MiB = (size_t)1 << 20;
        if (bytes >= 64 * MiB) {
            size_t p = 64 * MiB;
            while (2 * p <= bytes) p <<= 1;
            const size_t q = p / 4;
            return (bytes + q - 1) / q * q;
        }
        if (bytes >= MiB) return (bytes + 2 * MiB - 1) & ~(2 * MiB - 1);
        size_t c = 256;
        while (c < bytes) c <<= 1;
        return c;
    }

    hipError_t alloc(void **p, size_t bytes, size_t &block, int &dev) {
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        block = size_class(bytes);
        hipEvent_t wait = nullptr;
        {
            std::lock_guard<std::mutex> lock(mu_);
            Dev &D = dev_(dev);
            size_t best = D.blocks.size();
            bool best_ready = false;
            for (size_t i = 0; i < D.blocks.size(); i++) {
                const size_t b = D.blocks[i].bytes;
                if (b < block || b > 2 * block) continue;
                const bool ready = !D.blocks[i].ev || hipEventQuery(D.blocks[i].ev) == hipSuccess;
                if (best == D.blocks.size() || (ready && !best_ready) ||
                    (ready == best_ready && b < D.blocks[best].bytes)) {
                    best = i;
                    best_ready = ready;
                }
            }
            if (best < D.blocks.size()) {
                *p = D.blocks[best].p;
                block = D.blocks[best].bytes;
                wait = D.blocks[best].ev;
                D.kept -= block;
                D.blocks.erase(D.blocks.begin() + (ptrdiff_t)best);
            }
        }
        if (*p) {
            if (wait) {
                e = hipEventSynchronize(wait);
                recycle(wait);
            }
            return e;
        }
        e = hipMalloc(p, block);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            flush(dev);
            e = hipMalloc(p, block);
        }
        if (e == hipSuccess) {
            std::lock_guard<std::mutex> lock(mu_);
            Dev &D = dev_(dev);
            D.mallocs++;
            D.held += block;
        }
        return e;
    }

    void release(void *p, size_t block, int dev, const std::shared_ptr<StreamTag> &tag) {
        int cur = dev;
        (void)hipGetDevice(&cur);
        if (cur != dev) (void)hipSetDevice(dev);
        hipEvent_t ev = nullptr;
        if (tag && tag->alive) {
            ev = event();
            if (ev && hipEventRecord(ev, tag->st) != hipSuccess) {
                (void)hipGetLastError();
                (void)hipEventDestroy(ev);
                ev = nullptr;
                (void)hipStreamSynchronize(tag->st);
            }
        } else if (!tag) {
            (void)hipDeviceSynchronize();
        }
        if (block > kMaxBlock) {
            std::vector<Block> drop{{p, block, ev}};
            free_blocks(drop, dev);
        } else {
            std::vector<Block> drop;
            {
                std::lock_guard<std::mutex> lock(mu_);
                Dev &D = dev_(dev);
                D.blocks.push_back({p, block, ev});
                D.kept += block;
                while (D.kept > D.keep && !D.blocks.empty()) {
                    drop.push_back(D.blocks.front());
                    D.kept -= D.blocks.front().bytes;
                    D.blocks.erase(D.blocks.begin());
                }
            }
            free_blocks(drop, dev);
        }
        if (cur != dev) (void)hipSetDevice(cur);
    }

    /* frees every cached block of device `dev` (the current device) */
    void flush(int dev) {
        std::vector<Block> drop;
        {
            std::lock_guard<std::mutex> lock(mu_);
            Dev &D = dev_(dev);
            drop.swap(D.blocks);
            D.kept = 0;
        }
        free_blocks(drop, dev);
    }

    /* device `dev`: bytes and blocks kept for reuse, bytes held from the
     * device (in use + kept), hipMalloc and hipFree calls so far */
    void stats(int dev, int64_t out[5]) {
        std::lock_guard<std::mutex> lock(mu_);
        const Dev &D = dev_(dev);
        out[0] = (int64_t)D.kept;
        out[1] = (int64_t)D.blocks.size();
        out[2] = (int64_t)D.held;
        out[3] = D.mallocs;
        out[4] = D.frees;
    }

private:
    struct Block { void *p; size_t bytes; hipEvent_t ev; };
    struct Dev {
        std::vector<Block> blocks;
        size_t kept = 0, held = 0;
        size_t keep = 0;   /* 0: not yet set */
        int64_t mallocs = 0, frees = 0;
    };
    std::mutex mu_;
    std::vector<Dev> devs_;
    std::vector<hipEvent_t> spare_;   /* completed events for reuse */

    Dev &dev_(int d) {
        if ((size_t)d >= devs_.size()) devs_.resize((size_t)d + 1);
        Dev &D = devs_[(size_t)d];
        if (!D.keep) {
            size_t total = 0;
            if (hipDeviceTotalMem(&total, d) != hipSuccess) {
                (void)hipGetLastError();
                total = 0;
            }
            D.keep = std::max(kKeepMin, total / 3);
        }
        return D;
    }

    hipEvent_t event() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            if (!spare_.empty()) {
                hipEvent_t e = spare_.back();
                spare_.pop_back();
                return e;
            }
        }
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return e;
    }

    void recycle(hipEvent_t e) {
        std::lock_guard<std::mutex> lock(mu_);
        spare_.push_back(e);
    }

    void free_blocks(std::vector<Block> &drop, int dev) {
        for (const Block &b : drop) {
            if (b.ev) { (void)hipEventSynchronize(b.ev); recycle(b.ev); }
            (void)hipFree(b.p);
            std::lock_guard<std::mutex> lock(mu_);
            Dev &D = dev_(dev);
            D.frees++;
            D.held -= b.bytes;
        }
        drop.clear();
    }
};

/* owning device buffer (n elements; the block under it may be larger);
 * remembers the calling thread's current library stream at allocation */
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    size_t block = 0;
    int dev = 0;
    std::shared_ptr<StreamTag> tag;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) DevCache::get().release(p, block, dev, tag);
        p = nullptr;
        n = 0;
        block = 0;
        tag.reset();
    }
    int alloc(size_t count) {
        if (count <= n && p) {
            /* reused under another call's stream: the earlier call
             * synchronised its stream before it returned */
            if (current_stream_tag()) tag = current_stream_tag();
            return GF_OK;
        }
        release();
        if (count == 0) count = 1;
        hipError_t e = DevCache::get().alloc((void **)&p, count * sizeof(T), block, dev);
        if (e != hipSuccess) {
            p = nullptr;
            block = 0;
            set_error(std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
            return GF_ERR_ALLOC;
        }
        tag = current_stream_tag();
        n = count;
        return GF_OK;
    }
    void swap(DevBuf &o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
        std::swap(block, o.block);
        std::swap(dev, o.dev);
        std::swap(tag, o.tag);
    }
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

/* host memcpy split over a few threads (one core copies ~10 GB/s): the
 * calling thread and up to 7 pool workers, started once and kept until the
 * process ends */
class CopyPool {
public:
    static CopyPool &get() {
        static CopyPool *pool = new CopyPool();   /* never destroyed: workers outlive static teardown */
        return *pool;
    }
    void copy(char *dst, const char *src, size_t bytes) {
        const size_t min_part = (size_t)1 << 20;
        const size_t parts = std::min<size_t>(workers_.size() + 1, std::max<size_t>(1, bytes / min_part));
        if (parts <= 1) { std::memcpy(dst, src, bytes); return; }
        const size_t step = (bytes + parts - 1) / parts;
        std::lock_guard<std::mutex> one_job(job_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = dst; src_ = src; bytes_ = bytes; step_ = step;
            next_ = 1; parts_ = parts; pending_ = parts - 1;
            ++gen_;
        }
        cv_.notify_all();
        std::memcpy(dst, src, std::min(step, bytes));
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    CopyPool() {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned n = std::min(7u, hw > 1 ? hw - 1 : 0u);
        for (unsigned k = 0; k < n; k++) workers_.emplace_back([this] { run(); });
        for (auto &t : workers_) t.detach();
    }
    void run() {
        uint64_t seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return gen_ != seen && next_ < parts_; });
            const size_t k = next_++;
            if (next_ >= parts_) seen = gen_;
            char *d = dst_;
            const char *s = src_;
            const size_t off = k * step_, b = off < bytes_ ? std::min(step_, bytes_ - off) : 0;
            lk.unlock();
            if (b) std::memcpy(d + off, s + off, b);
            lk.lock();
            if (--pending_ == 0) done_cv_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex job_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0, step_ = 0, next_ = 0, parts_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
};

inline void parallel_copy(char *dst, const char *src, size_t bytes) { CopyPool::get().copy(dst, src, bytes); }

/* Device -> host copy through a process-wide pinned staging buffer (kept
 * between calls, grown on demand): a copy into pageable memory runs at a
 * fraction of the link's bandwidth.  Two halves: the link fills one while
 * the host copies the other out.  Serialised by a mutex. */
inline int readback(void *dst, const void *src, size_t bytes, hipStream_t st) {
    static std::mutex mu;
    static PinBuf<char> stage;
    if (bytes == 0) return GF_OK;
    std::lock_guard<std::mutex> lock(mu);
    const size_t half = std::min(bytes, (size_t)16 << 20);
    if (stage.n < 2 * half) GF_TRY(stage.alloc(2 * half));
    const size_t h = stage.n / 2;
    struct Events {   /* created on the stream's device, destroyed on return */
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() { for (auto x : e) if (x) (void)hipEventDestroy(x); }
    } ev;
    hipEvent_t *filled = ev.e;
    for (int k = 0; k < 2; k++) GF_HIP(hipEventCreateWithFlags(&filled[k], hipEventDisableTiming));
    const size_t nchunks = (bytes + h - 1) / h;
    auto issue = [&](size_t c) -> int {
        const size_t off = c * h, b = std::min(h, bytes - off);
        GF_HIP(hipMemcpyAsync(stage.p + (c & 1) * h, (const char *)src + off, b, hipMemcpyDeviceToHost, st));
        GF_HIP(hipEventRecord(filled[c & 1], st));
        return GF_OK;
    };
    GF_TRY(issue(0));
    for (size_t c = 0; c < nchunks; c++) {
        if (c + 1 < nchunks) GF_TRY(issue(c + 1));   /* the other half was copied out in the previous round */
        GF_HIP(hipEventSynchronize(filled[c & 1]));
        const size_t off = c * h;
        parallel_copy((char *)dst + off, stage.p + (c & 1) * h, std::min(h, bytes - off));
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
