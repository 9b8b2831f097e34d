// Device runtime for one GPU: a HIP stream, a stream-ordered caching arena and
// pinned staging for the few scalars the host loop control reads back.
// (Role of include/detail/scratch_pool.hpp:22-185 in the reference; here the
// arena lives in the context so repeated partitions reuse HBM without hipMalloc.)
#pragma once
#include "gp_common.hpp"

#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <map>
#include <string>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

namespace gp {

class Arena {
public:
    explicit Arena(hipStream_t stream) : stream_(stream) {}
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    ~Arena() { release(); }

    static std::size_t size_class(std::size_t bytes) {
        if (bytes < 256) bytes = 256;
        std::size_t p = 256;
        while (p < bytes && p < (std::size_t(1) << 20)) p <<= 1;
        if (p >= bytes) return p;
        // Above 1 MiB: round to 1/8 of the leading power of two.
        std::size_t lead = std::size_t(1) << 20;
        while (lead * 2 <= bytes) lead <<= 1;
        const std::size_t step = lead / 8;
        return (bytes + step - 1) / step * step;
    }

    void* alloc(std::size_t bytes) {
        if (!bytes) return nullptr;
        const std::size_t cls = size_class(bytes);
        std::lock_guard<std::mutex> lock(mutex_);
        auto it = free_.find(cls);
        if (it != free_.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            live_[p] = cls;
            return p;
        }
        void* p = nullptr;
        hipError_t err = hipMalloc(&p, cls);
        if (err != hipSuccess) {
            // Drop the cache once and retry before failing.
            (void)hipGetLastError();
            trim_locked();
            GP_HIP(hipMalloc(&p, cls));
        }
        total_ += cls;
        live_[p] = cls;
        return p;
    }

    void free(void* p) {
        if (!p) return;
        std::lock_guard<std::mutex> lock(mutex_);
        auto it = live_.find(p);
        if (it == live_.end()) throw std::logic_error("arena free of unknown pointer");
        free_[it->second].push_back(p);
        live_.erase(it);
    }

    // Return all cached (not live) blocks to HIP after the stream drains.
    void trim() {
        std::lock_guard<std::mutex> lock(mutex_);
        trim_locked();
    }
    std::size_t reserved_bytes() const {
        std::lock_guard<std::mutex> lock(mutex_);
        return total_;
    }

private:
    void trim_locked() {
        (void)hipStreamSynchronize(stream_);
        for (auto& kv : free_)
            for (void* p : kv.second) {
                (void)hipFree(p);
                total_ -= kv.first;
            }
        free_.clear();
    }
    void release() {
        (void)hipStreamSynchronize(stream_);
        for (auto& kv : free_)
            for (void* p : kv.second) (void)hipFree(p);
        for (auto& kv : live_) (void)hipFree(kv.first);
        free_.clear();
        live_.clear();
        total_ = 0;
    }
    hipStream_t stream_;
    mutable std::mutex mutex_;
    std::map<std::size_t, std::vector<void*>> free_;
    std::unordered_map<void*, std::size_t> live_;
    std::size_t total_ = 0;
};

// Move-only device array owned by an arena.
template <typename T>
class DBuf {
public:
    DBuf() = default;
    DBuf(Arena* a, std::size_t n) : a_(a), n_(n) { p_ = static_cast<T*>(a->alloc(n * sizeof(T))); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept { swap(o); }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            reset();
            swap(o);
        }
        return *this;
    }
    ~DBuf() { reset(); }
    void reset() {
        if (a_ && p_) a_->free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void swap(DBuf& o) noexcept {
        std::swap(a_, o.a_);
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }
    // Grow-only resize; contents are discarded on growth.
    void ensure(Arena* a, std::size_t n) {
        if (n <= n_ && p_) return;
        reset();
        a_ = a;
        n_ = n;
        p_ = static_cast<T*>(a->alloc((n ? n : 1) * sizeof(T)));
    }
    T* get() const { return p_; }
    std::size_t size() const { return n_; }
    // The arena the array came from (null when empty).
    Arena* arena() const { return a_; }
    explicit operator bool() const { return p_ != nullptr; }

private:
    Arena* a_ = nullptr;
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

// Copies n words from device memory into the context's read-back buffer in
// coherent host memory (one block; read-backs are at most 64 KiB).
template <typename T>
__global__ void __launch_bounds__(256) k_copy_out(const T* src, T* dst, std::size_t n) {
    for (std::size_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// The same for counter words, zeroing the first nclear of them on the
// device after they are copied (each word is read and cleared by one thread).
template <typename T>
__global__ void __launch_bounds__(256) k_copy_clear(T* src, T* dst, std::size_t n, std::size_t nclear) {
    for (std::size_t i = threadIdx.x; i < n; i += blockDim.x) {
        dst[i] = src[i];
        if (i < nclear) src[i] = 0;
    }
}

// Accumulated device time and algorithmic bytes per kernel class, measured
// with HIP events recorded on the context stream around each launch.
struct ProfileEntry {
    double ms = 0.0;
    double bytes = 0.0;
    long long launches = 0;
};

class Context {
public:
    explicit Context(int device) : device_(device) {
        GP_HIP(hipSetDevice(device));
        GP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        arena_ = new Arena(stream_);
        // Coherent: the copy kernel's stores reach host memory directly, so
        // the host reads them once the stream has passed the kernel.
        GP_HIP(hipHostMalloc(&pinned_, kPinnedBytes, hipHostMallocCoherent));
        // Created here, not on first wait, so that a wait (and every read-back
        // through the pinned buffer) creates no runtime objects.
        GP_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
        GP_HIP(hipHostMalloc(&mailbox_, kMailboxBytes, hipHostMallocCoherent));
        std::memset(mailbox_, 0, kMailboxBytes);
        for (auto& e : marks_) GP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        if (const char* e = std::getenv("GP_LIGHT_LANES")) set_light_lanes(std::atoi(e));
        if (const char* e = std::getenv("GP_MEDIUM_LANES")) set_medium_lanes(std::atoi(e));
        if (const char* e = std::getenv("GP_TIER_OVERLAP")) overlap_ = std::atoi(e) != 0;
        if (const char* e = std::getenv("GP_SPIN_SYNC")) spin_ = std::atoi(e) != 0;
        if (const char* e = std::getenv("GP_DEVICE_LOOP")) device_loop_ = std::atoi(e) != 0;
        if (const char* e = std::getenv("GP_LOOP_AHEAD")) loop_ahead_ = std::atoi(e) != 0 ? 1 : 0;
    }
    ~Context() {
        delete arena_;
        if (pinned_) (void)hipHostFree(pinned_);
        if (done_) (void)hipEventDestroy(done_);
        if (mailbox_) (void)hipHostFree(mailbox_);
        for (auto e : marks_)
            if (e) (void)hipEventDestroy(e);
        if (fork_) (void)hipEventDestroy(fork_);
        if (join_) (void)hipEventDestroy(join_);
        if (aux_) (void)hipStreamDestroy(aux_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    hipStream_t stream() const { return stream_; }
    Arena* arena() { return arena_; }

    // Lanes per row team for short rows (4, 8 or 16; GP_LIGHT_LANES).  Four
    // keeps the most rows in flight per wave, which the dependent gathers of
    // short rows need.
    int light_lanes() const { return light_lanes_; }
    // Lanes per row team for medium rows (32 or 64; GP_MEDIUM_LANES).
    int medium_lanes() const { return medium_lanes_; }
    void set_medium_lanes(int lanes) {
        if (lanes != 32 && lanes != 64) throw std::invalid_argument("medium lanes must be 32 or 64");
        medium_lanes_ = lanes;
    }
    void set_light_lanes(int lanes) {
        if (lanes != 4 && lanes != 8 && lanes != 16) throw std::invalid_argument("light lanes must be 4, 8 or 16");
        light_lanes_ = lanes;
    }
    // Side stream for work that overlaps the main stream between fork() and
    // join(); after join() the main stream is ordered after the side work,
    // so buffers freed later (main-stream order) are no longer in use.  The
    // side stream is created on first use: a context that never overlaps
    // holds one stream (one hardware queue).
    bool overlap() const { return overlap_; }
    void set_overlap(bool on) { overlap_ = on; }
    hipStream_t aux() const { return aux_; }
    void fork() {
        if (!aux_) {
            GP_HIP(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking));
            GP_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
            GP_HIP(hipEventCreateWithFlags(&join_, hipEventDisableTiming));
        }
        GP_HIP(hipEventRecord(fork_, stream_));
        GP_HIP(hipStreamWaitEvent(aux_, fork_, 0));
    }
    void join() {
        GP_HIP(hipEventRecord(join_, aux_));
        GP_HIP(hipStreamWaitEvent(stream_, join_, 0));
    }
    // Refinement loops decided on the device (GP_DEVICE_LOOP=0: on the host,
    // one counter read-back per round), and whether the host issues a round
    // before it has seen the previous one end (GP_LOOP_AHEAD=0: after).
    bool device_loop() const { return device_loop_; }
    int loop_ahead() const { return loop_ahead_; }
    int device() const { return device_; }
    void activate() const { GP_HIP(hipSetDevice(device_)); }

    // Waits for the stream.  With spin waits on (default; GP_SPIN_SYNC=0 turns
    // them off) the host polls an event recorded at the tail of the stream
    // instead of sleeping in hipStreamSynchronize: the loop control reads
    // counters back about a thousand times per partition, and each wake-up
    // is GPU idle time.
    // Host waits on the stream (read-backs included) since creation.
    std::uint64_t syncs() const { return syncs_; }

    void sync() {
        ++syncs_;
        if (!spin_) {
            GP_HIP(hipStreamSynchronize(stream_));
            return;
        }
        GP_HIP(hipEventRecord(done_, stream_));
        for (;;) {
            const hipError_t e = hipEventQuery(done_);
            if (e == hipSuccess) return;
            if (e != hipErrorNotReady) GP_HIP(e);
        }
    }

    // Words of coherent host memory that device-side loop control kernels
    // write their state to, so the host can follow a loop without a copy.
    unsigned long long* mailbox() const { return static_cast<unsigned long long*>(mailbox_); }
    // Stream positions: mark(i) records ring event i at the stream tail;
    // wait_mark(i) spins until the stream has passed it (counted as a host
    // wait).  Two events, so a loop can run one step ahead of the host.
    static constexpr int kMarks = 2;
    void mark(int i) { GP_HIP(hipEventRecord(marks_[i % kMarks], stream_)); }
    void wait_mark(int i) {
        ++syncs_;
        hipEvent_t e = marks_[i % kMarks];
        if (!spin_) {
            GP_HIP(hipEventSynchronize(e));
            return;
        }
        for (;;) {
            const hipError_t r = hipEventQuery(e);
            if (r == hipSuccess) return;
            if (r != hipErrorNotReady) GP_HIP(r);
        }
    }

    // D2H of a small region through the coherent pinned buffer, written by
    // a one-block kernel (cheaper than the runtime's copy for the counter
    // read-backs of the loop control); synchronizes the stream.
    void read(void* host, const void* dev, std::size_t bytes) {
        if (!bytes) return;
        if (bytes > kPinnedBytes) {
            ++::gp::thread_launches();
            GP_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream_));
            sync();
            return;
        }
        const auto addr = reinterpret_cast<std::uintptr_t>(dev);
        if (addr % 8 == 0 && bytes % 8 == 0)
            hipLaunchKernelGGL(k_copy_out<unsigned long long>, dim3(1), dim3(256), 0, stream_,
                               static_cast<const unsigned long long*>(dev), static_cast<unsigned long long*>(pinned_),
                               bytes / 8);
        else if (addr % 4 == 0 && bytes % 4 == 0)
            hipLaunchKernelGGL(k_copy_out<std::uint32_t>, dim3(1), dim3(256), 0, stream_,
                               static_cast<const std::uint32_t*>(dev), static_cast<std::uint32_t*>(pinned_), bytes / 4);
        else
            hipLaunchKernelGGL(k_copy_out<std::uint8_t>, dim3(1), dim3(256), 0, stream_,
                               static_cast<const std::uint8_t*>(dev), static_cast<std::uint8_t*>(pinned_), bytes);
        GP_LAUNCH_CHECK();
        sync();
        std::memcpy(host, pinned_, bytes);
    }
    // read() of counter words that also zeroes the first nclear words on the
    // device, so the next user of the counters needs no clearing launch.
    void read_clear(unsigned long long* host, unsigned long long* dev, std::size_t words, std::size_t nclear) {
        if (!words) return;
        if (words * sizeof(unsigned long long) > kPinnedBytes || nclear > words)
            throw std::invalid_argument("read_clear: region exceeds the read-back buffer");
        hipLaunchKernelGGL(k_copy_clear<unsigned long long>, dim3(1), dim3(256), 0, stream_, dev,
                           static_cast<unsigned long long*>(pinned_), words, nclear);
        GP_LAUNCH_CHECK();
        sync();
        std::memcpy(host, pinned_, words * sizeof(unsigned long long));
    }
    template <typename T>
    T read_scalar(const T* dev) {
        T v;
        read(&v, dev, sizeof(T));
        return v;
    }
    void upload(void* dev, const void* host, std::size_t bytes) {
        if (!bytes) return;
        ++::gp::thread_launches();  // copies count as stream commands
        GP_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, stream_));
        sync();
    }
    void download(void* host, const void* dev, std::size_t bytes) {
        if (!bytes) return;
        ++::gp::thread_launches();
        GP_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream_));
        sync();
    }
    void copy_d2d(void* dst, const void* src, std::size_t bytes) {
        if (!bytes) return;
        ++::gp::thread_launches();
        GP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_));
    }

    template <typename T>
    DBuf<T> buf(std::size_t n) { return DBuf<T>(arena_, n); }

    bool profiling() const { return profiling_; }
    void set_profiling(bool on) { profiling_ = on; }
    // Wraps one launch; bytes are the launch's algorithmic traffic.
    template <typename Fn>
    void timed(const char* name, double bytes, Fn&& fn) {
        if (!profiling_) {
            fn();
            return;
        }
        hipEvent_t a, b;
        GP_HIP(hipEventCreate(&a));
        GP_HIP(hipEventCreate(&b));
        GP_HIP(hipEventRecord(a, stream_));
        fn();
        GP_HIP(hipEventRecord(b, stream_));
        pending_.push_back(Pending{name, bytes, a, b});
    }
    // Position in the list of timed intervals not yet resolved, and removal
    // of the intervals recorded after such a position (work that a
    // device-side loop skipped: its launches returned at once).
    std::size_t pending_mark() const { return pending_.size(); }
    void drop_pending(std::size_t from) {
        while (pending_.size() > from) {
            (void)hipEventDestroy(pending_.back().a);
            (void)hipEventDestroy(pending_.back().b);
            pending_.pop_back();
        }
    }
    // Algorithmic bytes known only after the work ran (device counters).
    void add_bytes(const char* name, double bytes) {
        if (profiling_) profile_[name].bytes += bytes;
    }
    // Synchronizes and folds pending events into the totals.
    void resolve_profile() {
        if (pending_.empty()) return;
        sync();
        for (auto& p : pending_) {
            float ms = 0.f;
            GP_HIP(hipEventElapsedTime(&ms, p.a, p.b));
            ProfileEntry& e = profile_[p.name];
            e.ms += ms;
            e.bytes += p.bytes;
            e.launches += 1;
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        pending_.clear();
    }
    std::map<std::string, ProfileEntry>& profile() {
        resolve_profile();
        return profile_;
    }

private:
    struct Pending {
        std::string name;
        double bytes;
        hipEvent_t a, b;
    };
    static constexpr std::size_t kPinnedBytes = 1 << 16;
    static constexpr std::size_t kMailboxBytes = 4096;
    bool profiling_ = false;
    std::vector<Pending> pending_;
    std::map<std::string, ProfileEntry> profile_;
    int device_;
    hipStream_t stream_ = nullptr;
    hipStream_t aux_ = nullptr;
    hipEvent_t fork_ = nullptr, join_ = nullptr, done_ = nullptr;
    bool overlap_ = true;
    bool spin_ = true;
    bool device_loop_ = true;
    int loop_ahead_ = 0;
    Arena* arena_ = nullptr;
    void* pinned_ = nullptr;
    void* mailbox_ = nullptr;
    hipEvent_t marks_[kMarks] = {nullptr, nullptr};
    int light_lanes_ = 4;
    int medium_lanes_ = 64;
    std::uint64_t syncs_ = 0;
};

}  // namespace gp
