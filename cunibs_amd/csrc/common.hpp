// Host-side plumbing shared by every translation unit of the library: status codes, the HIP
// error check, an owning device buffer, launch-size helpers and the XCD-aware block remap.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>

namespace cn {

// Wavefront width on CDNA4. Every reduction tree in the library is written for 64 lanes.
constexpr int kWave = 64;
constexpr int kBlock = 256;

// Error kinds carried across the C ABI as status codes (see include/cunibs_amd.h).
struct InvalidArgument : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NotConverged : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(hipError_t err, const char* what) {
    if (err != hipSuccess) {
        throw HipError(std::string("HIP error in ") + what + ": " + hipGetErrorString(err));
    }
}

inline void require(bool ok, const std::string& what) {
    if (!ok) throw InvalidArgument(what);
}

inline unsigned grid_for(std::int64_t total, int per = kBlock) {
    if (total <= 0) return 0;
    return static_cast<unsigned>((total + per - 1) / per);
}

// Owning device allocation. Move-only; freed on destruction.
template <typename T>
class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t count, const char* what = "buffer") { reset(count, what); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }

    void reset(std::size_t count, const char* what = "buffer") {
        release();
        if (count == 0) return;
        void* p = nullptr;
        check(hipMalloc(&p, count * sizeof(T)), what);
        p_ = static_cast<T*>(p);
        n_ = count;
    }
    // Grow-only: keep the allocation when it already holds `count` elements.
    void ensure(std::size_t count, const char* what = "buffer") {
        if (count > n_) reset(count, what);
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    std::size_t size() const { return n_; }
    explicit operator bool() const { return p_ != nullptr; }

  private:
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

template <typename T>
void h2d(T* dst, const T* src, std::size_t count, hipStream_t s) {
    if (count) check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s), "copy H2D");
}
template <typename T>
void d2h(T* dst, const T* src, std::size_t count, hipStream_t s) {
    if (count) check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, s), "copy D2H");
}
template <typename T>
void d2d(T* dst, const T* src, std::size_t count, hipStream_t s) {
    if (count) check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToDevice, s), "copy D2D");
}

template <typename T>
DeviceBuffer<T> upload(const T* src, std::size_t count, hipStream_t s, const char* what) {
    DeviceBuffer<T> b(count, what);
    h2d(b.get(), src, count, s);
    return b;
}

inline void sync(hipStream_t s, const char* what) { check(hipStreamSynchronize(s), what); }

// Pinned host staging for host-to-device copies issued without a wait after them. A copy from
// pageable memory may be carried out by the runtime after hipMemcpyAsync has returned, and a
// function-local vector is freed (a large one unmapped) when the function returns; staging the
// bytes in pinned memory that the stage owns makes the source outlive the copy. One use: fill()
// waits until every copy of the previous use has completed (an event recorded after the last
// of them) and returns the pinned bytes; send() queues copies out of them; done() records the
// event. Grow-only.
class PinnedStage {
  public:
    PinnedStage() = default;
    PinnedStage(const PinnedStage&) = delete;
    PinnedStage& operator=(const PinnedStage&) = delete;
    ~PinnedStage() {
        if (ev_) {
            (void)hipEventSynchronize(ev_);
            (void)hipEventDestroy(ev_);
        }
        if (p_) (void)hipHostFree(p_);
    }
    void* fill(std::size_t bytes) {
        if (ev_) check(hipEventSynchronize(ev_), "staging wait");
        if (bytes > cap_) {
            if (p_) check(hipHostFree(p_), "staging free");
            p_ = nullptr;
            cap_ = 0;
            check(hipHostMalloc(&p_, bytes), "pinned staging");
            cap_ = bytes;
        }
        return p_;
    }
    void send(void* dst, std::size_t offset, std::size_t bytes, hipStream_t s) {
        if (bytes) check(hipMemcpyAsync(dst, static_cast<char*>(p_) + offset, bytes, hipMemcpyHostToDevice, s), "staged copy");
    }
    void done(hipStream_t s) {
        if (!ev_) check(hipEventCreateWithFlags(&ev_, hipEventDisableTiming), "staging event");
        check(hipEventRecord(ev_, s), "staging event");
    }
    // One array: stage `count` values of `src` and copy them to `dst` on `s`.
    template <typename T>
    void put(T* dst, const T* src, std::size_t count, hipStream_t s) {
        if (!count) return;
        std::memcpy(fill(count * sizeof(T)), src, count * sizeof(T));
        send(dst, 0, count * sizeof(T), s);
        done(s);
    }

  private:
    void* p_ = nullptr;
    std::size_t cap_ = 0;
    hipEvent_t ev_ = nullptr;
};

}  // namespace cn
