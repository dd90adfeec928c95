// HIP runtime helpers for the host engine: error checks that throw, RAII device
// and pinned host buffers, and a per-thread stream bound to a device.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

namespace kodama_hip {

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
  }
}

template <class T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { resize(n); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_; n_ = o.n_;
      o.p_ = nullptr; o.n_ = 0;
    }
    return *this;
  }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n > 0) check(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
    n_ = n;
  }
  void release() {
    if (p_ != nullptr) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  void upload(const T* src, size_t n, hipStream_t s, size_t offset = 0) {
    if (n) check(hipMemcpyAsync(p_ + offset, src, n * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  }
  void download(T* dst, size_t n, hipStream_t s, size_t offset = 0) const {
    if (n) check(hipMemcpyAsync(dst, p_ + offset, n * sizeof(T), hipMemcpyDeviceToHost, s), "download");
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Stream-ordered device memory from the device's default memory pool. The
// pool keeps freed memory (its release threshold is raised once per device),
// so an evolution that follows another of similar size allocates and frees
// without a trip to the driver.
class PoolBuffer {
 public:
  PoolBuffer() = default;
  ~PoolBuffer() { release(); }
  PoolBuffer(const PoolBuffer&) = delete;
  PoolBuffer& operator=(const PoolBuffer&) = delete;
  void allocate(size_t bytes, int device, hipStream_t s) {
    release();
    retain_pool(device);
    check(hipMallocAsync(&p_, bytes, s), "hipMallocAsync");
    bytes_ = bytes;
    s_ = s;
  }
  void release() {
    if (p_ != nullptr) (void)hipFreeAsync(p_, s_);
    p_ = nullptr;
    bytes_ = 0;
  }
  unsigned char* get() const { return static_cast<unsigned char*>(p_); }
  size_t size() const { return bytes_; }

 private:
  static void retain_pool(int device) {
    static std::atomic<bool> done[64];
    if (device < 0 || device >= 64 || done[device].load(std::memory_order_acquire)) return;
    hipMemPool_t pool = nullptr;
    check(hipDeviceGetDefaultMemPool(&pool, device), "hipDeviceGetDefaultMemPool");
    uint64_t threshold = UINT64_MAX;  // setting it twice from two threads is harmless
    check(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold), "hipMemPoolSetAttribute");
    done[device].store(true, std::memory_order_release);
  }

  void* p_ = nullptr;
  size_t bytes_ = 0;
  hipStream_t s_ = nullptr;
};

template <class T>
class PinnedBuffer {
 public:
  PinnedBuffer() = default;
  explicit PinnedBuffer(size_t n) { resize(n); }
  ~PinnedBuffer() { release(); }
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n > 0) check(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    n_ = n;
  }
  void release() {
    if (p_ != nullptr) (void)hipHostFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  T& operator[](size_t i) { return p_[i]; }
  const T& operator[](size_t i) const { return p_[i]; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// A stream on a chosen device, default priority.
class Stream {
 public:
  explicit Stream(int device) : device_(device) {
    check(hipSetDevice(device), "hipSetDevice");
    check(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate");
  }
  ~Stream() {
    if (s_ != nullptr) (void)hipStreamDestroy(s_);
  }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  hipStream_t get() const { return s_; }
  int device() const { return device_; }
  void sync() const { check(hipStreamSynchronize(s_), "hipStreamSynchronize"); }

 private:
  hipStream_t s_ = nullptr;
  int device_ = 0;
};

// Event-timed region on a stream (milliseconds).
class EventTimer {
 public:
  EventTimer() {
    check(hipEventCreate(&a_), "hipEventCreate");
    check(hipEventCreate(&b_), "hipEventCreate");
  }
  ~EventTimer() {
    (void)hipEventDestroy(a_);
    (void)hipEventDestroy(b_);
  }
  void start(hipStream_t s) { check(hipEventRecord(a_, s), "hipEventRecord"); }
  void stop(hipStream_t s) { check(hipEventRecord(b_, s), "hipEventRecord"); }
  void wait() { check(hipEventSynchronize(b_), "hipEventSynchronize"); }
  float ms() {
    check(hipEventSynchronize(b_), "hipEventSynchronize");
    float v = 0.0f;
    check(hipEventElapsedTime(&v, a_, b_), "hipEventElapsedTime");
    return v;
  }

 private:
  hipEvent_t a_ = nullptr, b_ = nullptr;
};

}  // namespace kodama_hip
