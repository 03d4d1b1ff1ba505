// Host runtime: device context with one stream, a caching device allocator, typed buffers and the
// error types the C ABI maps to status codes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace brook {

struct DeviceError : std::runtime_error {
  hipError_t code;
  DeviceError(hipError_t c, const std::string &what) : std::runtime_error(what), code(c) {}
};
struct Unsupported : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

#define BROOK_HIP(call)                                                                                 \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess)                                                                               \
      throw ::brook::DeviceError(e_, std::string(#call) + ": " + hipGetErrorString(e_) + " at " + __FILE__ + \
                                         ":" + std::to_string(__LINE__));                              \
  } while (0)
#define BROOK_LAUNCHED() BROOK_HIP(hipGetLastError())

class Context;

// The allocation size class of a request of `bytes` (the pool reuses blocks by class).
size_t size_class(size_t bytes);

// A block of device memory; returned to its context's cache when released.
class Block {
 public:
  Block() = default;
  Block(Context *ctx, void *ptr, size_t bytes) : ctx_(ctx), ptr_(ptr), bytes_(bytes) {}
  Block(const Block &) = delete;
  Block &operator=(const Block &) = delete;
  Block(Block &&o) noexcept { swap(o); }
  Block &operator=(Block &&o) noexcept { release(); swap(o); return *this; }
  ~Block() { release(); }
  void *get() const { return ptr_; }
  size_t bytes() const { return bytes_; }
  void release();
  void swap(Block &o) { std::swap(ctx_, o.ctx_); std::swap(ptr_, o.ptr_); std::swap(bytes_, o.bytes_); }

 private:
  Context *ctx_ = nullptr;
  void *ptr_ = nullptr;
  size_t bytes_ = 0;
};

// While a host thread has a FreeHold installed (FreeHoldScope), the blocks it releases that belong
// to other contexts' pools than the hold's own are kept back instead of going to their caches:
// those pools may serve other threads meanwhile, and their streams could reuse a block that this
// thread's stream still reads. release() gives them back; call it once no stream uses them and no
// other thread uses their pools.
class FreeHold {
 public:
  explicit FreeHold(Context &own) : own_(&own) {}
  FreeHold(const FreeHold &) = delete;
  FreeHold &operator=(const FreeHold &) = delete;
  ~FreeHold() { release(); }
  void release();

 private:
  friend class Block;
  Context *own_;
  std::vector<Block> kept_;
};

class FreeHoldScope {
 public:
  explicit FreeHoldScope(FreeHold &hold);
  FreeHoldScope(const FreeHoldScope &) = delete;
  FreeHoldScope &operator=(const FreeHoldScope &) = delete;
  ~FreeHoldScope();

 private:
  FreeHold *previous_;
};

template <class T> class Buffer {
 public:
  Buffer() = default;
  Buffer(Context &ctx, size_t count);
  T *data() const { return static_cast<T *>(block_.get()); }
  size_t size() const { return count_; }
  void clear(Context &ctx);                                 // bytes to 0
  void fill_bytes(Context &ctx, int value);
  void upload(Context &ctx, const T *host, size_t count);   // blocks
  std::vector<T> download(Context &ctx, size_t count) const;  // blocks
  std::vector<T> download(Context &ctx) const { return download(ctx, count_); }

 private:
  Block block_;
  size_t count_ = 0;
};

class Context {
 public:
  explicit Context(int device);
  ~Context();
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  int compute_units() const { return cus_; }
  void activate() const { BROOK_HIP(hipSetDevice(device_)); }
  void synchronize() { BROOK_HIP(hipStreamSynchronize(stream_)); }
  Block allocate(size_t bytes);
  void give_back(void *ptr, size_t bytes);
  void trim();
  std::mutex &mutex() { return mutex_; }
  std::map<std::string, double> stats;
  bool kernel_timing = false;  // time instrumented kernels with events on this stream
  // Auxiliary context i on the same device (a stream and an allocation cache of its own), made on
  // first use, for work that other host threads run beside this context's stream. Only the
  // thread that calls into this context makes them; each serves one thread at a time.
  Context &aux(int i = 0);

 private:
  friend class KernelTimer;

 private:
  std::vector<std::unique_ptr<Context>> aux_;
  int device_ = 0, cus_ = 0;
  hipStream_t stream_ = nullptr;
  std::multimap<size_t, void *> cache_;
  size_t cached_bytes_ = 0, cache_cap_ = ~size_t(0);
  std::mutex mutex_;
};

// Brackets the kernels launched in its scope with two events on the context's stream and adds
// the elapsed time to stats["k_<name>_ms"] (and a launch count to stats["k_<name>_n"]). Does
// nothing unless the context's kernel timing is on.
class KernelTimer {
 public:
  KernelTimer(Context &ctx, const char *name) : ctx_(ctx), name_(name) {
    if (!ctx_.kernel_timing) return;
    BROOK_HIP(hipEventCreate(&begin_));
    BROOK_HIP(hipEventCreate(&end_));
    BROOK_HIP(hipEventRecord(begin_, ctx_.stream()));
  }
  ~KernelTimer() {
    if (!ctx_.kernel_timing) return;
    float ms = 0.0f;
    if (hipEventRecord(end_, ctx_.stream()) == hipSuccess && hipEventSynchronize(end_) == hipSuccess &&
        hipEventElapsedTime(&ms, begin_, end_) == hipSuccess) {
      ctx_.stats[std::string("k_") + name_ + "_ms"] += ms;
      ctx_.stats[std::string("k_") + name_ + "_n"] += 1;
    }
    (void)hipEventDestroy(begin_);
    (void)hipEventDestroy(end_);
  }

 private:
  Context &ctx_;
  const char *name_;
  hipEvent_t begin_ = nullptr, end_ = nullptr;
};

// Serializes calls on one context and restores the caller's device afterwards.
class ContextCall {
 public:
  explicit ContextCall(Context &ctx) : lock_(ctx.mutex()) {
    BROOK_HIP(hipGetDevice(&previous_));
    ctx.activate();
  }
  ~ContextCall() { (void)hipSetDevice(previous_); }

 private:
  std::lock_guard<std::mutex> lock_;
  int previous_ = 0;
};

template <class T> Buffer<T>::Buffer(Context &ctx, size_t count) : block_(ctx.allocate(count * sizeof(T))), count_(count) {}
template <class T> void Buffer<T>::clear(Context &ctx) { fill_bytes(ctx, 0); }
template <class T> void Buffer<T>::fill_bytes(Context &ctx, int value) {
  if (count_) BROOK_HIP(hipMemsetAsync(data(), value, count_ * sizeof(T), ctx.stream()));
}
template <class T> void Buffer<T>::upload(Context &ctx, const T *host, size_t count) {
  if (count) BROOK_HIP(hipMemcpyAsync(data(), host, count * sizeof(T), hipMemcpyHostToDevice, ctx.stream()));
  ctx.synchronize();
}
template <class T> std::vector<T> Buffer<T>::download(Context &ctx, size_t count) const {
  std::vector<T> out(count);
  if (count) BROOK_HIP(hipMemcpyAsync(out.data(), data(), count * sizeof(T), hipMemcpyDeviceToHost, ctx.stream()));
  ctx.synchronize();
  return out;
}

}  // namespace brook
