#include "runtime.hpp"

namespace brook {

namespace {
thread_local FreeHold *t_hold = nullptr;  // the calling thread's installed hold, if any
}  // namespace

void Block::release() {
  if (ptr_ && ctx_) {
    if (t_hold && ctx_ != t_hold->own_) t_hold->kept_.emplace_back(ctx_, ptr_, bytes_);
    else ctx_->give_back(ptr_, bytes_);
  }
  ptr_ = nullptr;
  ctx_ = nullptr;
  bytes_ = 0;
}

void FreeHold::release() {
  std::vector<Block> kept;
  kept.swap(kept_);
  FreeHold *installed = t_hold;
  t_hold = nullptr;  // the kept blocks go to their own caches now
  kept.clear();
  t_hold = installed;
}

FreeHoldScope::FreeHoldScope(FreeHold &hold) : previous_(t_hold) { t_hold = &hold; }
FreeHoldScope::~FreeHoldScope() { t_hold = previous_; }

Context::Context(int device) : device_(device) {
  int count = 0;
  BROOK_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) throw std::invalid_argument("device index out of range");
  BROOK_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  BROOK_HIP(hipGetDeviceProperties(&prop, device));
  cus_ = prop.multiProcessorCount;
  cache_cap_ = prop.totalGlobalMem / 4;  // several contexts (serving slots) share the device
  BROOK_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

Context::~Context() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  trim();
  if (stream_) (void)hipStreamDestroy(stream_);
}

// Size classes: 1 KiB steps below 1 MiB, then four classes per octave (2^k times 1, 1.25, 1.5,
// 1.75), so buffers of volumes of other shapes are reused (at most 25 % above the request)
// without fresh hipMalloc/hipFree, each of which synchronizes the device.
size_t size_class(size_t bytes) {
  if (bytes == 0) bytes = 1;
  if (bytes < (size_t(1) << 20)) return (bytes + 1023) & ~size_t(1023);
  int k = 63 - __builtin_clzll((unsigned long long)bytes);  // 2^k <= bytes < 2^(k+1)
  const size_t step = size_t(1) << (k - 2);                  // a quarter octave
  return (bytes + step - 1) & ~(step - 1);
}

Block Context::allocate(size_t bytes) {
  const size_t size = size_class(bytes);
  // the smallest cached block of at least this class, if it is at most ~1.5 times the class
  auto it = cache_.lower_bound(size);
  if (it != cache_.end() && it->first <= size_class(size + size / 2)) {
    void *p = it->second;
    const size_t got = it->first;
    cache_.erase(it);
    cached_bytes_ -= got;
    return Block(this, p, got);
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, size);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    BROOK_HIP(hipMalloc(&p, size));
  }
  return Block(this, p, size);
}

void Context::give_back(void *ptr, size_t bytes) {
  cache_.emplace(bytes, ptr);
  cached_bytes_ += bytes;
  if (cached_bytes_ > cache_cap_) {  // free the largest blocks down to the cap
    (void)hipStreamSynchronize(stream_);
    while (cached_bytes_ > cache_cap_ && !cache_.empty()) {
      auto last = std::prev(cache_.end());
      (void)hipFree(last->second);
      cached_bytes_ -= last->first;
      cache_.erase(last);
    }
  }
}

Context &Context::aux(int i) {
  if (i < 0) throw std::invalid_argument("auxiliary context index");
  if (size_t(i) >= aux_.size()) aux_.resize(size_t(i) + 1);
  if (!aux_[size_t(i)]) aux_[size_t(i)] = std::make_unique<Context>(device_);
  return *aux_[size_t(i)];
}

void Context::trim() {
  for (auto &a : aux_)  // the auxiliary pools too, those no thread is using
    if (a && a->mutex().try_lock()) {
      a->trim();
      a->mutex().unlock();
    }
  if (!cache_.empty()) (void)hipStreamSynchronize(stream_);
  for (auto &kv : cache_) (void)hipFree(kv.second);
  cache_.clear();
  cached_bytes_ = 0;
}

}  // namespace brook
