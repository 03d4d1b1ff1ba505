#include "runtime.hpp"

namespace brook {

void Block::release() {
  if (ptr_ && ctx_) ctx_->give_back(ptr_, bytes_);
  ptr_ = nullptr;
  ctx_ = nullptr;
  bytes_ = 0;
}

Context::Context(int device) : device_(device) {
  int count = 0;
  BROOK_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) throw std::invalid_argument("device index out of range");
  BROOK_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  BROOK_HIP(hipGetDeviceProperties(&prop, device));
  cus_ = prop.multiProcessorCount;
  BROOK_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

Context::~Context() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  trim();
  if (stream_) (void)hipStreamDestroy(stream_);
}

// Sizes round up to 2 MiB multiples (1 KiB below 1 MiB), so repeated calls of one shape reuse
// their blocks without fresh hipMalloc/hipFree (each of which synchronizes the device).
static size_t rounded(size_t bytes) {
  if (bytes == 0) bytes = 1;
  if (bytes < (size_t(1) << 20)) return (bytes + 1023) & ~size_t(1023);
  return (bytes + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
}

Block Context::allocate(size_t bytes) {
  size_t size = rounded(bytes);
  auto it = cache_.find(size);
  if (it != cache_.end()) {
    void *p = it->second;
    cache_.erase(it);
    cached_bytes_ -= size;
    return Block(this, p, size);
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
}

void Context::trim() {
  if (!cache_.empty()) (void)hipStreamSynchronize(stream_);
  for (auto &kv : cache_) (void)hipFree(kv.second);
  cache_.clear();
  cached_bytes_ = 0;
}

}  // namespace brook
