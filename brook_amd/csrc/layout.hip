// Input layout normalization on the GPU: any byte-strided device volume (C order included) is
// gathered into the engine's Fortran order, and an object-id filter zeroes unlisted labels
// (runtime.cpp:106-150, layout.cu:11-61, skeletonize.cu:27-33).
#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace brook {
namespace {

template <class T>
__global__ void __launch_bounds__(256) gather(const char *base, long long s0, long long s1, long long s2, Dims d, T *out) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  long long n = (long long)d.sx * d.sy * d.sz;
  if (i >= n) return;
  long long x = i % d.sx, y = (i / d.sx) % d.sy, z = i / ((long long)d.sx * d.sy);
  out[i] = *reinterpret_cast<const T *>(base + x * s0 + y * s1 + z * s2);
}

// C order (z fastest) to Fortran order (x fastest): a transpose of x and z per y plane through a
// 64 x 64 LDS tile, so both the reads (along z) and the writes (along x) are coalesced.
template <class T>
__global__ void __launch_bounds__(256) transpose_xz(const T *in, Dims d, T *out) {
  __shared__ T tile[64][65];
  const int z0 = blockIdx.x * 64, x0 = blockIdx.y * 64, y = blockIdx.z;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {
    const int x = x0 + r, z = z0 + tx;
    if (x < d.sx && z < d.sz) tile[r][tx] = in[((long long)x * d.sy + y) * d.sz + z];
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int z = z0 + r, x = x0 + tx;
    if (x < d.sx && z < d.sz) out[x + (long long)d.sx * (y + (long long)d.sy * z)] = tile[tx][r];
  }
}

template <class T>
__global__ void __launch_bounds__(256) mask_ids(const T *in, long long n, const T *ids, int count, T *out) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  T v = in[i];
  int lo = 0, hi = count;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (ids[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  out[i] = (lo < count && ids[lo] == v) ? v : T(0);
}

template <class T>
void run_gather(Context &ctx, const void *base, const long long strides[3], Dims d, void *out) {
  long long n = (long long)d.sx * d.sy * d.sz;
  const long long b = sizeof(T);
  const bool c_order = (d.sz <= 1 || strides[2] == b) && (d.sy <= 1 || strides[1] == b * d.sz) &&
                       (d.sx <= 1 || strides[0] == b * d.sz * d.sy);
  if (c_order && (d.sx + 63) / 64 <= 65535 && d.sy <= 65535) {
    const dim3 grid(unsigned((d.sz + 63) / 64), unsigned((d.sx + 63) / 64), unsigned(d.sy));
    transpose_xz<T><<<grid, 256, 0, ctx.stream()>>>(static_cast<const T *>(base), d, static_cast<T *>(out));
    BROOK_LAUNCHED();
    return;
  }
  gather<T><<<unsigned((n + 255) / 256), 256, 0, ctx.stream()>>>(static_cast<const char *>(base), strides[0], strides[1],
                                                                  strides[2], d, static_cast<T *>(out));
  BROOK_LAUNCHED();
}

template <class T>
void run_mask(Context &ctx, const void *in, long long n, const std::vector<uint64_t> &ids, void *out) {
  std::vector<T> sorted;
  for (uint64_t v : ids) sorted.push_back(T(v));
  std::sort(sorted.begin(), sorted.end());
  sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
  Buffer<T> dev(ctx, std::max<size_t>(sorted.size(), 1));
  dev.upload(ctx, sorted.data(), sorted.size());
  mask_ids<T><<<unsigned((n + 255) / 256), 256, 0, ctx.stream()>>>(static_cast<const T *>(in), n, dev.data(),
                                                                    int(sorted.size()), static_cast<T *>(out));
  BROOK_LAUNCHED();
  ctx.synchronize();
}

}  // namespace

void gather_fortran(Context &ctx, const void *base, const long long byte_strides[3], int bytes, Dims d, void *out) {
  if ((long long)d.sx * d.sy * d.sz == 0) return;
  switch (bytes) {
    case 1: run_gather<uint8_t>(ctx, base, byte_strides, d, out); break;
    case 2: run_gather<uint16_t>(ctx, base, byte_strides, d, out); break;
    case 4: run_gather<uint32_t>(ctx, base, byte_strides, d, out); break;
    case 8: run_gather<unsigned long long>(ctx, base, byte_strides, d, out); break;
    default: throw std::invalid_argument("unsupported label width");
  }
}

void mask_labels(Context &ctx, const void *in, int bytes, long long n, const std::vector<uint64_t> &ids, void *out) {
  if (n == 0) return;
  switch (bytes) {
    case 1: run_mask<uint8_t>(ctx, in, n, ids, out); break;
    case 2: run_mask<uint16_t>(ctx, in, n, ids, out); break;
    case 4: run_mask<uint32_t>(ctx, in, n, ids, out); break;
    case 8: run_mask<unsigned long long>(ctx, in, n, ids, out); break;
    default: throw std::invalid_argument("unsupported label width");
  }
}

}  // namespace brook
