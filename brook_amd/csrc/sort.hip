// Library scans, selection and radix sorts (hipCUB over rocPRIM) for the path's sorting steps: the
// per-label target order and the assembly's vertex/edge keys. Kept in one translation unit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "assembly.hpp"
#include "fields.hpp"

namespace brook {

void exclusive_sum(Context &ctx, const int *in, int *out, long long n) {
  if (n <= 0) return;
  size_t bytes = 0;
  BROOK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, int(n), ctx.stream()));
  Buffer<unsigned char> temp(ctx, std::max<size_t>(bytes, 1));
  BROOK_HIP(hipcub::DeviceScan::ExclusiveSum(temp.data(), bytes, in, out, int(n), ctx.stream()));
}

void exclusive_sum(Context &ctx, const long long *in, long long *out, long long n) {
  if (n <= 0) return;
  size_t bytes = 0;
  BROOK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, int(n), ctx.stream()));
  Buffer<unsigned char> temp(ctx, std::max<size_t>(bytes, 1));
  BROOK_HIP(hipcub::DeviceScan::ExclusiveSum(temp.data(), bytes, in, out, int(n), ctx.stream()));
}

namespace {

// The foreground voxels in tiles of kTile (kPer consecutive voxels per thread): a count per tile,
// then, from the scanned counts, each tile's voxels in flat order.
constexpr int kPer = 16, kTile = 256 * kPer;

// Exclusive sum of v over the 256 threads of the block; *total gets the block's sum.
__device__ __forceinline__ int block_exclusive(int v, int *total) {
  __shared__ int wsum[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  int before = 0, all = 0;
  for (int w = 0; w < 4; ++w) {
    if (w < wave) before += wsum[w];
    all += wsum[w];
  }
  *total = all;
  return before + x - v;
}

template <class CC>
__global__ void __launch_bounds__(256) tile_counts(const CC *cc, long long n, int *counts) {
  const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kPer;
  int c = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) c += base + j < n && cc[base + j] != CC(0);
  int total = 0;
  block_exclusive(c, &total);
  if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// Keys (id, descending DAF with inf and 0 as 0) and voxel values of the foreground voxels, laid
// out in descending flat index so that a stable sort leaves equal keys in descending index.
template <class CC>
__global__ void __launch_bounds__(256) order_pairs(const CC *cc, const float *daf, long long n, long long F,
                                                   const int *tile_at, unsigned long long *keys, int *vals) {
  const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kPer;
  CC lab[kPer];
  int c = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    lab[j] = base + j < n ? cc[base + j] : CC(0);
    c += lab[j] != CC(0);
  }
  int total = 0;
  long long p = (long long)tile_at[blockIdx.x] + block_exclusive(c, &total);  // ascending position
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    if (lab[j] == CC(0)) continue;
    const int i = int(base + j);
    float f = daf[i];
    if (isinf(f) || f == 0.0f) f = 0.0f;
    const long long q = F - 1 - p++;
    if (q < 0) break;  // F counts every foreground voxel; never taken
    keys[q] = ((unsigned long long)lab[j] << 32) | (unsigned long long)(0xFFFFFFFFu - __float_as_uint(f));
    vals[q] = i;
  }
}

template <class CC>
void run_target_order(Context &ctx, const CC *cc, long long n, const float *daf, long long F, int64_t count, int *order) {
  const long long tiles = (n + kTile - 1) / kTile;
  Buffer<int> counts(ctx, size_t(tiles)), tile_at(ctx, size_t(tiles));
  tile_counts<CC><<<unsigned(tiles), 256, 0, ctx.stream()>>>(cc, n, counts.data());
  BROOK_LAUNCHED();
  exclusive_sum(ctx, counts.data(), tile_at.data(), tiles);
  Buffer<unsigned long long> k0(ctx, size_t(F)), k1(ctx, size_t(F));
  Buffer<int> v1(ctx, size_t(F));
  order_pairs<CC><<<unsigned(tiles), 256, 0, ctx.stream()>>>(cc, daf, n, F, tile_at.data(), k0.data(), order);
  BROOK_LAUNCHED();
  int id_bits = 1;
  while (id_bits < 32 && (int64_t(1) << id_bits) <= count) ++id_bits;
  hipcub::DoubleBuffer<unsigned long long> dk(k0.data(), k1.data());
  hipcub::DoubleBuffer<int> dv(order, v1.data());
  size_t bytes = 0;
  BROOK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, dk, dv, int(F), 0, 32 + id_bits, ctx.stream()));
  Buffer<unsigned char> temp(ctx, std::max<size_t>(bytes, 1));
  BROOK_HIP(hipcub::DeviceRadixSort::SortPairs(temp.data(), bytes, dk, dv, int(F), 0, 32 + id_bits, ctx.stream()));
  if (dv.Current() != order)
    BROOK_HIP(hipMemcpyAsync(order, dv.Current(), size_t(F) * 4, hipMemcpyDeviceToDevice, ctx.stream()));
}

}  // namespace

void target_order(Context &ctx, const void *cc, int cc_bytes, long long n, const float *daf, long long F, int64_t count,
                  int *order) {
  if (F <= 0) return;
  if (cc_bytes == 1) run_target_order(ctx, static_cast<const uint8_t *>(cc), n, daf, F, count, order);
  else if (cc_bytes == 2) run_target_order(ctx, static_cast<const uint16_t *>(cc), n, daf, F, count, order);
  else run_target_order(ctx, static_cast<const uint32_t *>(cc), n, daf, F, count, order);
}

void sort_keys(Context &ctx, unsigned long long *keys, unsigned long long *scratch, long long count, int end_bit) {
  if (count <= 1) return;
  hipcub::DoubleBuffer<unsigned long long> db(keys, scratch);
  size_t bytes = 0;
  BROOK_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, db, int(count), 0, end_bit, ctx.stream()));
  Buffer<unsigned char> temp(ctx, std::max<size_t>(bytes, 1));
  BROOK_HIP(hipcub::DeviceRadixSort::SortKeys(temp.data(), bytes, db, int(count), 0, end_bit, ctx.stream()));
  if (db.Current() != keys)
    BROOK_HIP(hipMemcpyAsync(keys, db.Current(), size_t(count) * 8, hipMemcpyDeviceToDevice, ctx.stream()));
}

}  // namespace brook
