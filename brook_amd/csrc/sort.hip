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

template <class CC> struct Foreground {
  const CC *cc;
  __host__ __device__ bool operator()(int i) const { return cc[i] != CC(0); }
};

// Keys (id, descending DAF with inf and 0 as 0) and voxel values of the foreground voxels, laid
// out in descending flat index so that a stable sort leaves equal keys in descending index.
template <class CC>
__global__ void order_pairs(const CC *cc, const float *daf, const int *ascending, long long F, unsigned long long *keys,
                            int *vals) {
  const long long j = (long long)blockIdx.x * 256 + threadIdx.x;
  if (j >= F) return;
  const int i = ascending[F - 1 - j];
  float f = daf[i];
  if (isinf(f) || f == 0.0f) f = 0.0f;
  keys[j] = ((unsigned long long)cc[i] << 32) | (unsigned long long)(0xFFFFFFFFu - __float_as_uint(f));
  vals[j] = i;
}

template <class CC>
void run_target_order(Context &ctx, const CC *cc, long long n, const float *daf, long long F, int64_t count, int *order) {
  Buffer<int> ascending(ctx, size_t(F)), selected(ctx, 1);
  hipcub::CountingInputIterator<int> index(0);
  size_t bytes = 0;
  BROOK_HIP(hipcub::DeviceSelect::If(nullptr, bytes, index, ascending.data(), selected.data(), int(n), Foreground<CC>{cc},
                                     ctx.stream()));
  {
    Buffer<unsigned char> temp(ctx, std::max<size_t>(bytes, 1));
    BROOK_HIP(hipcub::DeviceSelect::If(temp.data(), bytes, index, ascending.data(), selected.data(), int(n),
                                       Foreground<CC>{cc}, ctx.stream()));
  }
  Buffer<unsigned long long> k0(ctx, size_t(F)), k1(ctx, size_t(F));
  Buffer<int> v1(ctx, size_t(F));
  order_pairs<CC><<<unsigned((F + 255) / 256), 256, 0, ctx.stream()>>>(cc, daf, ascending.data(), F, k0.data(), order);
  BROOK_LAUNCHED();
  int id_bits = 1;
  while (id_bits < 32 && (int64_t(1) << id_bits) <= count) ++id_bits;
  hipcub::DoubleBuffer<unsigned long long> dk(k0.data(), k1.data());
  hipcub::DoubleBuffer<int> dv(order, v1.data());
  bytes = 0;
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
