// Library radix sorts (hipCUB over rocPRIM) for the two sorting steps of the path: the per-label
// target order and the assembly's vertex/edge keys. Kept in one translation unit.
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

void sort_segments(Context &ctx, const unsigned long long *in, unsigned long long *out, long long total,
                   const int *seg_begin, const int *seg_end, int segments) {
  if (total <= 0 || segments <= 0) return;
  size_t bytes = 0;
  BROOK_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, bytes, in, out, int(total), segments, seg_begin, seg_end,
                                                       0, 64, ctx.stream()));
  Buffer<unsigned char> temp(ctx, std::max<size_t>(bytes, 1));
  BROOK_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(temp.data(), bytes, in, out, int(total), segments, seg_begin,
                                                       seg_end, 0, 64, ctx.stream()));
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
