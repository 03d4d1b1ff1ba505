// Stage entry points of the skeletonization engine (host side of each HIP stage).
#pragma once
#include <array>
#include <cstdint>
#include <vector>

#include "common.hpp"
#include "runtime.hpp"

namespace brook {

// A device label volume in Fortran order (x fastest). `bytes` is the element width; signedness
// never matters on the path (labels are only compared for equality and with zero).
struct LabelVolume {
  const void *data = nullptr;
  int bytes = 4;
  bool is_signed = false;  // only changes how mapping values are widened to int64
  Dims dims{0, 0, 0};
  int64_t size() const { return int64_t(dims.sx) * dims.sy * dims.sz; }
};

struct Components {
  Buffer<uint8_t> labels;  // cc ids, Fortran order, `cc_bytes` wide (1, 2 or 4)
  int cc_bytes = 1;
  int64_t count = 0;
  std::vector<int64_t> mapping;  // [0] = 0, [id] = input value
  std::vector<int64_t> roots;    // [id - 1] = minimum flat index of the component
};

// One component label of one sample of a group (sample index, component id, voxel count).
struct LabelItem {
  int sample = 0, label = 0;
  int64_t size = 0;
};

struct Box {
  int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {-1, -1, -1};
  int64_t count = 0;
};

// Smallest id width that holds `count` ids (the reference's u8/u16/u32 rule, ccl.cu:57-60).
inline int id_bytes(int64_t count) { return count <= 255 ? 1 : (count <= 65535 ? 2 : 4); }

Components connected_components(Context &ctx, const LabelVolume &input);
// Per-id bounding boxes and voxel counts (index 0: background count).
std::vector<Box> analyze(Context &ctx, const Components &cc, Dims dims);
// Multi-label anisotropic EDT of cc ids with the reference's rounding rule; background 0.
// dims_used 2 treats the volume as a stack of independent xy planes (sz must be 1).
// When `label_max` is given (count+1 entries, zeroed), it receives the per-id maximum (float bits).
void edt(Context &ctx, const void *cc, int cc_bytes, Dims dims, const float aniso[3], bool black_border, int dims_used,
         float *out, unsigned *label_max = nullptr);
// Exclusive prefix sums on the context's stream (int and 64-bit element forms).
void exclusive_sum(Context &ctx, const int *in, int *out, long long n);
void exclusive_sum(Context &ctx, const long long *in, long long *out, long long n);
// Label-dependent copy of one volume face into a plane of shape (fy, fx, 1) (borders.cu:14-20).
void extract_face(Context &ctx, const void *cc, int cc_bytes, Dims dims, int a, int b, int c, int fixed, void *plane);

}  // namespace brook
