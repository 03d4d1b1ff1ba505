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
  // voxel_graph: one uint32 word per voxel on the device, Fortran order (null: none); a uint8
  // plane graph arrives widened value for value, with graph_compact_2d set
  const unsigned *graph = nullptr;
  bool graph_compact_2d = false;
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
// connected_components without device allocation: its working memory is a static device array of
// kNoAllocScratchBytes per device (ccl.hip). Device labels in Fortran order in, u32 ids to the
// caller's device buffer `ids` (n entries), mapping (capacity + 1) and roots (capacity) to host
// arrays. Returns the component count; when it exceeds `capacity` no table or id is written.
// Volumes whose working set (noalloc_scratch_need) exceeds the array are rejected. Calls on one
// device are serialized (each returns after its kernels finished).
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
void noalloc_scratch_init(int device);  // with `device` current: look up its copy of the array
size_t noalloc_scratch_need(int64_t voxels, int64_t capacity);
int64_t connected_components_noalloc(Context &ctx, const LabelVolume &input, uint32_t *ids, int64_t *mapping,
                                     int64_t *roots, int64_t capacity);
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
// The six faces of a volume on one canvas of `rows` x `cols` pixels (rows along the fast axis;
// the copy of each face is borders.cu:14-20's). Face f spans the volume's axes a[f] and b[f] at
// c[f] = 0 (f even) or c[f] = dims[c[f]] - 1 (f odd); its pixel (x along a, y along b) sits at
// row row0[f] + y of column col0[f] + x. Every other pixel is 0, and the layout leaves a zero row
// or column between any two faces, so no two faces touch.
struct FaceCanvas {
  int a[6], b[6], c[6], col0[6], row0[6];
  int rows, cols;
};
void extract_faces(Context &ctx, const void *cc, int cc_bytes, Dims dims, const FaceCanvas &fc, void *canvas);
// Faces f0 .. f1-1: their ids (`id_bytes` wide) and distances from the canvas, face after face,
// each face's pixel (x, y) at y + dims[b] * x.
void gather_faces(Context &ctx, const void *ids, int id_bytes, const float *dist, Dims dims, const FaceCanvas &fc,
                  int f0, int f1, void *face_ids, float *face_dist);

}  // namespace brook
