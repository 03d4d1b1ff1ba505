// Canonical postprocess of packed skeletons on the device: dust, loops, joins and ticks with the
// reference's tie rules (post.cu:421-651, post_kernels.cuh), plus the per-vertex components.
#pragma once
#include <array>
#include <cstdint>
#include <vector>

#include "assembly.hpp"
#include "merge.hpp"
#include "runtime.hpp"

namespace brook {

// Packed skeletons kept on the device between steps; offsets and labels on the host.
struct DevicePacked {
  std::vector<int64_t> labels, voff{0}, eoff{0};
  std::array<float, 3> anisotropy{1, 1, 1};
  Buffer<float> vertices, radii;  // 3 per vertex, 1 per vertex
  Buffer<uint32_t> edges;         // 2 per edge, local to the skeleton
  int64_t n() const { return voff.back(); }
  int64_t m() const { return eoff.back(); }
};

enum class PostStep { all = 0, dust = 1, loops = 2, join = 3, ticks = 4, prune = 5 };

DevicePacked upload_packed(Context &ctx, const PackedView &view);
Packed download_packed(Context &ctx, const DevicePacked &p);

// One step (dust and ticks take `a` as their threshold) or the whole postprocess (dust `a`,
// ticks `b`): remove_dust, remove_loops, join_components, remove_ticks, prune_unused.
DevicePacked postprocess(Context &ctx, DevicePacked p, PostStep step, double a = 0, double b = 0);

// Per vertex: the least vertex index of its connected component (post.cu:421-425).
std::vector<int64_t> skeleton_components(Context &ctx, const DevicePacked &p);

}  // namespace brook
