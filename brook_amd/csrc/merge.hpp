// Merging packed skeleton results (the reference's merge_fragments, packed.cpp:78-86).
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "assembly.hpp"

namespace brook {

// A borrowed packed result (the layout of Packed / brook_skeletons_view): labels and offsets in
// host memory, vertices, edges and radii in host memory or on the context's device.
struct PackedView {
  size_t skeletons = 0;
  const int64_t *labels = nullptr;
  const int64_t *v_off = nullptr, *e_off = nullptr;  // skeletons + 1 entries each
  const float *vertices = nullptr;                    // x, y, z per vertex
  const uint32_t *edges = nullptr;                    // local vertex pairs
  const float *radii = nullptr;
  std::array<float, 3> anisotropy{1, 1, 1};
};

// One packed result from the fragments, fragment i's vertices offset by origins[i] (packed_merge.cu:
// 88-166 with sorted labels, empty labels kept); the anisotropy is the first fragment's.
Packed merge_packed(Context &ctx, const std::vector<PackedView> &fragments, const std::vector<std::array<float, 3>> &origins);

}  // namespace brook
