// Packed skeletons and the on-device assembly of traced paths.
#pragma once
#include <array>
#include <cstdint>
#include <vector>

#include "stages.hpp"

namespace brook {

// Skeletons in flat host arrays (the reference's packed layout, packed.hpp:7-11): skeleton k has
// label labels[k], vertices [v_off[k], v_off[k+1]) (x,y,z float32, physical), the same range of
// radii, and edges [e_off[k], e_off[k+1]) (uint32 pairs local to the skeleton).
struct Packed {
  std::vector<int64_t> labels;
  std::vector<long long> v_off{0}, e_off{0};
  std::vector<float> vertices, radii;
  std::vector<uint32_t> edges;
  std::array<float, 3> anisotropy{1, 1, 1};
};

struct AssemblyInput {
  const void *cc = nullptr;
  Dims d{0, 0, 0};
  const float *dbf = nullptr;
  const int *store = nullptr;
  long long store_used = 0;
  const int *rec_start = nullptr, *rec_len = nullptr;
  int records = 0;
  std::vector<int> group_of;            // per component id: output group, -1 none
  std::vector<int64_t> group_labels;    // per group: input label
  float aniso[3] = {1, 1, 1};
  int origin[3] = {0, 0, 0};  // added to the voxel coordinates before the anisotropy (a private arena's box)
};

Packed assemble(Context &ctx, int cc_bytes, const AssemblyInput &in);

}  // namespace brook
