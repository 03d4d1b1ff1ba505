// Soma preparation of the batched path (batched.cu:58-108, 186-231; fill_voids' count,
// fillvoids_kernels.cuh): the enclosed-void check of a label's bounding box and the candidates of
// the soma roots.
#pragma once
#include <cstdint>
#include <vector>

#include "stages.hpp"

namespace brook {

// Voxels of id's bounding box that are not id and are not 6-connected through such voxels to a
// face of the box (the voids fill_voids would fill).
int64_t enclosed_voids(Context &ctx, const void *cc, int cc_bytes, Dims d, int64_t id, const Box &box);

// Every voxel v (ascending flat index) of an id with soma[id] != 0 whose DBF equals dbf_max[id]
// (device tables of count + 1 entries), with its id.
struct Maximum {
  int64_t voxel, id;
};
std::vector<Maximum> soma_maxima(Context &ctx, const void *cc, int cc_bytes, long long n, const float *dbf,
                                 const float *dbf_max, const uint8_t *soma);

}  // namespace brook
