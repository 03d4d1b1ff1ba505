// Soma preparation of the batched path (batched.cu:58-108, 186-231; fill_voids' count,
// fillvoids_kernels.cuh): the enclosed-void check of a label's bounding box and the candidates of
// the soma roots.
#pragma once
#include <cstdint>
#include <utility>
#include <vector>

#include "stages.hpp"

namespace brook {

// Voxels of id's bounding box that are not id and are not 6-connected through such voxels to a
// face of the box (the voids fill_voids would fill).
int64_t enclosed_voids(Context &ctx, const void *cc, int cc_bytes, Dims d, int64_t id, const Box &box);

// The voids of id's box, found once and then applied (soma.cu:141-237 for a soma without private
// arenas): taken()[o] = voxels of id o inside the voids; paint() writes id into every void voxel;
// dbf_refresh() replaces the DBF of id's filled box by that box's own EDT (black border when the
// filled mask fills the box; with a voxel graph, the graph EDT of the box's words) and returns its
// maximum.
class VoidFill {
 public:
  VoidFill(Context &ctx, void *cc, int cc_bytes, Dims d, int64_t id, const Box &box, size_t ids);
  int64_t voids() const { return voids_; }
  const std::vector<int64_t> &taken() const { return taken_; }
  void paint(Context &ctx);
  // `filled_count`: the label's voxels after paint() (the box's black border when it is the box)
  float dbf_refresh(Context &ctx, const float aniso[3], float *dbf, int64_t filled_count, const unsigned *graph = nullptr);
  // The box as a volume of its own (a private arena, soma.cu:54-72): the mask (1 = id or a void,
  // Fortran order over the box, moved out: paint() and dbf_refresh() need it) and the graph's words.
  Buffer<uint8_t> take_mask() { return std::move(mask_); }
  Buffer<unsigned> crop_graph(Context &ctx, const unsigned *graph) const;
  const int *lo() const { return lo_; }
  const int *ext() const { return ext_; }

 private:
  void *cc_;
  int cc_bytes_;
  Dims d_;
  int64_t id_;
  int lo_[3], ext_[3];
  int64_t voids_ = 0;
  std::vector<int64_t> taken_;
  Buffer<uint8_t> mask_;  // per box voxel: 1 = id or void
};

// Per id (count + 1 entries): its smallest flat index (-1: no voxel).
std::vector<int64_t> first_voxels(Context &ctx, const void *cc, int cc_bytes, long long n, int64_t count);

// Every voxel v (ascending flat index) of an id with soma[id] != 0 whose DBF equals dbf_max[id]
// (device tables of count + 1 entries), with its id.
struct Maximum {
  int64_t voxel, id;
};
std::vector<Maximum> soma_maxima(Context &ctx, const void *cc, int cc_bytes, long long n, const float *dbf,
                                 const float *dbf_max, const uint8_t *soma);

}  // namespace brook
