// Voxel graphs: components of a graph's permitted steps (graph_components.cu), the graph-aware EDT
// (graph_edt.cu) and the graph gates of the per-voxel neighbour masks (the push and pull gates of
// common.cuh:84-100).
#pragma once
#include "fields.hpp"
#include "stages.hpp"

namespace brook {

struct GraphComponents {
  Components cc;                // ids, mapping (label at each id's last run start), roots (lowest foreground voxel)
  std::vector<int64_t> keys;    // [id]: 1-based rank of the id's component among all of them ([0] = 0)
  bool input_uniform = true;    // every voxel of the input has the first voxel's label
};

// `graph`: one uint32 word per voxel, Fortran order, on the device (a uint8 plane graph widened value
// for value; compact_2d then moves its corner bits back for the components).
GraphComponents graph_components(Context &ctx, const LabelVolume &labels, const unsigned *graph, bool compact_2d);
// EDT of the foreground of `cc` on the doubled grid the graph defines; `label_max` (count + 1 zeroed
// entries) receives per-id maxima (float bits) when given. Volumes up to 2^28 voxels.
void graph_edt(Context &ctx, const void *cc, int cc_bytes, Dims dims, const unsigned *graph, const float aniso[3],
               bool black_border, float *out, unsigned *label_max = nullptr);
// A uint8 plane graph widened to one uint32 word per voxel, value for value.
void widen_words(Context &ctx, const uint8_t *in, long long n, unsigned *out);
// Gates the records' same-label masks by the graph (pushes) and writes the pull masks (n words).
void graph_gates(Context &ctx, VoxelRec *rec, const unsigned *graph, Dims dims, unsigned *in_masks);

}  // namespace brook
