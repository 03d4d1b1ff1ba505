// Per-label field preparation and the tracer's device tables.
#pragma once
#include <vector>

#include "stages.hpp"

namespace brook {

// Device tables indexed by component id (count + 1 entries).
struct LabelTables {
  const int *seg_start = nullptr;  // first entry of the id's voxel list
  const int *seg_count = nullptr;  // voxels of the id
  const int *vox = nullptr;        // voxel lists (unordered within an id)
  const long long *first = nullptr;  // the id's minimum flat index (root search source)
  long long *root = nullptr;       // in: border root or -1; out: root
  long long *target = nullptr;     // out: DAF argmax (lowest flat index on ties)
  float *max_daf = nullptr;        // out: DAF maximum
  const float *weights = nullptr;  // 26 geometric step weights
  // > 0: the id is a soma; its DAF flood starts from the free-space ball of this radius around
  // the root (null: no somata)
  const float *free_radius = nullptr;
  VoxelRec *rec = nullptr;         // voxel records (masks in; words ~0 in and out)
};

// rec[v]: neighbour mask (bit k = neighbour k of v is inside the volume and has v's nonzero id),
// PDRF +inf, word ~0, no flags.
void voxel_records(Context &ctx, const void *cc, int cc_bytes, Dims d, VoxelRec *rec);

void build_segments(Context &ctx, const void *cc, int cc_bytes, Dims d, const std::vector<int> &seg_start,
                    Buffer<int> &seg_start_dev, Buffer<int> &vox);
void fill_value(Context &ctx, float *p, long long n, float v);
// One sample's flood inputs: t.rec words ~0 in and out; daf receives the DAF on the traced ids'
// voxels; lists_*: scratch in segment layout.
struct FloodJob {
  Dims d{0, 0, 0};
  LabelTables t;
  float *daf = nullptr;
  int *lists_a = nullptr, *lists_b = nullptr;  // one entry per voxel (segment seg_start[l])
  int *far_a = nullptr, *far_b = nullptr;      // two entries per voxel (segment 2 * seg_start[l])
  int root_only = 0;  // 1: the root search alone (t.root out; no DAF flood, daf untouched)
};
// Root search and DAF floods of every item (one workgroup per item, in item order), at most
// max_grid workgroups at once (0: two per compute unit).
void prepare_fields(Context &ctx, const std::vector<FloodJob> &jobs, const std::vector<LabelItem> &items,
                    int max_grid = 0);
void pdrf_and_alive(Context &ctx, const void *cc, int cc_bytes, long long n, const float *dbf, const float *daf,
                    const float *M, const float *inv, const unsigned char *use, int squarings, float scale, float *pdrf,
                    VoxelRec *rec);
void root_rails(Context &ctx, VoxelRec *rec, const long long *root, const int *labels, int count);
// Per-id target order in segment layout: order[seg_start[id] + j] = the id's voxels by descending
// DAF (inf and 0 as 0), ties by descending flat index (lockstep_order.cu:21-43); F foreground voxels.
void target_order(Context &ctx, const void *cc, int cc_bytes, long long n, const float *daf, long long F, int64_t count,
                  int *order);
// Ascending sort of 64-bit keys over [0, end_bit).
void sort_keys(Context &ctx, unsigned long long *keys, unsigned long long *scratch, long long count, int end_bit);

}  // namespace brook
