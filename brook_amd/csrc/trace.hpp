// The per-label trace loop (target, railroad, backtrace, rails, invalidation).
#pragma once
#include <vector>

#include "stages.hpp"

namespace brook {

constexpr int kTraceProfileSlots = 16;

// The invalidation ball of a source s: radius scale * DBF[s] + constant.
struct Ball {
  float scale = 1.0f, constant = 0.0f;
};

struct TraceArgs {
  Dims d{0, 0, 0};
  VoxelRec *rec = nullptr;  // voxel records: PDRF (rails become 0 in place), flags, words (~0 between uses)
  const float *dbf = nullptr;
  const int *seg_start = nullptr, *seg_count = nullptr;
  const int *order = nullptr;  // per-label target order (voxels) in segment layout
  // manual targets: label L's are mflat[mstart[L] ..][0, mcount[L]) (before the DAF order), then
  // acount[L] more (after it; acount null: none)
  const int *mstart = nullptr, *mcount = nullptr, *mflat = nullptr, *acount = nullptr;
  int *maxp = nullptr;  // path budgets (a soma's is set on the device, after its ball)
  // int scratch in segment layout: lists 0, 1, 4, 5 hold one entry per voxel (segment seg_start[L]),
  // lists 2 and 3 (the railroad's far lists) two (segment 2 * seg_start[L])
  int *lists[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  unsigned long long *snaps[2] = {nullptr, nullptr};                       // u64 scratch, segment layout
  int *store = nullptr;
  long long store_cap = 0;
  int *rec_label = nullptr, *rec_len = nullptr, *rec_start = nullptr;
  int rec_cap = 0;
  // [0] store used [1] records [2] ejected labels [3] path store overflow; [4..5] u64
  // railroad-reached voxels, [6..7] u64 invalidated voxels; [8] failure bits (1: a railroad or an
  // invalidation did not converge within its round budget, 2: a far list overflowed)
  int *counters = nullptr;
  // per label (kTraceProfileSlots u64): wall-clock ticks total / target / railroad / backtrace /
  // invalidation, railroad rounds, invalidation rounds, paths, railroad phases, far entries
  // scanned, near entries expanded, invalidation entries expanded, waiting entries pulled, bands
  // (null: not recorded)
  unsigned long long *label_stats = nullptr;
  int *ejected = nullptr;
  int ejected_cap = 1;
  float wx = 1, wy = 1, wz = 1, delta = 1;
  Ball ball;  // the paths' invalidation ball (TEASAR scale and constant)
  // fix_branching=False: the parent field, built per label from `root` at the label's start (an
  // emitted vertex's entry becomes STOP, -2); paths are parent walks instead of railroads.
  // Null: fix_branching=True.
  int *parent = nullptr;
  const long long *root = nullptr;  // per label (both modes)
  // Somata, per label (null: none in the sample).
  const struct SomaLabel *soma = nullptr;
  // With a voxel graph, the pull masks (bit k of v: neighbour v + N[k] is of v's id and its word
  // permits the step to v) for backtraces, band-boundary pulls and parents; the records' masks are
  // then the push masks. Null: no graph (pulls read the records' masks).
  const unsigned *inmask = nullptr;
};

// A label's soma parameters: the trimming radius (< 0: not a soma), the soma's own ball and the
// path budget (< 0: the label's voxels minus those the soma ball took).
struct SomaLabel {
  double radius = -1.0;
  Ball ball;
  long long max_paths = -1;
};

// Traces every item (sample index into args, component id) with one workgroup per item, in item
// order, all in one launch; the soma labels among them (`somata`) lose their soma balls first.
void trace_all(Context &ctx, const std::vector<TraceArgs> &args, const std::vector<LabelItem> &items,
               const std::vector<LabelItem> &somata = {});

}  // namespace brook
