// The skeletonize pipeline (skeletonize.cu:70-201 restated for one GPU).
#pragma once
#include <vector>

#include "assembly.hpp"
#include "borders.hpp"
#include "stages.hpp"

namespace brook {

struct TeasarParams {
  // kimimaro.trace defaults (skeleton.hpp:24-25) for keys a caller leaves out
  double scale = 10, constant = 10, pdrf_scale = 5000, pdrf_exponent = 16;
  double soma_detection = 1100, soma_acceptance = 4000, soma_scale = 0.5, soma_constant = 0;
  int64_t max_paths = -1;  // < 0: none
};

struct SkeletonizeOptions {
  TeasarParams teasar;
  float aniso[3] = {1, 1, 1};
  double dust = 1000;
  bool fix_branching = true, fix_borders = true, fill_holes = false, fix_avocados = false, voxel_graph = false;
  std::vector<Point> before, after;
};

// Fields of the preamble, for stage-level checks.
struct FieldsOut {
  float *dbf = nullptr, *daf = nullptr, *pdrf = nullptr;  // host, n floats each (may be null)
  std::vector<int64_t> root, target;
  std::vector<float> dbf_max, max_daf;
  std::vector<uint8_t> active;
  int64_t count = 0;
  bool stop = false;  // return right after the fields are captured
};

Packed skeletonize(Context &ctx, const LabelVolume &labels, const SkeletonizeOptions &options, FieldsOut *fields = nullptr);
// Independent volumes with shared options: result[i] equals skeletonize(volumes[i]). Volumes are
// run in groups that fit half of the free device memory; a group's labels share one flood launch
// and one trace launch.
std::vector<Packed> skeletonize_batch(Context &ctx, const std::vector<LabelVolume> &volumes,
                                      const SkeletonizeOptions &options);
// The tracer's target orders of every nonzero id of a component volume (the engine's sort on the
// device), concatenated in ascending id: flat indices, one per nonzero voxel. daf: host, n floats.
std::vector<int> target_orders(Context &ctx, const LabelVolume &ids, const float *daf);

}  // namespace brook
