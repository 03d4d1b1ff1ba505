#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

#include "fields.hpp"
#include "trace.hpp"

namespace brook {
namespace {

using Clock = std::chrono::steady_clock;

struct Lap {
  Context &ctx;
  Clock::time_point t = Clock::now();
  void operator()(const char *key) {
    ctx.synchronize();
    auto now = Clock::now();
    ctx.stats[key] += std::chrono::duration<double, std::milli>(now - t).count();
    t = now;
  }
};

int64_t flat_of(const Point &p, Dims d) { return p[0] + int64_t(d.sx) * (p[1] + int64_t(d.sy) * p[2]); }

// The component id under a manual target point; negative coordinates count from the far edge for
// the lookup only (skeletonize.cu:16-24).
int64_t component_at(Context &ctx, const Components &cc, Dims d, Point p) {
  const int dims[3] = {d.sx, d.sy, d.sz};
  for (int a = 0; a < 3; ++a) {
    if (p[a] < 0) p[a] += dims[a];
    if (p[a] < 0 || p[a] >= dims[a]) throw std::invalid_argument("extra target is outside the input volume");
  }
  uint64_t value = 0;
  const char *src = reinterpret_cast<const char *>(cc.labels.data()) + flat_of(p, d) * cc.cc_bytes;
  BROOK_HIP(hipMemcpyAsync(&value, src, cc.cc_bytes, hipMemcpyDeviceToHost, ctx.stream()));
  ctx.synchronize();
  return int64_t(value);
}

Packed empty_result(const SkeletonizeOptions &o) {
  Packed p;
  p.anisotropy = {o.aniso[0], o.aniso[1], o.aniso[2]};
  return p;
}

}  // namespace

Packed skeletonize(Context &ctx, const LabelVolume &in, const SkeletonizeOptions &o, FieldsOut *fields) {
  ctx.stats.clear();
  Lap lap{ctx};
  const Dims d = in.dims;
  const int64_t n = in.size();
  for (float a : o.aniso)
    if (!(a > 0) || !std::isfinite(a)) throw std::invalid_argument("anisotropy must be finite and positive");
  if (n == 0 || double(n) <= o.dust) return empty_result(o);
  if (!o.fix_branching) throw Unsupported("fix_branching=False is not supported by this build");
  if (o.fill_holes || o.fix_avocados || o.voxel_graph)
    throw Unsupported("fill_holes, fix_avocados and voxel_graph are not supported by this build");
  if (!o.after.empty()) throw Unsupported("extra_targets_after is not supported by this build");
  const double exponent = o.teasar.pdrf_exponent;
  if (std::isfinite(exponent) && exponent == std::floor(exponent) && std::fmod(exponent, 2.0) == 1.0)
    throw std::invalid_argument("odd pdrf_exponent makes the background penalty negative");
  if (!(exponent > 0 && exponent < 65536 && exponent == std::floor(exponent) && !(int(exponent) & (int(exponent) - 1))))
    throw Unsupported("pdrf_exponent must be a power of two (1..32768) in this build");

  // ---- components, analysis, distance field -------------------------------------------------
  Components cc = connected_components(ctx, in);
  lap("prep_ccl_ms");
  const int64_t count = cc.count;
  if (count == 0) return empty_result(o);
  const size_t k = size_t(count) + 1;
  std::vector<Box> boxes = analyze(ctx, cc, d);
  lap("prep_analyze_ms");
  const bool black_border = count == 1 && boxes[1].count == n;
  Buffer<float> dbf(ctx, size_t(n));
  Buffer<unsigned> label_max(ctx, k);
  label_max.clear(ctx);
  edt(ctx, cc.labels.data(), cc.cc_bytes, d, o.aniso, black_border, 3, dbf.data(), label_max.data());
  std::vector<unsigned> max_bits = label_max.download(ctx);
  std::vector<float> dbf_max(k);
  for (size_t i = 0; i < k; ++i) std::memcpy(&dbf_max[i], &max_bits[i], 4);
  lap("prep_edt_ms");
  std::vector<std::vector<Point>> borders(k);
  if (o.fix_borders) borders = border_targets(ctx, cc.labels.data(), cc.cc_bytes, d, count, o.aniso);
  lap("prep_borders_ms");
  std::vector<std::vector<Point>> before(k);
  for (const Point &p : o.before) before.at(size_t(component_at(ctx, cc, d, p))).push_back(p);

  // ---- labels to trace (skeletonize.cu:128-133, batched.cu:186-203) ---------------------------
  std::vector<uint8_t> active(k, 0);
  std::vector<int> labels;
  for (size_t id = 1; id < k; ++id) {
    const Box &b = boxes[id];
    int64_t ext = int64_t(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
    if (double(b.count) <= o.dust || ext <= 1) continue;
    if (double(dbf_max[id]) > o.teasar.soma_detection)
      throw Unsupported("a component exceeds soma_detection_threshold; soma processing is not supported by this build");
    active[id] = 1;
    labels.push_back(int(id));
  }
  std::vector<int> seg_start(k + 1, 0), seg_count(k, 0);
  for (size_t id = 1; id < k; ++id) seg_count[id] = int(boxes[id].count);
  for (size_t id = 0; id < k; ++id) seg_start[id + 1] = seg_start[id] + seg_count[id];
  const int64_t F = seg_start[k];
  Buffer<int> seg_start_dev, vox;
  build_segments(ctx, cc.labels.data(), cc.cc_bytes, d, seg_start, seg_start_dev, vox);
  Buffer<int> seg_count_dev(ctx, k);
  seg_count_dev.upload(ctx, seg_count.data(), k);

  // ---- roots, DAF, targets -------------------------------------------------------------------
  std::vector<long long> first(k, -1), root(k, -1);
  for (size_t id = 1; id < k; ++id) first[id] = cc.roots[id - 1];
  for (int id : labels)
    if (!borders[id].empty()) root[id] = flat_of(borders[id].back(), d);
  int64_t flood_voxels = 0;  // voxels swept by the root-search and DAF floods
  for (int id : labels) flood_voxels += boxes[id].count * (root[id] < 0 ? 2 : 1);
  ctx.stats["flood_voxels"] = double(flood_voxels);
  ctx.stats["foreground_voxels"] = double(F);
  ctx.stats["cc_bytes"] = cc.cc_bytes;
  ctx.stats["components"] = double(count);
  std::vector<int> queue = labels;
  std::stable_sort(queue.begin(), queue.end(), [&](int a, int b) { return boxes[a].count > boxes[b].count; });
  Buffer<long long> first_dev(ctx, k), root_dev(ctx, k), target_dev(ctx, k);
  Buffer<float> max_daf_dev(ctx, k);
  first_dev.upload(ctx, first.data(), k);
  root_dev.upload(ctx, root.data(), k);
  Buffer<int> queue_dev(ctx, std::max<size_t>(queue.size(), 1));
  queue_dev.upload(ctx, queue.data(), queue.size());
  const double ax = o.aniso[0], ay = o.aniso[1], az = o.aniso[2];
  std::vector<float> W = {float(ax), float(ax), float(ay), float(ay), float(az), float(az)};
  for (double e : {std::sqrt(ax * ax + ay * ay), std::sqrt(ay * ay + az * az), std::sqrt(ax * ax + az * az)})
    for (int i = 0; i < 4; ++i) W.push_back(float(e));
  for (int i = 0; i < 8; ++i) W.push_back(float(std::sqrt(ax * ax + ay * ay + az * az)));
  Buffer<float> weights(ctx, 26);
  weights.upload(ctx, W.data(), 26);
  Buffer<float> daf(ctx, size_t(n));
  fill_value(ctx, daf.data(), n, kInf);
  Buffer<VoxelRec> rec(ctx, size_t(n));  // 16 bytes per voxel: PDRF, neighbour mask + flags, distance word
  voxel_records(ctx, cc.labels.data(), cc.cc_bytes, d, rec.data());
  std::vector<Buffer<int>> lists;
  for (int i = 0; i < 6; ++i) lists.emplace_back(ctx, size_t(std::max<int64_t>(F, 1)));
  LabelTables tables;
  tables.seg_start = seg_start_dev.data();
  tables.seg_count = seg_count_dev.data();
  tables.vox = vox.data();
  tables.first = first_dev.data();
  tables.root = root_dev.data();
  tables.target = target_dev.data();
  tables.max_daf = max_daf_dev.data();
  tables.weights = weights.data();
  tables.rec = rec.data();
  prepare_fields(ctx, d, tables, queue_dev.data(), int(queue.size()), daf.data(), lists[0].data(), lists[1].data());
  root = root_dev.download(ctx);
  std::vector<long long> target = target_dev.download(ctx);
  std::vector<float> max_daf = max_daf_dev.download(ctx);
  lap("prep_floods_ms");

  // ---- PDRF (exponent a power of two: squarings), alive mask, rails, target order -------------
  std::vector<float> M(k, 0.0f), inv(k, 0.0f);
  std::vector<uint8_t> use(k, 0);
  for (int id : labels) {
    M[id] = float(1.0 / std::pow(double(dbf_max[id]), 1.01));
    if (max_daf[id] != 0) { inv[id] = float(1.0 / double(max_daf[id])); use[id] = 2; }
    else use[id] = 1;
  }
  int squarings = 0;
  for (int e = int(exponent); e > 1; e >>= 1) ++squarings;
  Buffer<float> M_dev(ctx, k), inv_dev(ctx, k);
  Buffer<uint8_t> use_dev(ctx, k);
  M_dev.upload(ctx, M.data(), k);
  inv_dev.upload(ctx, inv.data(), k);
  use_dev.upload(ctx, use.data(), k);
  Buffer<float> pdrf(ctx, size_t(n));
  pdrf_and_alive(ctx, cc.labels.data(), cc.cc_bytes, n, dbf.data(), daf.data(), M_dev.data(), inv_dev.data(),
                 use_dev.data(), squarings, float(o.teasar.pdrf_scale), pdrf.data(), rec.data());
  Buffer<int> labels_dev(ctx, std::max<size_t>(labels.size(), 1));
  labels_dev.upload(ctx, labels.data(), labels.size());
  if (fields) {
    fields->count = count;
    fields->root.assign(root.begin(), root.end());
    fields->target.assign(target.begin(), target.end());
    fields->dbf_max = dbf_max;
    fields->max_daf = max_daf;
    fields->active = active;
    ctx.synchronize();
    if (fields->dbf) BROOK_HIP(hipMemcpy(fields->dbf, dbf.data(), size_t(n) * 4, hipMemcpyDeviceToHost));
    if (fields->daf) BROOK_HIP(hipMemcpy(fields->daf, daf.data(), size_t(n) * 4, hipMemcpyDeviceToHost));
    if (fields->pdrf) BROOK_HIP(hipMemcpy(fields->pdrf, pdrf.data(), size_t(n) * 4, hipMemcpyDeviceToHost));
    if (fields->stop) return empty_result(o);
  }
  root_rails(ctx, rec.data(), root_dev.data(), labels_dev.data(), int(labels.size()));
  Buffer<unsigned long long> order(ctx, size_t(std::max<int64_t>(F, 1))), snap_a(ctx, size_t(std::max<int64_t>(F, 1)));
  target_order_keys(ctx, vox.data(), F, daf.data(), snap_a.data());
  sort_segments(ctx, snap_a.data(), order.data(), F, seg_start_dev.data(), seg_start_dev.data() + 1, int(k));
  lap("prep_pdrf_order_ms");

  // ---- manual targets and path budgets (lockstep.cu:116-143) ----------------------------------
  std::vector<int> mstart(k, 0), mcount(k, 0), mflat, maxp(k, 0);
  for (int id : labels) {
    std::vector<Point> manual = borders[id];
    if (!manual.empty()) manual.pop_back();
    manual.insert(manual.end(), before[id].begin(), before[id].end());
    mstart[id] = int(mflat.size());
    if (manual.empty()) mflat.push_back(int(target[id]));
    else for (auto it = manual.rbegin(); it != manual.rend(); ++it) mflat.push_back(int(flat_of(*it, d)));
    mcount[id] = int(mflat.size()) - mstart[id];
    int64_t budget = o.teasar.max_paths >= 0 ? o.teasar.max_paths : boxes[id].count;
    maxp[id] = mcount[id] >= budget ? 0 : int(std::min<int64_t>(INT32_MAX, budget));
  }
  mflat.push_back(0);
  Buffer<int> mstart_dev(ctx, k), mcount_dev(ctx, k), maxp_dev(ctx, k), mflat_dev(ctx, mflat.size());
  mstart_dev.upload(ctx, mstart.data(), k);
  mcount_dev.upload(ctx, mcount.data(), k);
  maxp_dev.upload(ctx, maxp.data(), k);
  mflat_dev.upload(ctx, mflat.data(), mflat.size());

  // ---- trace -----------------------------------------------------------------------------------
  Buffer<unsigned long long> snap_b(ctx, size_t(std::max<int64_t>(F, 1)));
  const long long store_cap = std::max<long long>(int64_t(1) << 22, 4 * F);
  const int rec_cap = int(std::max<int64_t>(int64_t(1) << 16, F / 8 + 4 * int64_t(labels.size())));
  Buffer<int> store(ctx, size_t(store_cap)), rec_label(ctx, rec_cap), rec_len(ctx, rec_cap), rec_start(ctx, rec_cap);
  Buffer<int> counters(ctx, 8), ejected(ctx, 64), next(ctx, 1);
  counters.clear(ctx);
  next.clear(ctx);
  TraceArgs ta;
  ta.d = d;
  ta.rec = rec.data();
  ta.dbf = dbf.data();
  ta.seg_start = seg_start_dev.data();
  ta.seg_count = seg_count_dev.data();
  ta.order = order.data();
  ta.mstart = mstart_dev.data();
  ta.mcount = mcount_dev.data();
  ta.mflat = mflat_dev.data();
  ta.maxp = maxp_dev.data();
  for (int i = 0; i < 6; ++i) ta.lists[i] = lists[i].data();
  ta.snaps[0] = snap_a.data();
  ta.snaps[1] = snap_b.data();
  ta.store = store.data();
  ta.store_cap = store_cap;
  ta.rec_label = rec_label.data();
  ta.rec_len = rec_len.data();
  ta.rec_start = rec_start.data();
  ta.rec_cap = rec_cap;
  ta.counters = counters.data();
  Buffer<unsigned long long> label_stats(ctx, size_t(kTraceProfileSlots) * k);
  label_stats.clear(ctx);
  ta.label_stats = label_stats.data();
  ta.ejected = ejected.data();
  ta.ejected_cap = 64;
  ta.queue = queue_dev.data();
  ta.nqueue = int(queue.size());
  ta.next = next.data();
  ta.wx = o.aniso[0];
  ta.wy = o.aniso[1];
  ta.wz = o.aniso[2];
  ta.scale = float(o.teasar.scale);
  ta.constant = float(o.teasar.constant);
  ta.delta = std::min({o.aniso[0], o.aniso[1], o.aniso[2]});
  trace_all(ctx, cc.cc_bytes, ta);
  std::vector<int> cnt = counters.download(ctx);
  lap("trace_ms");
  if (cnt[3]) throw std::runtime_error("path store capacity exceeded");
  if (cnt[2]) throw Unsupported("a label has no reachable rail (the per-component fallback tracer is not part of this build)");
  ctx.stats["paths"] = cnt[1];
  ctx.stats["path_vertices"] = cnt[0];
  uint64_t reached = 0, claimed = 0;
  std::memcpy(&reached, &cnt[4], 8);
  std::memcpy(&claimed, &cnt[6], 8);
  ctx.stats["railroad_reached_voxels"] = double(reached);
  ctx.stats["invalidated_voxels"] = double(claimed);
  {  // per-label trace profile (device wall clock)
    int khz = 0;
    BROOK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx.device()));
    const double tick_ms = khz > 0 ? 1.0 / khz : 0.0;
    std::vector<unsigned long long> ls = label_stats.download(ctx);
    const size_t P = kTraceProfileSlots;
    double sums[kTraceProfileSlots] = {0}, worst = 0;
    size_t worst_id = 0;
    for (int id : labels) {
      for (size_t j = 0; j < P; ++j) sums[j] += double(ls[P * size_t(id) + j]);
      if (double(ls[P * size_t(id)]) > worst) { worst = double(ls[P * size_t(id)]); worst_id = size_t(id); }
    }
    const char *work[6] = {"railroad_phases", "far_scanned", "near_expanded", "invalidate_expanded",
                           "waiting_pulled", "bands"};
    for (int j = 0; j < 6; ++j) {
      ctx.stats[std::string(work[j]) + "_per_path"] = sums[8 + j] / std::max(1.0, sums[7]);
      ctx.stats[std::string("slowest_") + work[j]] = double(ls[P * worst_id + 8 + j]);
    }
    const double nl = std::max<double>(1.0, double(labels.size())), paths = std::max(1.0, sums[7]);
    ctx.stats["trace_label_max_ms"] = worst * tick_ms;
    ctx.stats["trace_label_mean_ms"] = sums[0] * tick_ms / nl;
    const char *phase[5] = {"total", "target", "railroad", "backtrace", "invalidate"};
    for (int j = 1; j < 5; ++j) ctx.stats[std::string("trace_") + phase[j] + "_ms_per_label"] = sums[j] * tick_ms / nl;
    ctx.stats["railroad_rounds_per_path"] = sums[5] / paths;
    ctx.stats["invalidate_rounds_per_path"] = sums[6] / paths;
    for (int j = 1; j < 5; ++j)
      ctx.stats[std::string("slowest_") + phase[j] + "_ms"] = double(ls[P * worst_id + j]) * tick_ms;
    ctx.stats["slowest_paths"] = double(ls[P * worst_id + 7]);
    ctx.stats["slowest_voxels"] = double(boxes[worst_id].count);
    ctx.stats["slowest_railroad_rounds"] = double(ls[P * worst_id + 5]);
    ctx.stats["slowest_invalidate_rounds"] = double(ls[P * worst_id + 6]);
    ctx.stats["slowest_id"] = double(worst_id);
    if (ctx.kernel_timing)  // per-label detail for diagnostics
      for (int id : labels) {
        const std::string s = std::to_string(id);
        ctx.stats["label_ms:" + s] = double(ls[P * id]) * tick_ms;
        ctx.stats["label_railroad_ms:" + s] = double(ls[P * id + 2]) * tick_ms;
        ctx.stats["label_paths:" + s] = double(ls[P * id + 7]);
        ctx.stats["label_near:" + s] = double(ls[P * id + 10]);
        ctx.stats["label_inval:" + s] = double(ls[P * id + 11]);
        ctx.stats["label_voxels:" + s] = double(boxes[id].count);
        ctx.stats["label_dbfmax:" + s] = double(dbf_max[id]);
      }
  }
  ctx.stats["labels_traced"] = double(labels.size());

  // ---- assembly: groups = input labels in first-skeleton-component order -----------------------
  const int records = cnt[1];
  std::vector<int> rl = rec_label.download(ctx, records), rn = rec_len.download(ctx, records);
  std::vector<uint8_t> has_edges(k, 0);
  for (int r = 0; r < records; ++r)
    if (rn[r] >= 2) has_edges[rl[r]] = 1;
  AssemblyInput ai;
  ai.cc = cc.labels.data();
  ai.d = d;
  ai.dbf = dbf.data();
  ai.store = store.data();
  ai.store_used = cnt[0];
  ai.rec_start = rec_start.data();
  ai.rec_len = rec_len.data();
  ai.records = records;
  ai.group_of.assign(k, 0);
  std::map<int64_t, int> group_index;
  for (size_t id = 1; id < k; ++id) {
    if (!has_edges[id]) continue;
    int64_t label = cc.mapping[id];
    auto it = group_index.find(label);
    if (it == group_index.end()) {
      it = group_index.emplace(label, int(ai.group_labels.size())).first;
      ai.group_labels.push_back(label);
    }
    ai.group_of[id] = it->second;
  }
  std::copy(o.aniso, o.aniso + 3, ai.aniso);
  Packed out = assemble(ctx, cc.cc_bytes, ai);
  lap("assembly_ms");
  return out;
}

}  // namespace brook
