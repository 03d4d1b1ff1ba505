#include "engine.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <map>
#include <memory>
#include <numeric>

#include "avocado.hpp"
#include "fields.hpp"
#include "graph.hpp"
#include "soma.hpp"
#include "trace.hpp"

namespace brook {
namespace {

using Clock = std::chrono::steady_clock;
constexpr float kInfHost = std::numeric_limits<float>::infinity();

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

void check_options(const SkeletonizeOptions &o) {
  for (float a : o.aniso)
    if (!(a > 0) || !std::isfinite(a)) throw std::invalid_argument("anisotropy must be finite and positive");
  const double exponent = o.teasar.pdrf_exponent;
  if (std::isfinite(exponent) && exponent == std::floor(exponent) && std::fmod(exponent, 2.0) == 1.0)
    throw std::invalid_argument("odd pdrf_exponent makes the background penalty negative");
  if (!(exponent > 0 && exponent < 65536 && exponent == std::floor(exponent) && !(int(exponent) & (int(exponent) - 1))))
    throw Unsupported("pdrf_exponent must be a power of two (1..32768) in this build");
}

// Geometric step weights of the 26 neighbours (float of the double length, batched.cu:186-188).
std::vector<float> step_weights(const float aniso[3]) {
  const double ax = aniso[0], ay = aniso[1], az = aniso[2];
  std::vector<float> W = {float(ax), float(ax), float(ay), float(ay), float(az), float(az)};
  for (double e : {std::sqrt(ax * ax + ay * ay), std::sqrt(ay * ay + az * az), std::sqrt(ax * ax + az * az)})
    for (int i = 0; i < 4; ++i) W.push_back(float(e));
  for (int i = 0; i < 8; ++i) W.push_back(float(std::sqrt(ax * ax + ay * ay + az * az)));
  return W;
}

// One volume through the pipeline, stage by stage, so that several volumes can share the flood
// and trace launches: preamble (CCL, analysis, EDT, border targets, voxel records), floods,
// PDRF and target order, trace, assembly. Device buffers live as long as the object.
// Auxiliary contexts (each serves one thread at a time): kBorderAux the border jobs of groups under
// kWorkerMin samples, kRootAux the early root searches (labels of at least kEarlyRootVoxels voxels
// in volumes of at least kEarlyRootVolume voxels); a group of at least kWorkerMin samples runs
// its per-sample stages on kWorkers host threads, each with auxiliary context kWorkerBase + w of
// its own.
constexpr int kBorderAux = 0, kRootAux = 1, kWorkerBase = 2;
constexpr int kWorkers = 16, kWorkerMin = 8;
constexpr int64_t kEarlyRootVoxels = int64_t(1) << 17, kEarlyRootVolume = int64_t(1) << 26;

class Sample {
 public:
  // A soma whose voids hold part of another traced component is traced in a private arena
  // (soma.cu:54-72, 141-237): its filled box as a volume of its own, whose one component is the
  // filled soma, with the label's border targets given (the last one the root) and its vertices
  // placed at the box's corner. Nothing of the main volume is painted or evicted for it.
  struct Arena {
    int64_t id = 0;               // the soma's component id in the main volume
    int origin[3] = {0, 0, 0};    // the box's corner in the input volume
    Buffer<uint8_t> mask;         // 1 on the filled soma, Fortran order over the box
    Buffer<unsigned> graph;       // the voxel graph's words of the box (empty: no graph)
    LabelVolume volume;           // the mask (and graph) as the arena's input
    SkeletonizeOptions o;         // the main options without border targets, hole filling or avocados
    std::vector<Point> borders;   // the label's border targets in box coordinates
    std::unique_ptr<Sample> sample;
  };

  Sample(Context &ctx, const LabelVolume &in, const SkeletonizeOptions &o, FieldsOut *fields, const Arena *arena = nullptr)
      : cx_(&ctx), main_(&ctx), in_(in), o_(o), fields_(fields), d_(in.dims), n_(in.size()), arena_(arena) {
    if (arena) std::copy(arena->origin, arena->origin + 3, origin_);
  }
  // The preamble's jobs read this sample's components and write its buffers: they end before any
  // member goes (when an error skipped the second half that joins them).
  ~Sample() {
    if (border_job_.valid()) border_job_.wait();
    if (root_job_.valid()) root_job_.wait();
  }
  Sample(const Sample &) = delete;
  Sample &operator=(const Sample &) = delete;

  bool done() const { return done_; }
  Packed take_result() { return std::move(result_); }
  // The context the sample's next work runs on (its buffers stay in the pool they came from);
  // use_context_all also moves its private arenas' samples.
  void use_context(Context &c) { cx_ = &c; }
  void use_context_all(Context &c) {
    cx_ = &c;
    for (auto &a : arenas_) a->sample->use_context_all(c);
  }
  // This sample and its arenas' samples, which share its flood and trace launches.
  void collect(std::vector<Sample *> &all) {
    all.push_back(this);
    for (auto &a : arenas_) a->sample->collect(all);
  }
  // Rough device bytes a sample holds from its preamble to its assembly (40 B per voxel, 56 B per
  // foreground voxel at the usual 40% occupancy, plus slack).
  static int64_t footprint(int64_t n) { return n * 40 + (n * 2 / 5) * 56 + (int64_t(64) << 20); }

  void preamble(Lap &lap) {
    preamble_components(lap, 0, false);
    preamble_labels(lap);
  }

  // The preamble in two halves, so that a group runs every sample's first half before any second
  // half: components, analysis, filled holes and the EDT (the border targets found meanwhile on
  // auxiliary context `aux`, or at the end of the half on this context when aux < 0, and with
  // early_roots the root searches, see below), then everything from the border targets on.
  void preamble_components(Lap &lap, int aux, bool early_roots) {
    if (n_ == 0 || double(n_) <= o_.dust) return finish_empty();
    bool uniform = false;
    if (in_.graph && !arena_) {  // components of the graph's steps (skeletonize.cu:76-78); an arena is one component
      GraphComponents g = graph_components(*cx_, in_, in_.graph, in_.graph_compact_2d);
      cc_ = std::move(g.cc);
      keys_ = std::move(g.keys);
      uniform = g.input_uniform;
    } else {
      cc_ = connected_components(*cx_, in_);
    }
    lap("prep_ccl_ms");
    count_ = cc_.count;
    if (count_ == 0) return finish_empty();
    k_ = size_t(count_) + 1;
    boxes_ = analyze(*cx_, cc_, d_);
    lap("prep_analyze_ms");
    // of the volume before any fill; with a graph, every input voxel has one label (skeletonize.cu:80)
    const bool black_border = in_.graph && !arena_ ? uniform : count_ == 1 && boxes_[1].count == n_;
    if (o_.fill_holes) {
      fill_holes();
      lap("prep_fill_holes_ms");
    }
    // Border targets read only the components, so unless fix_avocados relabels them first they
    // are found on an auxiliary context's stream, from a host thread of their own, while the EDT
    // (and, in a group, the other samples' first halves) run on this one. The jobs of one
    // auxiliary context take turns (its stream and buffer pool serve one job at a time).
    early_borders_ = o_.fix_borders && !o_.fix_avocados && aux >= 0;
    early_roots = early_roots && !in_.graph && !arena_ && !o_.fix_avocados && n_ >= kEarlyRootVolume;
    if (early_borders_ || early_roots) {  // the components are final: the jobs below start after them
      BROOK_HIP(hipEventCreateWithFlags(&cc_ready_.e, hipEventDisableTiming));
      BROOK_HIP(hipEventRecord(cc_ready_.e, cx_->stream()));
    }
    if (early_borders_) {
      Context &ax = main_->aux(aux);
      border_job_ = std::async(std::launch::async, [this, &ax, ready = cc_ready_.e] {
        std::lock_guard<std::mutex> turn(ax.mutex());
        ax.activate();
        BROOK_HIP(hipStreamWaitEvent(ax.stream(), ready, 0));
        auto targets = border_targets(ax, cc_.labels.data(), cc_.cc_bytes, d_, count_, o_.aniso);
        border_ms_[0] = ax.stats["borders_device_ms"];
        border_ms_[1] = ax.stats["borders_host_ms"];
        ax.stats.erase("borders_device_ms");
        ax.stats.erase("borders_host_ms");
        return targets;
      });
    }
    if (early_roots) search_roots_early();
    dbf_ = Buffer<float>(*cx_, size_t(n_));
    Buffer<unsigned> label_max(*cx_, k_);
    label_max.clear(*cx_);
    if (in_.graph)
      graph_edt(*cx_, cc_.labels.data(), cc_.cc_bytes, d_, in_.graph, o_.aniso, black_border, dbf_.data(), label_max.data());
    else
      edt(*cx_, cc_.labels.data(), cc_.cc_bytes, d_, o_.aniso, black_border, 3, dbf_.data(), label_max.data());
    std::vector<unsigned> max_bits = label_max.download(*cx_);
    dbf_max_.resize(k_);
    for (size_t i = 0; i < k_; ++i) std::memcpy(&dbf_max_[i], &max_bits[i], 4);
    lap("prep_edt_ms");
    if (o_.fix_avocados) {
      fix_avocados(black_border);
      lap("prep_avocados_ms");
      if (count_ == 0) return finish_empty();
    }
    if (aux < 0 && o_.fix_borders) {  // in place (a worker's context)
      borders_ = border_targets(*cx_, cc_.labels.data(), cc_.cc_bytes, d_, count_, o_.aniso);
      borders_ready_ = true;
      lap("prep_borders_ms");
    }
  }

  // A label with neither a border target nor a soma root takes a root search flood before its DAF
  // flood, and the flood launch waits for the largest of those two-flood labels. The search needs
  // only the components, so the labels that can have no border target (box off every face of the
  // volume, or no fix_borders) and are large enough to end the flood launch that way (at least half
  // the largest label, and at least kEarlyRootVoxels voxels: below that the flood is short and the
  // auxiliary buffers cost more) are searched here, one workgroup each, on auxiliary context kRootAux with
  // records, voxel lists and scratch of its own, from a host thread of their own, while the EDT and
  // the border targets run; the second half gives the roots found to the labels that end up
  // without one (the same flood, so the same root), and the flood launch floods them once.
  void search_roots_early() {
    std::vector<LabelItem> items;
    int64_t largest = 0;
    auto traced = [&](const Box &b) {
      const int64_t ext = int64_t(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
      return double(b.count) > o_.dust && ext > 1;
    };
    for (size_t id = 1; id < k_; ++id)
      if (traced(boxes_[id])) largest = std::max(largest, boxes_[id].count);
    for (size_t id = 1; id < k_; ++id) {
      const Box &b = boxes_[id];
      if (!traced(b) || 2 * b.count < largest || b.count < kEarlyRootVoxels) continue;
      const int dims[3] = {d_.sx, d_.sy, d_.sz};
      bool face = false;
      for (int a = 0; a < 3; ++a) face |= b.lo[a] == 0 || b.hi[a] == dims[a] - 1;
      if (o_.fix_borders && face) continue;
      items.push_back(LabelItem{0, int(id), b.count});
    }
    if (items.empty()) return;
    std::stable_sort(items.begin(), items.end(), [](const LabelItem &a, const LabelItem &b) { return a.size > b.size; });
    Context &ax = main_->aux(kRootAux);
    std::vector<int64_t> counts(k_, 0);
    for (size_t id = 1; id < k_; ++id) counts[id] = boxes_[id].count;
    std::vector<long long> first(k_, -1);
    for (size_t id = 1; id < k_; ++id) first[id] = cc_.roots[id - 1];
    root_job_ = std::async(std::launch::async, [this, &ax, ready = cc_ready_.e, items = std::move(items),
                                                counts = std::move(counts), first = std::move(first)]() -> std::vector<long long> {
      try {  // an early search that fails (memory, a stalled flood) leaves the search to the flood launch
        std::lock_guard<std::mutex> turn(ax.mutex());
        ax.activate();
        BROOK_HIP(hipStreamWaitEvent(ax.stream(), ready, 0));
        const size_t k = counts.size();
        std::vector<int> seg_start(k + 1, 0), seg_count(k, 0);
        for (size_t id = 1; id < k; ++id) seg_count[id] = int(counts[id]);
        for (size_t id = 0; id < k; ++id) seg_start[id + 1] = seg_start[id] + seg_count[id];
        const size_t F = size_t(std::max(seg_start[k], 1));
        Buffer<int> seg_start_dev, vox, seg_count_dev(ax, k);
        build_segments(ax, cc_.labels.data(), cc_.cc_bytes, d_, seg_start, seg_start_dev, vox);
        seg_count_dev.upload(ax, seg_count.data(), k);
        std::vector<long long> none(k, -1);
        Buffer<long long> first_dev(ax, k), root_dev(ax, k), target_dev(ax, k);
        Buffer<float> max_daf_dev(ax, k), weights(ax, 26);
        first_dev.upload(ax, first.data(), k);
        root_dev.upload(ax, none.data(), k);
        const std::vector<float> W = step_weights(o_.aniso);
        weights.upload(ax, W.data(), 26);
        Buffer<VoxelRec> rec(ax, size_t(n_));
        voxel_records(ax, cc_.labels.data(), cc_.cc_bytes, d_, rec.data());
        Buffer<int> la(ax, F), lb(ax, F), fa(ax, 2 * F), fb(ax, 2 * F);
        FloodJob j;
        j.d = d_;
        j.t.seg_start = seg_start_dev.data();
        j.t.seg_count = seg_count_dev.data();
        j.t.vox = vox.data();
        j.t.first = first_dev.data();
        j.t.root = root_dev.data();
        j.t.target = target_dev.data();
        j.t.max_daf = max_daf_dev.data();
        j.t.weights = weights.data();
        j.t.rec = rec.data();
        j.lists_a = la.data();
        j.lists_b = lb.data();
        j.far_a = fa.data();
        j.far_b = fb.data();
        j.root_only = 1;
        // one workgroup per label, at most a quarter of the chip: the EDT runs beside it
        prepare_fields(ax, {j}, items, std::max(1, ax.compute_units() / 4));
        std::vector<long long> roots = root_dev.download(ax);  // (synchronizes the auxiliary stream)
        early_rec_ = std::move(rec);
        early_seg_start_ = std::move(seg_start_dev);
        early_vox_ = std::move(vox);
        return roots;
      } catch (...) {
        return {};
      }
    });
  }

  void preamble_labels(Lap &lap) {
    if (done_) return;
    if (!borders_ready_) borders_.assign(k_, {});
    if (early_borders_) {
      borders_ = border_job_.get();
      cx_->stats["borders_device_ms"] += border_ms_[0];
      cx_->stats["borders_host_ms"] += border_ms_[1];
    } else if (o_.fix_borders && !borders_ready_) {
      borders_ = border_targets(*cx_, cc_.labels.data(), cc_.cc_bytes, d_, count_, o_.aniso);
    }
    if (arena_ && count_ == 1) borders_[1] = arena_->borders;
    lap("prep_borders_ms");
    // Each point is filed under the component its wrapped coordinates fall on and traced at the
    // coordinates as written (skeletonize.cu:59-65). A point whose written coordinates flatten off
    // the volume or onto another component is rejected: the trace starts at that flat index.
    auto file_points = [&](const std::vector<Point> &points, std::vector<std::vector<Point>> &out) {
      out.assign(k_, {});
      for (const Point &p : points) {
        const int64_t id = component_at(*cx_, cc_, d_, p);
        if (id) {
          const int64_t flat = flat_of(p, d_), plane = int64_t(d_.sx) * d_.sy;
          const Point at = {flat % d_.sx, (flat / d_.sx) % d_.sy, flat / plane};
          if (flat < 0 || flat >= n_ || component_at(*cx_, cc_, d_, at) != id)
            throw std::invalid_argument("extra target's coordinates do not flatten onto its component");
        }
        out.at(size_t(id)).push_back(p);
      }
    };
    file_points(o_.before, before_);
    file_points(o_.after, after_);

    // labels to trace (skeletonize.cu:128-133, batched.cu:186-203)
    active_.assign(k_, 0);
    for (size_t id = 1; id < k_; ++id) {
      const Box &b = boxes_[id];
      int64_t ext = int64_t(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
      if (double(b.count) <= o_.dust || ext <= 1) continue;
      active_[id] = 1;
      labels_.push_back(int(id));
    }
    // Somata (batched.cu:194-205, soma.cu:141-237), in ascending id: a label above the detection
    // threshold whose box encloses voids takes them, evicting every traced label wholly inside
    // them, and its DBF becomes the EDT of its filled box; when a traced label lies partly inside
    // them, the soma goes to a private arena instead and leaves the main volume as it is. A label
    // is a soma when its (refreshed) DBF maximum exceeds the acceptance threshold.
    soma_.assign(k_, 0);
    bool any_soma = false, any_fill = false;
    std::vector<uint8_t> evicted(k_, 0);
    for (int id : labels_) {
      if (evicted[id] || !(double(dbf_max_[id]) > o_.teasar.soma_detection)) continue;
      if (enclosed_voids(*cx_, cc_.labels.data(), cc_.cc_bytes, d_, id, boxes_[id])) {
        VoidFill fill(*cx_, cc_.labels.data(), cc_.cc_bytes, d_, id, boxes_[id], k_);
        const std::vector<int64_t> &taken = fill.taken();
        auto owner = [&](size_t o) { return taken[o] && active_[o] && !evicted[o]; };
        bool partial = false;
        for (size_t o = 1; o < k_; ++o) partial |= owner(o) && taken[o] != boxes_[o].count;
        if (partial) {
          arenas_.push_back(make_arena(fill, id));
          evicted[id] = 1;
          active_[id] = 0;
          any_fill = true;
          continue;
        }
        for (size_t o = 1; o < k_; ++o) {
          if (!owner(o)) continue;
          evicted[o] = 1;
          active_[o] = 0;
        }
        fill.paint(*cx_);
        for (size_t o = 1; o < k_; ++o) boxes_[o].count -= taken[o];
        boxes_[id].count += fill.voids();
        dbf_max_[id] = fill.dbf_refresh(*cx_, o_.aniso, dbf_.data(), boxes_[id].count, in_.graph);
        any_fill = true;
      }
      soma_[id] = double(dbf_max_[id]) > o_.teasar.soma_acceptance;
      any_soma |= soma_[id] != 0;
    }
    if (any_fill) {
      std::vector<int> kept;
      for (int id : labels_)
        if (!evicted[id]) kept.push_back(id);
      labels_.swap(kept);
    }
    // The early root searches built the records and voxel lists of the components as they were
    // before any soma painting: they serve the floods too unless a soma's voids were painted.
    std::vector<long long> early_roots;
    if (root_job_.valid()) early_roots = root_job_.get();
    const bool reuse = !any_fill && early_roots.size() == k_ && early_rec_.size() == size_t(n_);
    std::vector<int> seg_start(k_ + 1, 0), seg_count(k_, 0);
    for (size_t id = 1; id < k_; ++id) seg_count[id] = int(boxes_[id].count);
    for (size_t id = 0; id < k_; ++id) seg_start[id + 1] = seg_start[id] + seg_count[id];
    F_ = seg_start[k_];
    if (reuse) {
      seg_start_dev_ = std::move(early_seg_start_);
      vox_ = std::move(early_vox_);
    } else {
      build_segments(*cx_, cc_.labels.data(), cc_.cc_bytes, d_, seg_start, seg_start_dev_, vox_);
    }
    seg_count_dev_ = Buffer<int>(*cx_, k_);
    seg_count_dev_.upload(*cx_, seg_count.data(), k_);

    // roots, DAF, targets: inputs of the floods
    std::vector<long long> first(k_, -1);
    root_.assign(k_, -1);
    for (size_t id = 1; id < k_; ++id) first[id] = cc_.roots[id - 1];
    for (int id : labels_)
      if (!soma_[id] && !borders_[id].empty()) root_[id] = flat_of(borders_[id].back(), d_);
    if (any_soma) {
      // the soma root (batched.cu:58-108): among the label's DBF maxima, in box-relative
      // coordinates sorted by (x, y, z), the first of least squared distance to their centroid
      Buffer<uint8_t> soma_dev(*cx_, k_);
      soma_dev.upload(*cx_, soma_.data(), k_);
      Buffer<float> dbf_max_dev(*cx_, k_);
      dbf_max_dev.upload(*cx_, dbf_max_.data(), k_);
      std::vector<Maximum> maxima = soma_maxima(*cx_, cc_.labels.data(), cc_.cc_bytes, n_, dbf_.data(), dbf_max_dev.data(),
                                                soma_dev.data());
      std::map<int64_t, std::vector<std::pair<Point, int64_t>>> by_label;
      for (const Maximum &m : maxima) {
        const Box &b = boxes_[size_t(m.id)];
        const int64_t x = m.voxel % d_.sx, y = (m.voxel / d_.sx) % d_.sy, z = m.voxel / (int64_t(d_.sx) * d_.sy);
        by_label[m.id].push_back({Point{x - b.lo[0], y - b.lo[1], z - b.lo[2]}, m.voxel});
      }
      for (auto &entry : by_label) {
        auto &pts = entry.second;
        std::sort(pts.begin(), pts.end());
        int64_t sum[3] = {0, 0, 0};
        for (auto &p : pts) for (int a = 0; a < 3; ++a) sum[a] += p.first[a];
        float center[3];
        for (int a = 0; a < 3; ++a) center[a] = float(double(sum[a]) / double(pts.size()));
        double best = INFINITY;
        for (auto &p : pts) {
          double cost = 0;
          for (int a = 0; a < 3; ++a) { const double dd = double(p.first[a]) - center[a]; cost += dd * dd; }
          if (cost < best) { best = cost; root_[size_t(entry.first)] = p.second; }
        }
      }
      std::vector<float> radius(k_, 0.0f);
      for (int id : labels_)
        if (soma_[id]) radius[id] = dbf_max_[id];
      free_radius_dev_ = Buffer<float>(*cx_, k_);
      free_radius_dev_.upload(*cx_, radius.data(), k_);
    }
    if (early_roots.size() == k_)  // roots searched beside the EDT (search_roots_early)
      for (int id : labels_)
        if (root_[id] < 0 && early_roots[size_t(id)] >= 0) root_[id] = early_roots[size_t(id)];
    flood_voxels_ = 0;
    for (int id : labels_) flood_voxels_ += boxes_[id].count * (root_[id] < 0 ? 2 : 1);
    first_dev_ = Buffer<long long>(*cx_, k_);
    root_dev_ = Buffer<long long>(*cx_, k_);
    target_dev_ = Buffer<long long>(*cx_, k_);
    max_daf_dev_ = Buffer<float>(*cx_, k_);
    first_dev_.upload(*cx_, first.data(), k_);
    root_dev_.upload(*cx_, root_.data(), k_);
    std::vector<float> W = step_weights(o_.aniso);
    weights_ = Buffer<float>(*cx_, 26);
    weights_.upload(*cx_, W.data(), 26);
    daf_ = Buffer<float>(*cx_, size_t(n_));
    fill_value(*cx_, daf_.data(), n_, kInf);
    if (reuse) {
      rec_ = std::move(early_rec_);  // the root searches left every word ~0
    } else {
      rec_ = Buffer<VoxelRec>(*cx_, size_t(n_));  // 16 bytes per voxel: PDRF, neighbour mask + flags, distance word
      voxel_records(*cx_, cc_.labels.data(), cc_.cc_bytes, d_, rec_.data());
    }
    early_rec_ = Buffer<VoxelRec>();
    early_seg_start_ = Buffer<int>();
    early_vox_ = Buffer<int>();
    if (in_.graph) {  // push masks in the records, pull masks beside them
      inmask_ = Buffer<unsigned>(*cx_, size_t(n_));
      graph_gates(*cx_, rec_.data(), in_.graph, d_, inmask_.data());
    }
    // lists 2 and 3 are the railroad's far lists: two entries per voxel (trace.hip far_compact)
    for (int i = 0; i < 6; ++i) lists_.emplace_back(*cx_, size_t(std::max<int64_t>(F_, 1)) * (i == 2 || i == 3 ? 2 : 1));
    if (!o_.fix_branching) parent_ = Buffer<int>(*cx_, size_t(n_));
    for (auto &a : arenas_) a->sample->preamble(lap);
  }

  std::unique_ptr<Arena> make_arena(VoidFill &fill, int id) {
    auto a = std::make_unique<Arena>();
    a->id = id;
    const int *lo = fill.lo(), *ext = fill.ext();
    for (int ax = 0; ax < 3; ++ax) a->origin[ax] = origin_[ax] + lo[ax];
    a->mask = fill.take_mask();
    if (in_.graph) a->graph = fill.crop_graph(*cx_, in_.graph);
    a->volume.data = a->mask.data();
    a->volume.bytes = 1;
    a->volume.dims = Dims{ext[0], ext[1], ext[2]};
    a->volume.graph = in_.graph ? a->graph.data() : nullptr;
    a->o = o_;
    a->o.fix_borders = a->o.fill_holes = a->o.fix_avocados = false;
    auto shift = [&](const std::vector<Point> &points) {
      std::vector<Point> out = points;
      for (Point &p : out) for (int ax = 0; ax < 3; ++ax) p[ax] -= lo[ax];
      return out;
    };
    a->o.before = shift(before_[id]);
    a->o.after = shift(after_[id]);
    a->borders = shift(borders_[id]);
    a->sample = std::make_unique<Sample>(*cx_, a->volume, a->o, nullptr, a.get());
    return a;
  }

  // fix_avocados (skeletonize.cu:91-95): the correction runs on host copies of the ids and the DBF
  // (as the reference's does); the DBF is refreshed on the device after every pass that changed
  // labels; then the ids are renumbered by first voxel and the boxes recomputed.
  void fix_avocados(bool black_border) {
    const size_t n = size_t(n_);
    std::vector<uint32_t> labels(n);
    {
      const std::vector<uint8_t> raw = cc_.labels.download(*cx_, n * size_t(cc_.cc_bytes));
      for (size_t i = 0; i < n; ++i) {
        uint32_t v = 0;
        std::memcpy(&v, raw.data() + i * cc_.cc_bytes, size_t(cc_.cc_bytes));
        labels[i] = v;
      }
    }
    const std::vector<uint32_t> original = labels;
    std::vector<float> distance = dbf_.download(*cx_, n);
    Buffer<uint32_t> scratch;
    auto refresh = [&](const std::vector<uint32_t> &current, std::vector<float> &field) {
      if (!scratch.size()) scratch = Buffer<uint32_t>(*cx_, n);
      scratch.upload(*cx_, current.data(), n);
      if (in_.graph) graph_edt(*cx_, scratch.data(), 4, d_, in_.graph, o_.aniso, black_border, dbf_.data());
      else edt(*cx_, scratch.data(), 4, d_, o_.aniso, black_border, 3, dbf_.data());
      field = dbf_.download(*cx_, n);
    };
    cx_->stats["avocado_passes"] = avocado_correct(labels, distance, d_, k_, o_.teasar.soma_detection, refresh, keys_);
    keys_.clear();  // the renumbered ids are their own keys
    std::vector<int64_t> mapping, roots;
    const size_t count = avocado_renumber(labels, original, cc_.mapping, k_, mapping, roots);
    const int bytes = id_bytes(int64_t(count));
    std::vector<uint8_t> raw(n * size_t(bytes));
    for (size_t i = 0; i < n; ++i) std::memcpy(raw.data() + i * bytes, &labels[i], size_t(bytes));  // little endian
    cc_.labels = Buffer<uint8_t>(*cx_, std::max<size_t>(raw.size(), 1));
    cc_.labels.upload(*cx_, raw.data(), raw.size());
    cc_.cc_bytes = bytes;
    cc_.count = int64_t(count);
    cc_.mapping = mapping;
    cc_.roots = roots;
    count_ = int64_t(count);
    k_ = count + 1;
    dbf_max_.assign(k_, 0.0f);
    for (size_t i = 0; i < n; ++i)
      if (labels[i] && distance[i] < kInfHost) dbf_max_[labels[i]] = std::max(dbf_max_[labels[i]], distance[i]);
    if (count) boxes_ = analyze(*cx_, cc_, d_);
  }

  // fill_all_holes (holes.cu:37-62, skeletonize.cu:83-90): in ascending id and with the original
  // boxes, a component's box voids become the component; a component painted over, even in part,
  // fills nothing later; then the boxes and first voxels are recomputed.
  void fill_holes() {
    std::vector<uint8_t> skip(k_, 0);
    bool painted = false;
    for (size_t id = 1; id < k_; ++id) {
      const Box &b = boxes_[id];
      if (!b.count || skip[id]) continue;
      const int64_t ex = b.hi[0] - b.lo[0] + 1, ey = b.hi[1] - b.lo[1] + 1, ez = b.hi[2] - b.lo[2] + 1;
      if (ex <= 2 || ey <= 2 || ez <= 2 || ex * ey * ez == b.count) continue;  // no voxel can be enclosed
      if (!enclosed_voids(*cx_, cc_.labels.data(), cc_.cc_bytes, d_, int64_t(id), b)) continue;
      VoidFill fill(*cx_, cc_.labels.data(), cc_.cc_bytes, d_, int64_t(id), b, k_);
      fill.paint(*cx_);
      for (size_t o = 1; o < k_; ++o)
        if (fill.taken()[o] && o != id) skip[o] = 1;
      cx_->stats["hole_filled_voxels"] += double(fill.voids());
      painted = true;
    }
    if (!painted) return;
    boxes_ = analyze(*cx_, cc_, d_);
    const std::vector<int64_t> first = first_voxels(*cx_, cc_.labels.data(), cc_.cc_bytes, n_, count_);
    for (size_t id = 1; id < k_; ++id) cc_.roots[id - 1] = first[id];
  }

  FloodJob flood_job() const {
    FloodJob j;
    j.d = d_;
    j.t.seg_start = seg_start_dev_.data();
    j.t.seg_count = seg_count_dev_.data();
    j.t.vox = vox_.data();
    j.t.first = first_dev_.data();
    j.t.root = root_dev_.data();
    j.t.target = target_dev_.data();
    j.t.max_daf = max_daf_dev_.data();
    j.t.weights = weights_.data();
    j.t.free_radius = free_radius_dev_.size() ? free_radius_dev_.data() : nullptr;
    j.t.rec = rec_.data();
    j.daf = daf_.data();
    j.lists_a = lists_[0].data();
    j.lists_b = lists_[1].data();
    j.far_a = lists_[2].data();
    j.far_b = lists_[3].data();
    return j;
  }

  // The sample's labels as work items (sorted largest first over all samples of a group).
  void items(int s, std::vector<LabelItem> &q) const {
    if (done_) return;
    for (int id : labels_) q.push_back(LabelItem{s, id, boxes_[id].count});
  }
  void soma_items(int s, std::vector<LabelItem> &q) const {
    if (done_ || soma_.empty()) return;
    for (int id : labels_)
      if (soma_[id]) q.push_back(LabelItem{s, id, boxes_[id].count});
  }

  void after_floods(Lap &lap) {
    if (done_) return;
    root_ = root_dev_.download(*cx_);
    target_ = target_dev_.download(*cx_);
    max_daf_ = max_daf_dev_.download(*cx_);

    // PDRF (exponent a power of two: squarings), alive bits, rails, target order
    std::vector<float> M(k_, 0.0f), inv(k_, 0.0f);
    std::vector<uint8_t> use(k_, 0);
    for (int id : labels_) {
      M[id] = float(1.0 / std::pow(double(dbf_max_[id]), 1.01));
      if (max_daf_[id] != 0) { inv[id] = float(1.0 / double(max_daf_[id])); use[id] = 2; }
      else use[id] = 1;
    }
    int squarings = 0;
    for (int e = int(o_.teasar.pdrf_exponent); e > 1; e >>= 1) ++squarings;
    Buffer<float> M_dev(*cx_, k_), inv_dev(*cx_, k_);
    Buffer<uint8_t> use_dev(*cx_, k_);
    M_dev.upload(*cx_, M.data(), k_);
    inv_dev.upload(*cx_, inv.data(), k_);
    use_dev.upload(*cx_, use.data(), k_);
    Buffer<float> pdrf;  // the PDRF as a volume of its own only when the fields are asked for
    if (fields_ && fields_->pdrf) pdrf = Buffer<float>(*cx_, size_t(n_));
    pdrf_and_alive(*cx_, cc_.labels.data(), cc_.cc_bytes, n_, dbf_.data(), daf_.data(), M_dev.data(), inv_dev.data(),
                   use_dev.data(), squarings, float(o_.teasar.pdrf_scale), pdrf.size() ? pdrf.data() : nullptr, rec_.data());
    Buffer<int> labels_dev(*cx_, std::max<size_t>(labels_.size(), 1));
    labels_dev.upload(*cx_, labels_.data(), labels_.size());
    if (fields_) {
      fields_->count = count_;
      fields_->root.assign(root_.begin(), root_.end());
      fields_->target.assign(target_.begin(), target_.end());
      fields_->dbf_max = dbf_max_;
      fields_->max_daf = max_daf_;
      fields_->active = active_;
      cx_->synchronize();
      if (fields_->dbf) BROOK_HIP(hipMemcpy(fields_->dbf, dbf_.data(), size_t(n_) * 4, hipMemcpyDeviceToHost));
      if (fields_->daf) BROOK_HIP(hipMemcpy(fields_->daf, daf_.data(), size_t(n_) * 4, hipMemcpyDeviceToHost));
      if (fields_->pdrf) BROOK_HIP(hipMemcpy(fields_->pdrf, pdrf.data(), size_t(n_) * 4, hipMemcpyDeviceToHost));
      if (fields_->stop) return finish_empty();
    }
    // the root becomes a rail (lockstep.cu:183); the fix_branching=False field floods from it
    if (o_.fix_branching) root_rails(*cx_, rec_.data(), root_dev_.data(), labels_dev.data(), int(labels_.size()));
    const size_t fl = size_t(std::max<int64_t>(F_, 1));
    order_ = Buffer<int>(*cx_, fl);
    snap_a_ = Buffer<unsigned long long>(*cx_, fl);
    target_order(*cx_, cc_.labels.data(), cc_.cc_bytes, n_, daf_.data(), F_, count_, order_.data());
    lap("prep_pdrf_order_ms");

    // manual targets and path budgets (lockstep.cu:116-143)
    std::vector<int> mstart(k_, 0), mcount(k_, 0), acount(k_, 0), mflat, maxp(k_, 0);
    for (int id : labels_) {
      std::vector<Point> manual = borders_[id];
      if (!manual.empty()) manual.pop_back();
      if (soma_[id] && !borders_[id].empty()) manual.insert(manual.begin(), borders_[id].back());
      manual.insert(manual.end(), before_[id].begin(), before_[id].end());
      mstart[id] = int(mflat.size());
      if (manual.empty() && !soma_[id]) mflat.push_back(int(target_[id]));
      else for (auto it = manual.rbegin(); it != manual.rend(); ++it) mflat.push_back(int(flat_of(*it, d_)));
      mcount[id] = int(mflat.size()) - mstart[id];
      // the targets after the DAF order (per-component tracer, trace.cu:179-199): last given first
      for (auto it = after_[id].rbegin(); it != after_[id].rend(); ++it) mflat.push_back(int(flat_of(*it, d_)));
      acount[id] = int(after_[id].size());
      int64_t budget = o_.teasar.max_paths >= 0 ? o_.teasar.max_paths : boxes_[id].count;
      maxp[id] = mcount[id] + acount[id] >= budget ? 0 : int(std::min<int64_t>(INT32_MAX, budget));
    }
    mflat.push_back(0);
    mstart_dev_ = Buffer<int>(*cx_, k_);
    mcount_dev_ = Buffer<int>(*cx_, k_);
    maxp_dev_ = Buffer<int>(*cx_, k_);
    mflat_dev_ = Buffer<int>(*cx_, mflat.size());
    mstart_dev_.upload(*cx_, mstart.data(), k_);
    mcount_dev_.upload(*cx_, mcount.data(), k_);
    acount_dev_ = Buffer<int>(*cx_, k_);
    acount_dev_.upload(*cx_, acount.data(), k_);
    maxp_dev_.upload(*cx_, maxp.data(), k_);
    mflat_dev_.upload(*cx_, mflat.data(), mflat.size());

    // trace buffers, sized by what the trace can emit (the reference grows its store mid-trace,
    // lockstep.cu:389-391): a label's paths number at most its ordered targets (alive voxels of
    // the label, each path kills its own) plus its manual ones; a path's vertices are new voxels
    // (not yet rail or emitted, and they become so) plus at most two it ends on (the rail reached
    // and the join's rail), or a soma's root; so records <= F + manual targets and the store
    // <= F + 2 * records. (A soma's parent walks do not stop at emitted vertices, so they may
    // repeat voxels: the store's overflow check stays and reports such a case.)
    snap_b_ = Buffer<unsigned long long>(*cx_, fl);
    const int64_t manual_targets = int64_t(mflat.size());
    const int64_t paths_bound = F_ + manual_targets + int64_t(labels_.size());
    if (paths_bound > INT32_MAX / 2) throw Unsupported("path records exceed the int32 store index");
    rec_cap_ = int(std::max<int64_t>(1, paths_bound));
    store_cap_ = std::max<long long>(1, F_ + 2 * paths_bound);
    store_ = Buffer<int>(*cx_, size_t(store_cap_));
    rec_label_ = Buffer<int>(*cx_, rec_cap_);
    rec_len_ = Buffer<int>(*cx_, rec_cap_);
    rec_start_ = Buffer<int>(*cx_, rec_cap_);
    counters_ = Buffer<int>(*cx_, 16);
    ejected_ = Buffer<int>(*cx_, 64);
    counters_.clear(*cx_);
    label_stats_ = Buffer<unsigned long long>(*cx_, size_t(kTraceProfileSlots) * k_);
    label_stats_.clear(*cx_);
    if (free_radius_dev_.size()) {  // the somata's trimming radii and balls (lockstep.cu:137-141)
      std::vector<SomaLabel> table(k_);
      for (int id : labels_)
        if (soma_[id]) {
          table[id].radius = double(dbf_max_[id]) * o_.teasar.soma_scale + o_.teasar.soma_constant;
          table[id].ball = Ball{float(o_.teasar.soma_scale), float(o_.teasar.soma_constant)};
          table[id].max_paths = o_.teasar.max_paths;
        }
      soma_dev_ = Buffer<SomaLabel>(*cx_, k_);
      soma_dev_.upload(*cx_, table.data(), k_);
    }
  }

  TraceArgs trace_args() const {
    TraceArgs ta;
    ta.d = d_;
    ta.rec = rec_.data();
    ta.dbf = dbf_.data();
    ta.seg_start = seg_start_dev_.data();
    ta.seg_count = seg_count_dev_.data();
    ta.order = order_.data();
    ta.mstart = mstart_dev_.data();
    ta.mcount = mcount_dev_.data();
    ta.mflat = mflat_dev_.data();
    ta.acount = acount_dev_.data();
    ta.maxp = maxp_dev_.data();
    for (int i = 0; i < 6; ++i) ta.lists[i] = lists_[i].data();
    ta.snaps[0] = snap_a_.data();
    ta.snaps[1] = snap_b_.data();
    ta.store = store_.data();
    ta.store_cap = store_cap_;
    ta.rec_label = rec_label_.data();
    ta.rec_len = rec_len_.data();
    ta.rec_start = rec_start_.data();
    ta.rec_cap = rec_cap_;
    ta.counters = counters_.data();
    ta.label_stats = label_stats_.data();
    ta.ejected = ejected_.data();
    ta.ejected_cap = 64;
    ta.wx = o_.aniso[0];
    ta.wy = o_.aniso[1];
    ta.wz = o_.aniso[2];
    ta.ball = Ball{float(o_.teasar.scale), float(o_.teasar.constant)};
    ta.delta = std::min({o_.aniso[0], o_.aniso[1], o_.aniso[2]});
    ta.parent = o_.fix_branching ? nullptr : parent_.data();
    ta.root = root_dev_.data();
    if (soma_dev_.size()) ta.soma = soma_dev_.data();
    if (inmask_.size()) ta.inmask = inmask_.data();
    return ta;
  }

  // Trace outcome, statistics (per-label profiles only when `profile`) and assembly.
  void finish(Lap &lap, bool profile) {
    if (done_) return;
    for (auto &a : arenas_) a->sample->finish(lap, false);
    std::vector<int> cnt = counters_.download(*cx_);
    if (cnt[8] & 1) throw std::runtime_error("trace did not converge: a railroad or an invalidation exceeded its round budget");
    if (cnt[8] & 2) throw std::runtime_error("trace far list overflowed its capacity");
    if (cnt[3]) throw std::runtime_error("path store capacity exceeded");
    if (cnt[2]) throw Unsupported("a label has no reachable rail (the per-component fallback tracer is not part of this build)");
    cx_->stats["paths"] += cnt[1];
    cx_->stats["path_vertices"] += cnt[0];
    uint64_t reached = 0, claimed = 0;
    std::memcpy(&reached, &cnt[4], 8);
    std::memcpy(&claimed, &cnt[6], 8);
    cx_->stats["railroad_reached_voxels"] += double(reached);
    cx_->stats["invalidated_voxels"] += double(claimed);
    cx_->stats["labels_traced"] += double(labels_.size());
    if (profile) record_profile();

    // assembly: groups = input labels in first-skeleton-component order
    const int records = cnt[1];
    std::vector<int> rl = rec_label_.download(*cx_, records), rn = rec_len_.download(*cx_, records);
    std::vector<uint8_t> has_edges(k_, 0);
    for (int r = 0; r < records; ++r)
      if (rn[r] >= 2) has_edges[rl[r]] = 1;
    const std::vector<uint8_t> main_edges = has_edges;
    for (auto &a : arenas_)
      if (!a->sample->result_.labels.empty()) has_edges[size_t(a->id)] = 1;
    AssemblyInput ai;
    ai.cc = cc_.labels.data();
    ai.d = d_;
    ai.dbf = dbf_.data();
    ai.store = store_.data();
    ai.store_used = cnt[0];
    ai.rec_start = rec_start_.data();
    ai.rec_len = rec_len_.data();
    ai.records = records;
    ai.group_of.assign(k_, 0);
    std::map<int64_t, int> group_index;
    for (size_t id = 1; id < k_; ++id) {
      if (!has_edges[id]) continue;
      int64_t label = cc_.mapping[id];
      auto it = group_index.find(label);
      if (it == group_index.end()) {
        it = group_index.emplace(label, int(ai.group_labels.size())).first;
        ai.group_labels.push_back(label);
      }
      ai.group_of[id] = it->second;
    }
    std::copy(o_.aniso, o_.aniso + 3, ai.aniso);
    std::copy(origin_, origin_ + 3, ai.origin);
    result_ = assemble(*cx_, cc_.cc_bytes, ai);
    if (!arenas_.empty()) join_arenas(ai.group_of, main_edges, rl, rn, cnt[0]);
    lap("assembly_ms");
    done_ = true;
  }

  void record_counts() const {
    if (!count_) return;
    cx_->stats["flood_voxels"] += double(flood_voxels_);
    cx_->stats["foreground_voxels"] += double(F_);
    cx_->stats["cc_bytes"] = cc_.cc_bytes;
    cx_->stats["components"] += double(count_);
  }

 private:
  void finish_empty() {
    result_ = empty_result(o_);
    done_ = true;
    for (auto &a : arenas_) a->sample->finish_empty();
  }

  // The arenas' skeletons joined to their labels' groups as the final merge joins per-component
  // skeletons (skeletonize.cu:184-198): concatenated in component id order, vertices made unique by
  // position with the first one's radius kept, edges sorted and unique. A group's main-volume part
  // comes merged already; where an arena vertex meets a main vertex, the lower component id (the
  // main one's looked up in the volume) gives the radius. A main path of one vertex has no edge, yet
  // its vertex is in that concatenation when its component has edges, so it takes part as well.
  void join_arenas(const std::vector<int> &group_of, const std::vector<uint8_t> &main_edges, const std::vector<int> &rl,
                   const std::vector<int> &rn, long long store_used) {
    struct Vert {
      std::array<float, 3> p;
      float r;
      int64_t owner;  // component id; -1: a main vertex not looked up yet
    };
    using Edge = std::array<uint32_t, 2>;
    Packed &m = result_;
    const size_t groups = m.labels.size();
    std::vector<std::vector<Vert>> gv(groups);
    std::vector<std::vector<Edge>> ge(groups);
    for (size_t g = 0; g < groups; ++g) {
      for (long long v = m.v_off[g]; v < m.v_off[g + 1]; ++v)
        gv[g].push_back(Vert{{m.vertices[3 * v], m.vertices[3 * v + 1], m.vertices[3 * v + 2]}, m.radii[size_t(v)], -1});
      for (long long e = m.e_off[g]; e < m.e_off[g + 1]; ++e) ge[g].push_back(Edge{m.edges[2 * e], m.edges[2 * e + 1]});
    }
    std::vector<uint8_t> arena_group(groups, 0);
    for (auto &a : arenas_)
      if (!a->sample->result_.labels.empty()) arena_group[size_t(group_of[size_t(a->id)])] = 1;
    std::map<int64_t, int64_t> single;  // voxel of a one-vertex main path in an arena's group -> its component
    const int records = int(rl.size());
    std::vector<int> starts;
    std::vector<int> store;
    for (int r = 0; r < records; ++r) {
      if (rn[r] != 1 || !main_edges[size_t(rl[r])] || !arena_group[size_t(group_of[size_t(rl[r])])]) continue;
      if (starts.empty()) {
        starts = rec_start_.download(*cx_, size_t(records));
        store = store_.download(*cx_, size_t(store_used));
      }
      single[store[size_t(starts[size_t(r)])]] = rl[r];
    }
    auto voxel_of = [&](const std::array<float, 3> &p) {
      Point q;
      for (int ax = 0; ax < 3; ++ax) q[ax] = std::llround(double(p[ax]) / double(o_.aniso[ax])) - origin_[ax];
      return q;
    };
    auto main_radius = [&](int64_t flat) {
      float r = 0;
      BROOK_HIP(hipMemcpyAsync(&r, dbf_.data() + flat, 4, hipMemcpyDeviceToHost, cx_->stream()));
      cx_->synchronize();
      return r;
    };
    for (auto &a : arenas_) {  // ascending id
      const Packed &ar = a->sample->result_;
      if (ar.labels.empty()) continue;
      const size_t g = size_t(group_of[size_t(a->id)]);
      std::vector<Vert> &cur = gv[g];
      std::vector<Vert> in;
      for (long long v = ar.v_off[0]; v < ar.v_off[1]; ++v) {
        Vert x{{ar.vertices[3 * v], ar.vertices[3 * v + 1], ar.vertices[3 * v + 2]}, ar.radii[size_t(v)], a->id};
        const int64_t flat = flat_of(voxel_of(x.p), d_);
        auto it = single.find(flat);
        if (it != single.end() && it->second < a->id) { x.owner = it->second; x.r = main_radius(flat); }
        in.push_back(x);
      }
      std::vector<Vert> out;
      std::vector<uint32_t> map_cur(cur.size()), map_in(in.size());
      size_t i = 0, j = 0;
      while (i < cur.size() || j < in.size()) {
        if (j == in.size() || (i < cur.size() && cur[i].p < in[j].p)) {
          map_cur[i] = uint32_t(out.size());
          out.push_back(cur[i++]);
        } else if (i == cur.size() || in[j].p < cur[i].p) {
          map_in[j] = uint32_t(out.size());
          out.push_back(in[j++]);
        } else {
          Vert &c = cur[i];
          if (c.owner < 0) c.owner = component_at(*cx_, cc_, d_, voxel_of(c.p));
          map_cur[i] = map_in[j] = uint32_t(out.size());
          out.push_back(c.owner < in[j].owner ? c : in[j]);
          ++i, ++j;
        }
      }
      std::vector<Edge> edges;
      for (const Edge &e : ge[g]) edges.push_back(Edge{map_cur[e[0]], map_cur[e[1]]});
      for (long long e = ar.e_off[0]; e < ar.e_off[1]; ++e)
        edges.push_back(Edge{map_in[ar.edges[2 * e]], map_in[ar.edges[2 * e + 1]]});
      for (Edge &e : edges)
        if (e[0] > e[1]) std::swap(e[0], e[1]);
      std::sort(edges.begin(), edges.end());
      edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
      cur.swap(out);
      ge[g].swap(edges);
    }
    m.vertices.clear();
    m.radii.clear();
    m.edges.clear();
    m.v_off.assign(1, 0);
    m.e_off.assign(1, 0);
    for (size_t g = 0; g < groups; ++g) {
      for (const Vert &v : gv[g]) {
        m.vertices.insert(m.vertices.end(), v.p.begin(), v.p.end());
        m.radii.push_back(v.r);
      }
      for (const Edge &e : ge[g]) m.edges.insert(m.edges.end(), e.begin(), e.end());
      m.v_off.push_back(m.v_off.back() + (long long)gv[g].size());
      m.e_off.push_back(m.e_off.back() + (long long)ge[g].size());
    }
  }

  void record_profile() {  // per-label trace profile (device wall clock)
    int khz = 0;
    BROOK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, cx_->device()));
    const double tick_ms = khz > 0 ? 1.0 / khz : 0.0;
    std::vector<unsigned long long> ls = label_stats_.download(*cx_);
    const size_t P = kTraceProfileSlots;
    double sums[kTraceProfileSlots] = {0}, worst = 0;
    size_t worst_id = 0;
    for (int id : labels_) {
      for (size_t j = 0; j < P; ++j) sums[j] += double(ls[P * size_t(id) + j]);
      if (double(ls[P * size_t(id)]) > worst) { worst = double(ls[P * size_t(id)]); worst_id = size_t(id); }
    }
    auto &st = cx_->stats;
    const char *work[6] = {"railroad_phases", "far_scanned", "near_expanded", "invalidate_expanded",
                           "waiting_pulled", "bands"};
    for (int j = 0; j < 6; ++j) {
      st[std::string(work[j]) + "_per_path"] = sums[8 + j] / std::max(1.0, sums[7]);
      st[std::string("slowest_") + work[j]] = double(ls[P * worst_id + 8 + j]);
    }
    const double nl = std::max<double>(1.0, double(labels_.size())), paths = std::max(1.0, sums[7]);
    st["trace_label_max_ms"] = worst * tick_ms;
    st["trace_label_mean_ms"] = sums[0] * tick_ms / nl;
    const char *phase[5] = {"total", "target", "railroad", "backtrace", "invalidate"};
    for (int j = 1; j < 5; ++j) st[std::string("trace_") + phase[j] + "_ms_per_label"] = sums[j] * tick_ms / nl;
    st["railroad_rounds_per_path"] = sums[5] / paths;
    st["invalidate_rounds_per_path"] = sums[6] / paths;
    for (int j = 1; j < 5; ++j) st[std::string("slowest_") + phase[j] + "_ms"] = double(ls[P * worst_id + j]) * tick_ms;
    st["slowest_paths"] = double(ls[P * worst_id + 7]);
    st["slowest_voxels"] = double(boxes_[worst_id].count);
    st["slowest_railroad_rounds"] = double(ls[P * worst_id + 5]);
    st["slowest_invalidate_rounds"] = double(ls[P * worst_id + 6]);
    st["slowest_id"] = double(worst_id);
    if (cx_->kernel_timing)  // per-label detail for diagnostics
      for (int id : labels_) {
        const std::string s = std::to_string(id);
        st["label_ms:" + s] = double(ls[P * id]) * tick_ms;
        st["label_railroad_ms:" + s] = double(ls[P * id + 2]) * tick_ms;
        st["label_paths:" + s] = double(ls[P * id + 7]);
        st["label_near:" + s] = double(ls[P * id + 10]);
        st["label_inval:" + s] = double(ls[P * id + 11]);
        st["label_voxels:" + s] = double(boxes_[id].count);
        st["label_dbfmax:" + s] = double(dbf_max_[id]);
      }
  }

  Context *cx_;    // the context the sample's work runs on (a worker's during a group's first halves)
  Context *main_;  // the caller's context (its auxiliary contexts serve the border and root jobs)
  // The border job of the first preamble half (joined in the second, or by the destructor).
  struct EventGuard {
    hipEvent_t e = nullptr;
    ~EventGuard() { if (e) (void)hipEventDestroy(e); }
  } cc_ready_;
  bool early_borders_ = false, borders_ready_ = false;
  double border_ms_[2] = {0, 0};  // device and host milliseconds of the border job
  std::future<std::vector<std::vector<Point>>> border_job_;
  // The early root searches: the records and voxel lists they built, and per id the root found, or
  // -1 (written by the job, read after joining it).
  Buffer<VoxelRec> early_rec_;
  Buffer<int> early_seg_start_, early_vox_;
  std::future<std::vector<long long>> root_job_;
  LabelVolume in_;
  const SkeletonizeOptions &o_;
  FieldsOut *fields_;
  const Dims d_;
  const int64_t n_;
  const Arena *arena_;        // this sample is that arena (null: an input volume)
  int origin_[3] = {0, 0, 0};  // the sample's corner in the input volume
  bool done_ = false;
  Packed result_;
  std::vector<std::unique_ptr<Arena>> arenas_;  // somata traced in private arenas, ascending id

  Components cc_;
  std::vector<int64_t> keys_;  // voxel graph: Kimimaro's ids of the components (until fix_avocados)
  int64_t count_ = 0, F_ = 0, flood_voxels_ = 0;
  size_t k_ = 0;
  std::vector<Box> boxes_;
  std::vector<float> dbf_max_, max_daf_;
  std::vector<std::vector<Point>> borders_, before_, after_;
  std::vector<uint8_t> active_;
  std::vector<int> labels_;
  std::vector<long long> root_, target_;
  Buffer<float> dbf_, daf_, weights_, max_daf_dev_;
  Buffer<int> seg_start_dev_, vox_, seg_count_dev_;
  Buffer<long long> first_dev_, root_dev_, target_dev_;
  Buffer<VoxelRec> rec_;
  Buffer<unsigned> inmask_;  // voxel graph: pull masks (empty: no graph)
  std::vector<Buffer<int>> lists_;
  Buffer<int> order_;
  Buffer<int> parent_;  // fix_branching=False: parent field
  std::vector<uint8_t> soma_;       // per id: a soma
  Buffer<float> free_radius_dev_;   // per id: the soma's free-space radius (0: not a soma); empty: no somata
  Buffer<SomaLabel> soma_dev_;      // per id: soma parameters; empty: no somata
  Buffer<unsigned long long> snap_a_, snap_b_;
  Buffer<int> mstart_dev_, mcount_dev_, acount_dev_, maxp_dev_, mflat_dev_;
  long long store_cap_ = 0;
  int rec_cap_ = 0;
  Buffer<int> store_, rec_label_, rec_len_, rec_start_, counters_, ejected_;
  Buffer<unsigned long long> label_stats_;
};

bool larger_first(const LabelItem &a, const LabelItem &b) { return a.size > b.size; }

// Runs fn(i, context, lap) for every i < n on kWorkers host threads, worker w taking i = w,
// w + kWorkers, ... on auxiliary context kWorkerBase + w (each a stream and buffer pool of its
// own); adds the workers' stage times to ctx.stats and rethrows the first error once all ended.
// Blocks a worker frees from other pools than its own (this context's, another worker's) go back
// once every worker's stream is idle.
template <class F> void on_workers(Context &ctx, size_t n, F fn) {
  std::vector<std::unique_ptr<FreeHold>> holds;
  for (int w = 0; w < kWorkers; ++w) {  // made here: the threads only look them up
    Context &wc = ctx.aux(kWorkerBase + w);
    wc.kernel_timing = ctx.kernel_timing;
    holds.push_back(std::make_unique<FreeHold>(wc));
  }
  ctx.synchronize();  // what this context's stream did is complete before the workers read it
  std::vector<std::future<void>> workers;
  for (int w = 0; w < kWorkers; ++w)
    workers.push_back(std::async(std::launch::async, [&ctx, &fn, &holds, n, w] {
      Context &wc = ctx.aux(kWorkerBase + w);
      std::lock_guard<std::mutex> turn(wc.mutex());
      wc.activate();
      FreeHoldScope held(*holds[size_t(w)]);
      Lap wl{wc};
      for (size_t i = size_t(w); i < n; i += kWorkers) fn(i, wc, wl);
      wc.synchronize();
    }));
  std::exception_ptr err;
  for (auto &f : workers) {
    try {
      f.get();
    } catch (...) {
      if (!err) err = std::current_exception();
    }
  }
  for (int w = 0; w < kWorkers; ++w) {
    Context &wc = ctx.aux(kWorkerBase + w);
    if (err) (void)hipStreamSynchronize(wc.stream());  // a worker that threw left its stream running
    for (const auto &kv : wc.stats) ctx.stats[kv.first] += kv.second;  // stage times and counts, summed
    wc.stats.clear();
  }
  for (auto &h : holds) h->release();
  if (err) std::rethrow_exception(err);
}

// Runs a group of samples together: every preamble, one flood launch over all their labels
// (largest first), every PDRF/order pass, one trace launch, every assembly. In a group of at
// least kWorkerMin samples the per-sample stages (preambles, PDRF/order, assembly; latency-bound
// for small volumes: kernel launches and host reads) of different samples overlap on the worker
// contexts; the launches shared by the group run on this context.
void run_group(Context &ctx, std::vector<std::unique_ptr<Sample>> &group, Lap &lap) {
  const bool workers = group.size() >= size_t(kWorkerMin);
  if (workers) {
    on_workers(ctx, group.size(), [&group](size_t i, Context &wc, Lap &wl) {
      group[i]->use_context(wc);
      group[i]->preamble_components(wl, -1, false);  // border targets in place
      group[i]->preamble_labels(wl);                 // arenas made here run their preambles here too
    });
    for (auto &s : group) s->use_context_all(ctx);
    lap("prep_preambles_ms");
  } else {
    for (auto &s : group) s->preamble_components(lap, kBorderAux, group.size() == 1);
    for (auto &s : group) s->preamble_labels(lap);
  }
  std::vector<Sample *> all;  // the samples and their private arenas
  for (auto &s : group) s->collect(all);
  for (Sample *s : all) s->record_counts();
  std::vector<LabelItem> items;
  for (size_t i = 0; i < all.size(); ++i) all[i]->items(int(i), items);
  std::stable_sort(items.begin(), items.end(), larger_first);
  std::vector<FloodJob> jobs;
  for (Sample *s : all) jobs.push_back(s->done() ? FloodJob{} : s->flood_job());
  prepare_fields(ctx, jobs, items);
  lap("prep_floods_ms");
  if (workers) {
    on_workers(ctx, all.size(), [&all](size_t i, Context &wc, Lap &wl) {
      all[i]->use_context(wc);
      all[i]->after_floods(wl);
    });
    for (Sample *s : all) s->use_context(ctx);
    lap("prep_after_floods_ms");
  } else {
    for (Sample *s : all) s->after_floods(lap);
  }
  items.clear();
  for (size_t i = 0; i < all.size(); ++i) all[i]->items(int(i), items);
  std::stable_sort(items.begin(), items.end(), larger_first);
  std::vector<TraceArgs> args;
  for (Sample *s : all) args.push_back(s->done() ? TraceArgs{} : s->trace_args());
  std::vector<LabelItem> somata;
  for (size_t i = 0; i < all.size(); ++i) all[i]->soma_items(int(i), somata);
  trace_all(ctx, args, items, somata);
  lap("trace_ms");
  if (workers) {
    on_workers(ctx, group.size(), [&group](size_t i, Context &wc, Lap &wl) {
      group[i]->use_context_all(wc);  // its arenas assemble in its finish
      group[i]->finish(wl, false);
    });
    for (auto &s : group) s->use_context_all(ctx);
    lap("assembly_wall_ms");
  } else {
    const bool profile = group.size() == 1;
    for (auto &s : group) s->finish(lap, profile);
  }
}

}  // namespace

Packed skeletonize(Context &ctx, const LabelVolume &in, const SkeletonizeOptions &o, FieldsOut *fields) {
  ctx.stats.clear();
  check_options(o);
  Lap lap{ctx};
  std::vector<std::unique_ptr<Sample>> group;
  group.push_back(std::make_unique<Sample>(ctx, in, o, fields));
  run_group(ctx, group, lap);
  return group[0]->take_result();
}

std::vector<int> target_orders(Context &ctx, const LabelVolume &ids, const float *daf) {
  if (ids.bytes > 4) throw std::invalid_argument("target orders take component ids of at most 32 bits");
  const int64_t n = ids.size();
  std::vector<int> out;
  if (n == 0) return out;
  const size_t b = size_t(ids.bytes);
  std::vector<uint8_t> raw(size_t(n) * b);
  BROOK_HIP(hipMemcpyAsync(raw.data(), ids.data, raw.size(), hipMemcpyDeviceToHost, ctx.stream()));
  ctx.synchronize();
  int64_t F = 0, top = 0;
  for (int64_t i = 0; i < n; ++i) {
    uint32_t v = 0;
    std::memcpy(&v, raw.data() + size_t(i) * b, b);
    F += v != 0;
    top = std::max<int64_t>(top, v);
  }
  if (F == 0) return out;
  Buffer<float> daf_dev(ctx, size_t(n));
  daf_dev.upload(ctx, daf, size_t(n));
  Buffer<int> order(ctx, size_t(F));
  target_order(ctx, ids.data, ids.bytes, n, daf_dev.data(), F, top, order.data());
  return order.download(ctx);
}

std::vector<Packed> skeletonize_batch(Context &ctx, const std::vector<LabelVolume> &volumes, const SkeletonizeOptions &o) {
  ctx.stats.clear();
  check_options(o);
  Lap lap{ctx};
  size_t free_bytes = 0, total_bytes = 0;
  BROOK_HIP(hipMemGetInfo(&free_bytes, &total_bytes));
  const int64_t budget = std::max<int64_t>(int64_t(free_bytes) / 2, int64_t(1) << 30);
  std::vector<Packed> out;
  out.reserve(volumes.size());
  size_t i = 0;
  while (i < volumes.size()) {  // groups within half of the free device memory (at least one sample)
    std::vector<std::unique_ptr<Sample>> group;
    int64_t bytes = 0;
    while (i < volumes.size()) {
      const int64_t need = Sample::footprint(volumes[i].size());
      if (!group.empty() && bytes + need > budget) break;
      group.push_back(std::make_unique<Sample>(ctx, volumes[i], o, nullptr));
      bytes += need;
      ++i;
    }
    run_group(ctx, group, lap);
    for (auto &s : group) out.push_back(s->take_result());
  }
  return out;
}

}  // namespace brook
