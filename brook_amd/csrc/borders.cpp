// Border targets (fix_borders): for every volume face, 2-D components of the face's component
// ids and their 2-D EDT are computed on the GPU with the volume kernels (the six faces on one
// canvas: one CCL, one EDT per block of a face or a face pair); the farthest-pixel pick with its tie keys and the
// insertion-ordered point sets are host work, as in the reference's serial plan (borders.cu:53-116,
// point_set.hpp:11-55).
#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <future>
#include <limits>

#include "borders.hpp"

namespace brook {
namespace {

// Insertion-only set of integer 3-tuples with CPython's tuple hash and set probing (the order in
// which Kimimaro's Python set yields its border targets): xxHash-style tuple hash, 9 linear
// probes then perturbed jumps, growth to 4x (2x above 50000) of the used count at 3/5 fill.
class TupleSet {
 public:
  void add(const Point &p) {
    if (!place(p, hash(p))) return;
    ++used_;
    if (used_ * 5 < (slots_.size() - 1) * 3) return;
    size_t want = used_ * (used_ > 50000 ? 2 : 4), size = 8;
    while (size <= want) size <<= 1;
    std::vector<Slot> old(size);
    old.swap(slots_);
    for (const Slot &s : old)
      if (s.used) place(s.p, s.h);
  }
  std::vector<Point> items() const {
    std::vector<Point> out;
    for (const Slot &s : slots_)
      if (s.used) out.push_back(s.p);
    return out;
  }

 private:
  struct Slot {
    Point p{};
    uint64_t h = 0;
    bool used = false;
  };
  std::vector<Slot> slots_ = std::vector<Slot>(8);
  size_t used_ = 0;

  static uint64_t int_hash(int64_t v) {
    const uint64_t modulus = (uint64_t(1) << 61) - 1;
    uint64_t magnitude = v < 0 ? uint64_t(-(v + 1)) + 1 : uint64_t(v);
    int64_t h = int64_t(magnitude % modulus);
    if (v < 0) h = -h;
    return uint64_t(h == -1 ? -2 : h);
  }
  static uint64_t hash(const Point &p) {
    const uint64_t prime1 = 11400714785074694791ull, prime2 = 14029467366897019727ull, prime5 = 2870177450012600261ull;
    uint64_t acc = prime5;
    for (int64_t v : p) {
      acc += int_hash(v) * prime2;
      acc = (acc << 31) | (acc >> 33);
      acc *= prime1;
    }
    acc += uint64_t(p.size()) ^ (prime5 ^ 3527539ull);
    return acc == ~uint64_t(0) ? 1546275796ull : acc;
  }
  bool place(const Point &p, uint64_t h) {
    const size_t mask = slots_.size() - 1;
    size_t i = h & mask;
    uint64_t perturb = h;
    for (;;) {
      const size_t run = i + 9 <= mask ? 9 : 0;
      for (size_t j = 0; j <= run; ++j) {
        Slot &s = slots_[i + j];
        if (!s.used) { s.p = p; s.h = h; s.used = true; return true; }
        if (s.h == h && s.p == p) return false;
      }
      perturb >>= 5;
      i = (i * 5 + 1 + perturb) & mask;
    }
  }
};

float sq(float x, float y, float a, float b, float wx, float wy) {
  float dx = wx * (x - a), dy = wy * (y - b);
  return dx * dx + dy * dy;
}

// Tie keys of a farthest pixel: distance to the component's centre pixel, to the face centre,
// to the nearest corner, then the edge distance (borders.cu:42-51).
std::array<float, 4> tie_keys(int64_t ix, int64_t iy, float cx, float cy, int64_t nx, int64_t ny, float wx, float wy) {
  float x = float(ix), y = float(iy), fx = float(nx), fy = float(ny);
  float mid_x = float(double(wx * fx) / 2.0), mid_y = float(double(wy * fy) / 2.0);
  float far_x = float(double(fx) - 0.5), far_y = float(double(fy) - 0.5);
  float corner = std::min({sq(x, y, -0.5f, -0.5f, wx, wy), sq(x, y, far_x, -0.5f, wx, wy), sq(x, y, far_x, far_y, wx, wy),
                           sq(x, y, -0.5f, far_x, wx, wy)});
  float edge = float(std::min({double(wx) * (double(x) - 0.5), double(wx) * (double(fx) - 0.5 - double(x)),
                               double(wy) * (double(y) - 0.5), double(wy) * (double(fy) - 0.5 - double(y))}));
  return {sq(x, y, cx, cy, wx, wy), sq(x, y, mid_x, mid_y, wx, wy), corner, edge};
}

// One face's host pass (borders.cu:53-116): float32 centroid sums in Kimimaro's order (x outer),
// the farthest pixels' tie keys, and the picks ordered by each component's first pixel with a
// positive distance (x + fx*y). Pixel i = y + fy*x, so both loops walk memory in order; equal keys
// keep the pixel first in (y, x) order, as the reference's y-outer scan does.
struct FacePick {
  int64_t first;
  int64_t id;  // the 2-D component
  int64_t x, y;
};
template <class T>
std::vector<FacePick> face_picks(const std::vector<uint8_t> &raw, const std::vector<float> &dv, size_t k, int64_t fx,
                                 int64_t fy, float wx, float wy) {
  const T *lab = reinterpret_cast<const T *>(raw.data());
  std::vector<float> maxdist(k, 0.0f), sumx(k, 0.0f), sumy(k, 0.0f);
  std::vector<uint32_t> size(k, 0);
  std::vector<int64_t> first(k, std::numeric_limits<int64_t>::max());
  for (int64_t x = 0; x < fx; ++x)
    for (int64_t y = 0; y < fy; ++y) {
      const size_t i = size_t(y + fy * x);
      const size_t l = size_t(lab[i]);
      if (!l) continue;
      sumx[l] += float(x);
      sumy[l] += float(y);
      ++size[l];
      if (dv[i] > 0) {
        maxdist[l] = std::max(maxdist[l], dv[i]);
        first[l] = std::min(first[l], x + fx * y);
      }
    }
  std::vector<float> cxs(k, 0.0f), cys(k, 0.0f);
  for (size_t l = 1; l < k; ++l) {
    if (!size[l]) continue;
    float px = (wx * sumx[l]) / float(size[l]), py = (wy * sumy[l]) / float(size[l]);
    if (px - (wx * float(fx)) / 2.0f < 0) px = px + wx;
    if (py - (wy * float(fy)) / 2.0f < 0) py = py + wy;
    cxs[l] = float(int(px / wx));
    cys[l] = float(int(py / wy));
  }
  std::vector<std::array<float, 4>> best(k);
  std::vector<int64_t> bx(k, -1), by(k, -1);
  for (int64_t x = 0; x < fx; ++x)
    for (int64_t y = 0; y < fy; ++y) {
      const size_t i = size_t(y + fy * x);
      const size_t l = size_t(lab[i]);
      if (!l || dv[i] == 0 || dv[i] != maxdist[l]) continue;
      const auto key = tie_keys(x, y, cxs[l], cys[l], fx, fy, wx, wy);
      if (bx[l] < 0 || key < best[l] || (key == best[l] && (y < by[l] || (y == by[l] && x < bx[l])))) {
        best[l] = key;
        bx[l] = x;
        by[l] = y;
      }
    }
  std::vector<FacePick> out;
  for (size_t l = 1; l < k; ++l)
    if (bx[l] >= 0) out.push_back(FacePick{first[l], int64_t(l), bx[l], by[l]});
  std::sort(out.begin(), out.end(), [](const FacePick &u, const FacePick &v) { return u.first < v.first; });
  return out;
}

}  // namespace

std::vector<std::vector<Point>> border_targets(Context &ctx, const void *cc, int cc_bytes, Dims d, int64_t count,
                                               const float aniso[3]) {
  std::vector<TupleSet> sets(size_t(count) + 1);
  auto items = [&sets]() {
    std::vector<std::vector<Point>> out;
    out.reserve(sets.size());
    for (auto &s : sets) out.push_back(s.items());
    return out;
  };
  const int axes[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
  const int dims[3] = {d.sx, d.sy, d.sz};
  using Clock = std::chrono::steady_clock;
  auto ms = [](Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  // The faces' CCL and EDT kernels share the volume kernels' timer names; their time belongs to
  // this stage, so per-kernel timing pauses here (the stage's own laps still count it).
  struct TimingPause {
    Context &ctx;
    bool was;
    explicit TimingPause(Context &c) : ctx(c), was(c.kernel_timing) { c.kernel_timing = false; }
    ~TimingPause() { ctx.kernel_timing = was; }
  } pause(ctx);
  const auto t0 = Clock::now();
  // All six faces on one canvas, zero rows and columns between them: one 2-D CCL labels every
  // face's components at once. The canvas is cut into blocks of whole columns, each holding one
  // face or a pair of opposite faces stacked along the rows (near face, a zero row, far face);
  // every block gets one EDT, which sees the same lines as a face's own EDT would (x runs only by
  // their offsets, y lines from the face's first column), so the distances are the per-face ones.
  // The row count is the tallest block: a pair is stacked when it fits in rows the EDT's x pass
  // takes a wave per row for (kWaveRows); otherwise its faces get a block each.
  constexpr int kWaveRows = 2048;
  FaceCanvas fc{};
  for (int f = 0; f < 6; ++f) {
    fc.a[f] = axes[f / 2][0];
    fc.b[f] = axes[f / 2][1];
    fc.c[f] = axes[f / 2][2];
  }
  for (int o = 0; o < 3; ++o) {
    const int fy = dims[fc.b[2 * o]];
    fc.rows = std::max(fc.rows, 2 * fy + 1 <= kWaveRows ? 2 * fy + 1 : fy);
  }
  struct Block {
    int col0, cols, f0, f1;  // columns, faces f0 .. f1-1
  };
  std::vector<Block> blocks;
  int col = 0;
  for (int o = 0; o < 3; ++o) {
    const int f = 2 * o, fx = dims[fc.a[f]], fy = dims[fc.b[f]];
    if (2 * fy + 1 <= fc.rows) {
      fc.col0[f] = fc.col0[f + 1] = col;
      fc.row0[f] = 0;
      fc.row0[f + 1] = fy + 1;
      blocks.push_back(Block{col, fx, f, f + 2});
      col += fx + 1;
    } else {
      for (int h = 0; h < 2; ++h) {
        fc.col0[f + h] = col;
        fc.row0[f + h] = 0;
        blocks.push_back(Block{col, fx, f + h, f + h + 1});
        col += fx + 1;
      }
    }
  }
  fc.cols = col - 1;
  const size_t np = size_t(fc.rows) * size_t(fc.cols);
  Buffer<uint8_t> canvas(ctx, np * size_t(cc_bytes));
  extract_faces(ctx, cc, cc_bytes, d, fc, canvas.data());
  LabelVolume pv;
  pv.data = canvas.data();
  pv.bytes = cc_bytes;
  pv.dims = Dims{fc.rows, fc.cols, 1};
  Components c2 = connected_components(ctx, pv);
  if (c2.count == 0) return items();
  const int bytes = c2.cc_bytes;
  Buffer<float> dist(ctx, np);
  // Block by block: its EDT, its faces gathered (pixel y + fy * x) and read, and their host passes
  // started (the larger faces on threads of their own) while the next block's device work runs.
  struct Face {
    std::vector<uint8_t> raw;
    std::vector<float> dv;
    int a, b, c, fixed, fx, fy;
  };
  std::vector<Face> faces(6);
  size_t most = 0;
  for (int f = 0; f < 6; ++f) {
    Face &face = faces[size_t(f)];
    face.a = fc.a[f];
    face.b = fc.b[f];
    face.c = fc.c[f];
    face.fixed = f % 2 ? dims[face.c] - 1 : 0;
    face.fx = dims[face.a];
    face.fy = dims[face.b];
    most = std::max(most, 2 * size_t(face.fx) * size_t(face.fy));
  }
  const size_t k = size_t(c2.count) + 1;
  auto pass = [&faces, &aniso, k, bytes](int f) {
    const Face &face = faces[size_t(f)];
    const float wx = aniso[face.a], wy = aniso[face.b];
    if (bytes == 1) return face_picks<uint8_t>(face.raw, face.dv, k, face.fx, face.fy, wx, wy);
    if (bytes == 2) return face_picks<uint16_t>(face.raw, face.dv, k, face.fx, face.fy, wx, wy);
    return face_picks<uint32_t>(face.raw, face.dv, k, face.fx, face.fy, wx, wy);
  };
  constexpr size_t kThreadPixels = size_t(1) << 16;
  std::vector<std::future<std::vector<FacePick>>> jobs(6);
  std::vector<std::vector<FacePick>> picks(6);
  Buffer<uint8_t> block_ids(ctx, most * size_t(bytes));
  Buffer<float> block_dist(ctx, most);
  for (const Block &blk : blocks) {
    const int a = fc.a[blk.f0], b = fc.b[blk.f0];
    const size_t at = size_t(blk.col0) * size_t(fc.rows);
    const float pa[3] = {aniso[b], aniso[a], 1.0f};
    edt(ctx, c2.labels.data() + at * size_t(bytes), bytes, Dims{fc.rows, blk.cols, 1}, pa, true, 2, dist.data() + at);
    gather_faces(ctx, c2.labels.data(), bytes, dist.data(), d, fc, blk.f0, blk.f1, block_ids.data(), block_dist.data());
    const size_t pixels = size_t(dims[a]) * size_t(dims[b]), nf = size_t(blk.f1 - blk.f0);
    const std::vector<uint8_t> ids = block_ids.download(ctx, nf * pixels * size_t(bytes));
    const std::vector<float> dv = block_dist.download(ctx, nf * pixels);
    for (size_t h = 0; h < nf; ++h) {
      const int f = blk.f0 + int(h);
      Face &face = faces[size_t(f)];
      face.raw.assign(ids.begin() + ptrdiff_t(h * pixels * size_t(bytes)), ids.begin() + ptrdiff_t((h + 1) * pixels * size_t(bytes)));
      face.dv.assign(dv.begin() + ptrdiff_t(h * pixels), dv.begin() + ptrdiff_t((h + 1) * pixels));
      if (pixels >= kThreadPixels) jobs[size_t(f)] = std::async(std::launch::async, pass, f);
    }
  }
  const auto t1 = Clock::now();
  ctx.stats["borders_device_ms"] += ms(t0, t1);  // canvas, 2-D CCL and EDTs, downloads
  for (int f = 0; f < 6; ++f)
    if (!jobs[size_t(f)].valid()) picks[size_t(f)] = pass(f);
  for (int f = 0; f < 6; ++f)
    if (jobs[size_t(f)].valid()) picks[size_t(f)] = jobs[size_t(f)].get();
  for (int f = 0; f < 6; ++f) {
    const Face &face = faces[size_t(f)];
    for (const FacePick &q : picks[size_t(f)]) {
      Point p{};
      p[face.a] = q.x;
      p[face.b] = q.y;
      p[face.c] = face.fixed;
      sets.at(size_t(c2.mapping[size_t(q.id)])).add(p);
    }
  }
  ctx.stats["borders_host_ms"] += ms(t1, Clock::now());  // the host passes
  return items();
}

}  // namespace brook
