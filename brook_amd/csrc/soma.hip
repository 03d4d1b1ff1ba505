// Soma preparation kernels: the enclosed-void check (6-connected union-find over the background of
// a label's bounding box, with the components that touch the box's faces marked) and the DBF
// maxima of soma labels (a flagged selection in flat order).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstring>

#include "graph.hpp"
#include "soma.hpp"

namespace brook {
namespace {

constexpr int kThreads = 256;
unsigned blocks_for(long long n) { return unsigned(std::max<long long>((n + kThreads - 1) / kThreads, 1)); }

struct Crop {
  int lo[3], ext[3];
  __host__ __device__ long long size() const { return (long long)ext[0] * ext[1] * ext[2]; }
};

template <class CC> __device__ bool is_label(const CC *cc, Dims d, const Crop &c, int x, int y, int z, CC id) {
  return cc[(c.lo[0] + x) + (long long)d.sx * ((c.lo[1] + y) + (long long)d.sy * (c.lo[2] + z))] == id;
}

// The box voxels that are not the label, labelled by 6-connected union-find over box-local
// indices i = x + ex * (y + ey * z): every x-run hooks to its first voxel by a wave ballot, runs
// touching in the y-1 and z-1 rows are united (larger root -> smaller), the roots of face voxels
// are marked, and a voxel is reached when its root is marked. No host round trip.
constexpr int kOut = -1;  // parent of a label voxel (not part of the background mask)

template <class CC>
__device__ __forceinline__ bool open_at(const CC *cc, Dims d, const Crop &c, long long i, CC id) {
  const int x = int(i % c.ext[0]), y = int((i / c.ext[0]) % c.ext[1]), z = int(i / ((long long)c.ext[0] * c.ext[1]));
  return !is_label(cc, d, c, x, y, z, id);
}

template <class CC>
__global__ void box_runs(const CC *cc, Dims d, Crop c, CC id, int *parent) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = i < c.size();
  const bool open = valid && open_at(cc, d, c, i, id);
  const int x = valid ? int(i % c.ext[0]) : 0;
  int prev = shfl_up_any(int(open), 1);
  if (lane == 0 && valid && x > 0) prev = int(open_at(cc, d, c, i - 1, id));
  const bool start = open && (x == 0 || !prev);
  const unsigned long long mask = __ballot(start);
  if (!valid) return;
  if (!open) { parent[i] = kOut; return; }
  const unsigned long long below = mask & lanes_upto(lane);
  const long long base = i - lane;
  parent[i] = int(below ? base + (63 - __clzll((long long)below)) : base - 1);  // base - 1: same run
}

__device__ __forceinline__ int box_root(int *parent, int v) {
  int cur = ld_i(parent + v);
  if (cur != v) {
    int prev = v, next;
    while (cur > (next = ld_i(parent + cur))) {
      st_i(parent + prev, next);  // path halving: only ever points closer to the root
      prev = cur;
      cur = next;
    }
  }
  return cur;
}

__device__ __forceinline__ void box_unite(int *parent, int a, int b) {
  int va = box_root(parent, a), vb = box_root(parent, b);
  while (va != vb) {
    if (va < vb) {
      const int ret = atomicCAS(parent + vb, vb, va);
      if (ret == vb) return;
      vb = box_root(parent, ret);
    } else {
      const int ret = atomicCAS(parent + va, va, vb);
      if (ret == va) return;
      va = box_root(parent, ret);
    }
  }
}

// Face contacts between runs: voxel x of a run unites with the run at x in the y-1 row and in the
// z-1 plane where that run begins at x or x is this run's first voxel.
__global__ void box_union(Crop c, int *parent) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= c.size() || parent[i] == kOut) return;
  const int ex = c.ext[0], ey = c.ext[1];
  const int x = int(i % ex), y = int((i / ex) % ey), z = int(i / ((long long)ex * ey));
  const bool first = x == 0 || parent[i - 1] == kOut;
  const long long rows[2] = {y > 0 ? i - ex : -1, z > 0 ? i - (long long)ex * ey : -1};
  for (int r = 0; r < 2; ++r) {
    const long long j = rows[r];
    if (j < 0 || parent[j] == kOut) continue;
    if (first || parent[j - 1] == kOut) box_unite(parent, int(i), int(j));  // x > 0 when !first
  }
}

__global__ void box_mark_faces(Crop c, int *parent, unsigned *reached) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= c.size() || parent[i] == kOut) return;
  const int x = int(i % c.ext[0]), y = int((i / c.ext[0]) % c.ext[1]), z = int(i / ((long long)c.ext[0] * c.ext[1]));
  const bool face = x == 0 || y == 0 || z == 0 || x == c.ext[0] - 1 || y == c.ext[1] - 1 || z == c.ext[2] - 1;
  if (face) g_st(reached + box_root(parent, int(i)), 1u);
}

__global__ void box_resolve(Crop c, int *parent, unsigned *reached) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= c.size() || parent[i] == kOut) return;
  const int r = box_root(parent, int(i));
  if (r != int(i) && g_ld(reached + r)) g_st(reached + i, 1u);
}

// reached[i] = 1 for the box voxels that are not the label and reach a face of the box through
// such voxels (face neighbours).
template <class CC> void reach_faces(Context &ctx, const CC *cc, Dims d, CC id, const Crop &c, Buffer<unsigned> &reached) {
  auto s = ctx.stream();
  const long long n = c.size();
  if (n > INT32_MAX) throw std::invalid_argument("label bounding box exceeds int32 capacity");
  reached = Buffer<unsigned>(ctx, size_t(n));
  Buffer<int> parent(ctx, size_t(n));
  reached.clear(ctx);
  const unsigned grid = blocks_for(n);
  box_runs<CC><<<grid, kThreads, 0, s>>>(cc, d, c, id, parent.data());
  box_union<<<grid, kThreads, 0, s>>>(c, parent.data());
  box_mark_faces<<<grid, kThreads, 0, s>>>(c, parent.data(), reached.data());
  box_resolve<<<grid, kThreads, 0, s>>>(c, parent.data(), reached.data());
  BROOK_LAUNCHED();
}

template <class CC>
__global__ void void_tally(const CC *cc, Dims d, Crop c, CC id, const unsigned *reached, unsigned long long *voids) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  bool hole = false;
  if (i < c.size()) {
    const int x = int(i % c.ext[0]), y = int((i / c.ext[0]) % c.ext[1]), z = int(i / ((long long)c.ext[0] * c.ext[1]));
    hole = !is_label(cc, d, c, x, y, z, id) && reached[i] == 0u;
  }
  const int n = __popcll(__ballot(hole));
  if ((threadIdx.x & 63) == 0 && n) g_add(voids, (unsigned long long)n);
}

template <class CC>
int64_t voids_of(Context &ctx, const CC *cc, Dims d, CC id, const Crop &c) {
  Buffer<unsigned> reached;
  reach_faces(ctx, cc, d, id, c, reached);
  Buffer<unsigned long long> voids(ctx, 1);
  voids.clear(ctx);
  void_tally<CC><<<blocks_for(c.size()), kThreads, 0, ctx.stream()>>>(cc, d, c, id, reached.data(), voids.data());
  BROOK_LAUNCHED();
  return int64_t(voids.download(ctx)[0]);
}

// Box voxels: mask = the label or a void; a void voxel's previous owner is counted.
template <class CC>
__global__ void void_owners(const CC *cc, Dims d, Crop c, CC id, const unsigned *reached, uint8_t *mask,
                            unsigned long long *taken) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= c.size()) return;
  const int x = int(i % c.ext[0]), y = int((i / c.ext[0]) % c.ext[1]), z = int(i / ((long long)c.ext[0] * c.ext[1]));
  const CC l = cc[(c.lo[0] + x) + (long long)d.sx * ((c.lo[1] + y) + (long long)d.sy * (c.lo[2] + z))];
  const bool hole = l != id && reached[i] == 0u;
  mask[i] = (l == id || hole) ? 1 : 0;
  if (hole && l != CC(0)) g_add(taken + (unsigned long long)l, 1ULL);
}

template <class CC> __global__ void paint_voids(CC *cc, Dims d, Crop c, CC id, const uint8_t *mask) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= c.size() || !mask[i]) return;
  const int x = int(i % c.ext[0]), y = int((i / c.ext[0]) % c.ext[1]), z = int(i / ((long long)c.ext[0] * c.ext[1]));
  cc[(c.lo[0] + x) + (long long)d.sx * ((c.lo[1] + y) + (long long)d.sy * (c.lo[2] + z))] = id;
}

// The graph words of the box (Fortran order over the box).
__global__ void crop_words(Dims d, Crop c, const unsigned *graph, unsigned *out) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= c.size()) return;
  const int x = int(i % c.ext[0]), y = int((i / c.ext[0]) % c.ext[1]), z = int(i / ((long long)c.ext[0] * c.ext[1]));
  out[i] = graph[(c.lo[0] + x) + (long long)d.sx * ((c.lo[1] + y) + (long long)d.sy * (c.lo[2] + z))];
}

__global__ void overlay_dbf(Dims d, Crop c, const uint8_t *mask, const float *box_dbf, float *dbf) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= c.size() || !mask[i]) return;
  const int x = int(i % c.ext[0]), y = int((i / c.ext[0]) % c.ext[1]), z = int(i / ((long long)c.ext[0] * c.ext[1]));
  dbf[(c.lo[0] + x) + (long long)d.sx * ((c.lo[1] + y) + (long long)d.sy * (c.lo[2] + z))] = box_dbf[i];
}

template <class CC>
int64_t fill_of(Context &ctx, const CC *cc, Dims d, CC id, const Crop &c, size_t ids, Buffer<uint8_t> &mask,
                std::vector<int64_t> &taken) {
  auto s = ctx.stream();
  const long long n = c.size();
  Buffer<unsigned> reached;
  reach_faces(ctx, cc, d, id, c, reached);
  Buffer<unsigned long long> owners(ctx, ids), holes(ctx, 1);
  owners.clear(ctx);
  holes.clear(ctx);
  mask = Buffer<uint8_t>(ctx, size_t(n));
  void_owners<CC><<<blocks_for(n), kThreads, 0, s>>>(cc, d, c, id, reached.data(), mask.data(), owners.data());
  BROOK_LAUNCHED();
  void_tally<CC><<<blocks_for(n), kThreads, 0, s>>>(cc, d, c, id, reached.data(), holes.data());
  BROOK_LAUNCHED();
  const std::vector<unsigned long long> o = owners.download(ctx);
  taken.assign(o.begin(), o.end());
  return int64_t(holes.download(ctx)[0]);
}

template <class CC> __global__ void first_of(const CC *cc, long long n, unsigned long long *first) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const CC l = cc[i];
  if (l != CC(0)) g_min(first + (unsigned long long)l, (unsigned long long)i);
}

template <class CC> struct AtMaximum {
  const CC *cc;
  const float *dbf, *dbf_max;
  const uint8_t *soma;
  __device__ bool operator()(const long long &v) const {
    const CC l = cc[v];
    return l != CC(0) && soma[l] && dbf[v] == dbf_max[l];
  }
};

template <class CC>
__global__ void ids_at(const CC *cc, const long long *voxels, const int *count, long long *ids) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j < *count) ids[j] = (long long)cc[voxels[j]];
}

template <class CC>
std::vector<Maximum> maxima_of(Context &ctx, const CC *cc, long long n, const float *dbf, const float *dbf_max,
                               const uint8_t *soma) {
  auto s = ctx.stream();
  Buffer<long long> voxels(ctx, size_t(n)), ids(ctx, size_t(n));
  Buffer<int> count(ctx, 1);
  hipcub::CountingInputIterator<long long> index(0);
  const AtMaximum<CC> pick{cc, dbf, dbf_max, soma};
  size_t bytes = 0;
  BROOK_HIP(hipcub::DeviceSelect::If(nullptr, bytes, index, voxels.data(), count.data(), n, pick, s));
  Buffer<unsigned char> scratch(ctx, std::max<size_t>(bytes, 1));
  BROOK_HIP(hipcub::DeviceSelect::If(scratch.data(), bytes, index, voxels.data(), count.data(), n, pick, s));
  ids_at<CC><<<blocks_for(n), kThreads, 0, s>>>(cc, voxels.data(), count.data(), ids.data());
  BROOK_LAUNCHED();
  const int total = count.download(ctx)[0];
  std::vector<long long> v = voxels.download(ctx, size_t(total)), l = ids.download(ctx, size_t(total));
  std::vector<Maximum> out;
  out.reserve(size_t(total));
  for (int i = 0; i < total; ++i) out.push_back(Maximum{v[size_t(i)], l[size_t(i)]});
  return out;
}

}  // namespace

int64_t enclosed_voids(Context &ctx, const void *cc, int cc_bytes, Dims d, int64_t id, const Box &box) {
  Crop c;
  for (int a = 0; a < 3; ++a) {
    c.lo[a] = box.lo[a];
    c.ext[a] = box.hi[a] - box.lo[a] + 1;
  }
  if (cc_bytes == 1) return voids_of(ctx, static_cast<const uint8_t *>(cc), d, uint8_t(id), c);
  if (cc_bytes == 2) return voids_of(ctx, static_cast<const uint16_t *>(cc), d, uint16_t(id), c);
  return voids_of(ctx, static_cast<const uint32_t *>(cc), d, uint32_t(id), c);
}

VoidFill::VoidFill(Context &ctx, void *cc, int cc_bytes, Dims d, int64_t id, const Box &box, size_t ids)
    : cc_(cc), cc_bytes_(cc_bytes), d_(d), id_(id) {
  Crop c;
  for (int a = 0; a < 3; ++a) {
    c.lo[a] = lo_[a] = box.lo[a];
    c.ext[a] = ext_[a] = box.hi[a] - box.lo[a] + 1;
  }
  if (cc_bytes == 1) voids_ = fill_of(ctx, static_cast<const uint8_t *>(cc), d, uint8_t(id), c, ids, mask_, taken_);
  else if (cc_bytes == 2) voids_ = fill_of(ctx, static_cast<const uint16_t *>(cc), d, uint16_t(id), c, ids, mask_, taken_);
  else voids_ = fill_of(ctx, static_cast<const uint32_t *>(cc), d, uint32_t(id), c, ids, mask_, taken_);
}

void VoidFill::paint(Context &ctx) {
  Crop c;
  for (int a = 0; a < 3; ++a) { c.lo[a] = lo_[a]; c.ext[a] = ext_[a]; }
  const unsigned grid = blocks_for(c.size());
  auto s = ctx.stream();
  if (cc_bytes_ == 1) paint_voids<uint8_t><<<grid, kThreads, 0, s>>>(static_cast<uint8_t *>(cc_), d_, c, uint8_t(id_), mask_.data());
  else if (cc_bytes_ == 2) paint_voids<uint16_t><<<grid, kThreads, 0, s>>>(static_cast<uint16_t *>(cc_), d_, c, uint16_t(id_), mask_.data());
  else paint_voids<uint32_t><<<grid, kThreads, 0, s>>>(static_cast<uint32_t *>(cc_), d_, c, uint32_t(id_), mask_.data());
  BROOK_LAUNCHED();
}

float VoidFill::dbf_refresh(Context &ctx, const float aniso[3], float *dbf, int64_t filled_count, const unsigned *graph) {
  Crop c;
  for (int a = 0; a < 3; ++a) { c.lo[a] = lo_[a]; c.ext[a] = ext_[a]; }
  const Dims box{ext_[0], ext_[1], ext_[2]};
  Buffer<float> box_dbf(ctx, size_t(c.size()));
  Buffer<unsigned> maximum(ctx, 2);
  maximum.clear(ctx);
  if (graph) {  // the graph EDT of the box with the box's graph words (trace.cu:107)
    Buffer<unsigned> words(ctx, size_t(c.size()));
    crop_words<<<blocks_for(c.size()), kThreads, 0, ctx.stream()>>>(d_, c, graph, words.data());
    BROOK_LAUNCHED();
    graph_edt(ctx, mask_.data(), 1, box, words.data(), aniso, filled_count == c.size(), box_dbf.data(), maximum.data());
  } else {
    edt(ctx, mask_.data(), 1, box, aniso, filled_count == c.size(), 3, box_dbf.data(), maximum.data());
  }
  overlay_dbf<<<blocks_for(c.size()), kThreads, 0, ctx.stream()>>>(d_, c, mask_.data(), box_dbf.data(), dbf);
  BROOK_LAUNCHED();
  const unsigned bits = maximum.download(ctx)[1];
  float m;
  std::memcpy(&m, &bits, 4);
  return m;
}

Buffer<unsigned> VoidFill::crop_graph(Context &ctx, const unsigned *graph) const {
  Crop c;
  for (int a = 0; a < 3; ++a) { c.lo[a] = lo_[a]; c.ext[a] = ext_[a]; }
  Buffer<unsigned> words(ctx, size_t(c.size()));
  crop_words<<<blocks_for(c.size()), kThreads, 0, ctx.stream()>>>(d_, c, graph, words.data());
  BROOK_LAUNCHED();
  return words;
}

std::vector<int64_t> first_voxels(Context &ctx, const void *cc, int cc_bytes, long long n, int64_t count) {
  Buffer<unsigned long long> first(ctx, size_t(count + 1));
  first.fill_bytes(ctx, 0xFF);
  const unsigned grid = blocks_for(n);
  auto s = ctx.stream();
  if (cc_bytes == 1) first_of<uint8_t><<<grid, kThreads, 0, s>>>(static_cast<const uint8_t *>(cc), n, first.data());
  else if (cc_bytes == 2) first_of<uint16_t><<<grid, kThreads, 0, s>>>(static_cast<const uint16_t *>(cc), n, first.data());
  else first_of<uint32_t><<<grid, kThreads, 0, s>>>(static_cast<const uint32_t *>(cc), n, first.data());
  BROOK_LAUNCHED();
  std::vector<int64_t> out;
  for (unsigned long long f : first.download(ctx)) out.push_back(f == ~0ULL ? -1 : int64_t(f));
  return out;
}

std::vector<Maximum> soma_maxima(Context &ctx, const void *cc, int cc_bytes, long long n, const float *dbf,
                                 const float *dbf_max, const uint8_t *soma) {
  if (cc_bytes == 1) return maxima_of(ctx, static_cast<const uint8_t *>(cc), n, dbf, dbf_max, soma);
  if (cc_bytes == 2) return maxima_of(ctx, static_cast<const uint16_t *>(cc), n, dbf, dbf_max, soma);
  return maxima_of(ctx, static_cast<const uint32_t *>(cc), n, dbf, dbf_max, soma);
}

}  // namespace brook
