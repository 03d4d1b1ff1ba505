// Soma preparation kernels: the enclosed-void check (a background flood of a label's bounding box
// through face neighbours from the box's faces, one grid-wide round per launch) and the DBF maxima
// of soma labels (a flagged selection in flat order).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstring>

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

// Box voxels on a face that are not the label: reached, first front.
template <class CC>
__global__ void void_seed(const CC *cc, Dims d, Crop c, CC id, unsigned *reached, int *front, int *count) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= c.size()) return;
  const int x = int(i % c.ext[0]), y = int((i / c.ext[0]) % c.ext[1]), z = int(i / ((long long)c.ext[0] * c.ext[1]));
  const bool face = x == 0 || y == 0 || z == 0 || x == c.ext[0] - 1 || y == c.ext[1] - 1 || z == c.ext[2] - 1;
  if (!face || is_label(cc, d, c, x, y, z, id)) return;
  g_st(reached + i, 1u);
  g_st(front + wave_slot(count), int(i));
}

// One launch of the flood: every front entry explores depth-first from itself (a small stack per
// thread, a budget of pops), claiming box voxels that are not the label with an exchange; what a
// full stack or a spent budget cannot hold goes to the next front. Each voxel is claimed once, so
// a front never exceeds the box.
constexpr int kDfsStack = 16, kDfsBudget = 512;
template <class CC>
__global__ void void_round(const CC *cc, Dims d, Crop c, CC id, unsigned *reached, const int *front, int n, int *next,
                           int *count) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= n) return;
  const long long sx = c.ext[0], sxy = (long long)c.ext[0] * c.ext[1];
  int stack[kDfsStack];
  int sp = 0, budget = kDfsBudget;
  stack[sp++] = front[j];
  while (sp > 0) {
    const long long i = stack[--sp];
    if (--budget < 0) {  // hand the rest back
      g_st(next + atomicAdd(count, 1), int(i));
      continue;
    }
    const int x = int(i % c.ext[0]), y = int((i / c.ext[0]) % c.ext[1]), z = int(i / sxy);
    for (int k = 0; k < 6; ++k) {
      const int nx = x + (k == 0 ? -1 : k == 1 ? 1 : 0), ny = y + (k == 2 ? -1 : k == 3 ? 1 : 0),
                nz = z + (k == 4 ? -1 : k == 5 ? 1 : 0);
      if (nx < 0 || ny < 0 || nz < 0 || nx >= c.ext[0] || ny >= c.ext[1] || nz >= c.ext[2]) continue;
      if (is_label(cc, d, c, nx, ny, nz, id)) continue;
      const long long v = nx + sx * ny + sxy * nz;
      if (g_xchg(reached + v, 1u) != 0u) continue;
      if (sp < kDfsStack) stack[sp++] = int(v);
      else g_st(next + atomicAdd(count, 1), int(v));
    }
  }
}

// reached[i] = 1 for the box voxels that are not the label and reach a face of the box through
// such voxels (face neighbours).
template <class CC> void reach_faces(Context &ctx, const CC *cc, Dims d, CC id, const Crop &c, Buffer<unsigned> &reached) {
  auto s = ctx.stream();
  const long long n = c.size();
  if (n > INT32_MAX) throw std::invalid_argument("label bounding box exceeds int32 capacity");
  reached = Buffer<unsigned>(ctx, size_t(n));
  Buffer<int> a(ctx, size_t(n)), b(ctx, size_t(n)), counts(ctx, 2);
  reached.clear(ctx);
  counts.clear(ctx);
  void_seed<CC><<<blocks_for(n), kThreads, 0, s>>>(cc, d, c, id, reached.data(), a.data(), counts.data());
  BROOK_LAUNCHED();
  int front = 0;
  BROOK_HIP(hipMemcpyAsync(&front, counts.data(), 4, hipMemcpyDeviceToHost, s));
  ctx.synchronize();
  int slot = 0;  // counts[slot] holds the current front's length
  while (front > 0) {
    const int other = 1 - slot;
    BROOK_HIP(hipMemsetAsync(counts.data() + other, 0, 4, s));
    void_round<CC><<<blocks_for(front), kThreads, 0, s>>>(cc, d, c, id, reached.data(), a.data(), front, b.data(),
                                                          counts.data() + other);
    BROOK_LAUNCHED();
    BROOK_HIP(hipMemcpyAsync(&front, counts.data() + other, 4, hipMemcpyDeviceToHost, s));
    ctx.synchronize();
    std::swap(a, b);
    slot = other;
  }
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

float VoidFill::dbf_refresh(Context &ctx, const float aniso[3], float *dbf, int64_t filled_count) {
  Crop c;
  for (int a = 0; a < 3; ++a) { c.lo[a] = lo_[a]; c.ext[a] = ext_[a]; }
  const Dims box{ext_[0], ext_[1], ext_[2]};
  Buffer<float> box_dbf(ctx, size_t(c.size()));
  Buffer<unsigned> maximum(ctx, 2);
  maximum.clear(ctx);
  edt(ctx, mask_.data(), 1, box, aniso, filled_count == c.size(), 3, box_dbf.data(), maximum.data());
  overlay_dbf<<<blocks_for(c.size()), kThreads, 0, ctx.stream()>>>(d_, c, mask_.data(), box_dbf.data(), dbf);
  BROOK_LAUNCHED();
  const unsigned bits = maximum.download(ctx)[1];
  float m;
  std::memcpy(&m, &bits, 4);
  return m;
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
