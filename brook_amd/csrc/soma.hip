// Soma preparation kernels: the enclosed-void check (a background flood of a label's bounding box
// through face neighbours from the box's faces, one grid-wide round per launch) and the DBF maxima
// of soma labels (a flagged selection in flat order).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

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

// One round: the face neighbours of the front that are not the label and not yet reached.
template <class CC>
__global__ void void_round(const CC *cc, Dims d, Crop c, CC id, unsigned *reached, const int *front, int n, int *next,
                           int *count) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= n) return;
  const long long i = front[j];
  const int x = int(i % c.ext[0]), y = int((i / c.ext[0]) % c.ext[1]), z = int(i / ((long long)c.ext[0] * c.ext[1]));
  const long long sx = c.ext[0], sxy = (long long)c.ext[0] * c.ext[1];
  const int step[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  for (int k = 0; k < 6; ++k) {
    const int nx = x + step[k][0], ny = y + step[k][1], nz = z + step[k][2];
    if (nx < 0 || ny < 0 || nz < 0 || nx >= c.ext[0] || ny >= c.ext[1] || nz >= c.ext[2]) continue;
    if (is_label(cc, d, c, nx, ny, nz, id)) continue;
    const long long v = nx + sx * ny + sxy * nz;
    if (g_xchg(reached + v, 1u) == 0u) g_st(next + wave_slot(count), int(v));
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
  auto s = ctx.stream();
  const long long n = c.size();
  if (n > INT32_MAX) throw std::invalid_argument("soma bounding box exceeds int32 capacity");
  Buffer<unsigned> reached(ctx, size_t(n));
  Buffer<int> a(ctx, size_t(n)), b(ctx, size_t(n)), counts(ctx, 2);
  Buffer<unsigned long long> voids(ctx, 1);
  reached.clear(ctx);
  counts.clear(ctx);
  voids.clear(ctx);
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
  void_tally<CC><<<blocks_for(n), kThreads, 0, s>>>(cc, d, c, id, reached.data(), voids.data());
  BROOK_LAUNCHED();
  return int64_t(voids.download(ctx)[0]);
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

std::vector<Maximum> soma_maxima(Context &ctx, const void *cc, int cc_bytes, long long n, const float *dbf,
                                 const float *dbf_max, const uint8_t *soma) {
  if (cc_bytes == 1) return maxima_of(ctx, static_cast<const uint8_t *>(cc), n, dbf, dbf_max, soma);
  if (cc_bytes == 2) return maxima_of(ctx, static_cast<const uint16_t *>(cc), n, dbf, dbf_max, soma);
  return maxima_of(ctx, static_cast<const uint32_t *>(cc), n, dbf, dbf_max, soma);
}

}  // namespace brook
