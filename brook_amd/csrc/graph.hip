// Voxel graphs (the `voxel_graph` option) on gfx950: components of the graph's permitted steps, the
// graph-aware EDT on the doubled grid, and the graph gates of the per-voxel neighbour masks.
//
// A graph word holds a bit per 26-neighbour step in cc3d's convention (kGraphBit): the step
// v -> v + N[k] is permitted when bit kGraphBit[k] of graph[v] is set. A uint8 graph of one XY plane
// arrives widened value for value; its components read the corner bits 6..9 moved down by 2
// (graph_components.cu:18-31), everything else reads the widened word as it is
// (graph_components.cu:167-176).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "graph.hpp"

namespace brook {
namespace {

// cc3d's bit of neighbour k (common.cuh:49-56), and the opposite neighbour at compile time.
constexpr int kGraphBit[26] = {1, 0, 3, 2, 5, 4, 9, 7, 8, 6, 17, 13, 16, 12, 15, 11, 14, 10, 25, 21, 23, 19, 24, 20, 22, 18};
constexpr int kOppositeOf[26] = {1, 0, 3, 2, 5, 4, 9, 8, 7, 6, 13, 12, 11, 10, 17, 16, 15, 14, 25, 24, 23, 22, 21, 20, 19, 18};

// Bit k of the result: graph word w permits step k (compact: a uint8 plane graph's corner bits).
__device__ __forceinline__ unsigned step_mask(unsigned w, bool compact) {
  unsigned m = 0;
  for_each_nb([&](auto K) {
    constexpr int k = decltype(K)::value;
    const int bit = compact && kGraphBit[k] >= 6 ? kGraphBit[k] - 2 : kGraphBit[k];
    m |= ((w >> bit) & 1u) << k;
  });
  return m;
}

unsigned blocks_of(long long n) { return unsigned((n + 255) / 256); }

// ---- components ------------------------------------------------------------------------------
__device__ __forceinline__ int find_root(const int *parent, int v) {
  int p = ld_i(parent + v);
  while (p != v) {
    v = p;
    p = ld_i(parent + v);
  }
  return v;
}

// Union by index: the smaller root stays a root, so each component's root is its lowest voxel.
__device__ __forceinline__ void unite(int *parent, int a, int b) {
  a = find_root(parent, a);
  b = find_root(parent, b);
  while (a != b) {
    if (a > b) { const int t = a; a = b; b = t; }
    const int got = g_cas(parent + b, b, a);
    if (got == b) return;
    b = find_root(parent, got);
    a = find_root(parent, a);
  }
}

__global__ void __launch_bounds__(256) g_init(int *parent, int n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) parent[i] = int(i);
}

// Each voxel unites with its lower-index neighbours its own word permits a step to.
__global__ void __launch_bounds__(256) g_union(const unsigned *graph, bool compact, int *parent, Dims d, int n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int v = int(i);
  unsigned m = step_mask(graph[v], compact);
  if (!m) return;
  int x, y, z;
  unravel(v, d, x, y, z);
  while (m) {
    const int k = __ffs(m) - 1;
    m &= m - 1;
    const int nx = x + nb_dx(k), ny = y + nb_dy(k), nz = z + nb_dz(k);
    if (!inside(nx, ny, nz, d)) continue;
    const int u = flat(nx, ny, nz, d);
    if (u < v) unite(parent, v, u);
  }
}

// Roots, and the roots of components that hold a foreground voxel; label changes (input_uniform).
template <class T>
__global__ void __launch_bounds__(256) g_flags(const T *labels, int *parent, int n, int *is_root, int *has_fg,
                                               int *nonuniform) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int r = find_root(parent, int(i));
  is_root[i] = r == int(i);
  const T l = labels[i];
  if (l != T(0)) st_i(has_fg + r, 1);
  if (l != labels[0]) st_i(nonuniform, 1);
}

__global__ void __launch_bounds__(256) g_flatten(int *parent, int n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) parent[i] = find_root(parent, int(i));
}

// ids (1-based compact) of the foreground voxels; keys[id] = 1-based rank of the root among all
// roots; roots[id - 1] = lowest foreground voxel; last[id] = start of the id's last run.
template <class T>
__global__ void __launch_bounds__(256) g_color(const T *labels, const int *parent, const int *is_root, const int *rank,
                                               const int *has_fg, const int *compact, int n, unsigned *ids,
                                               long long *keys) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  ids[i] = labels[i] != T(0) ? unsigned(compact[parent[i]] + 1) : 0u;
  if (is_root[i] && has_fg[i]) keys[compact[i] + 1] = (long long)rank[i] + 1;
}

__global__ void __launch_bounds__(256) g_positions(const unsigned *ids, int n, int *roots, int *last) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const unsigned id = ids[i];
  if (!id) return;
  atomicMin(roots + (id - 1), int(i));
  if (i == 0 || ids[i - 1] != id) atomicMax(last + id, int(i));
}

template <class T>
__global__ void __launch_bounds__(256) g_mapping(const T *labels, const int *last, int count, long long *mapping) {
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  if (id > count) return;
  mapping[id] = id ? (long long)labels[last[id]] : 0;
}

// A uint8 plane graph's words, value for value (graph_components.cu:167-170).
__global__ void __launch_bounds__(256) g_widen(const uint8_t *in, long long n, unsigned *out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = in[i];
}

template <class Out>
__global__ void __launch_bounds__(256) g_pack(const unsigned *ids, int n, Out *out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = Out(ids[i]);
}

template <class T>
GraphComponents run_components(Context &ctx, const LabelVolume &in, const unsigned *graph, bool compact_2d) {
  const int n = int(in.size());
  const Dims d = in.dims;
  const T *lab = static_cast<const T *>(in.data);
  auto s = ctx.stream();
  Buffer<int> parent(ctx, size_t(n)), is_root(ctx, size_t(n)), rank(ctx, size_t(n)), has_fg(ctx, size_t(n)),
      compact(ctx, size_t(n)), flag(ctx, 1);
  has_fg.clear(ctx);
  flag.clear(ctx);
  {
    KernelTimer kt(ctx, "graph_components");
    g_init<<<blocks_of(n), 256, 0, s>>>(parent.data(), n);
    g_union<<<blocks_of(n), 256, 0, s>>>(graph, compact_2d, parent.data(), d, n);
    g_flags<T><<<blocks_of(n), 256, 0, s>>>(lab, parent.data(), n, is_root.data(), has_fg.data(), flag.data());
    g_flatten<<<blocks_of(n), 256, 0, s>>>(parent.data(), n);
    BROOK_LAUNCHED();
  }
  exclusive_sum(ctx, is_root.data(), rank.data(), n);
  exclusive_sum(ctx, has_fg.data(), compact.data(), n);
  int tail[2] = {0, 0};
  BROOK_HIP(hipMemcpyAsync(&tail[0], compact.data() + n - 1, 4, hipMemcpyDeviceToHost, s));
  BROOK_HIP(hipMemcpyAsync(&tail[1], has_fg.data() + n - 1, 4, hipMemcpyDeviceToHost, s));
  int nonuniform = 0;
  BROOK_HIP(hipMemcpyAsync(&nonuniform, flag.data(), 4, hipMemcpyDeviceToHost, s));
  ctx.synchronize();
  const int count = tail[0] + tail[1];
  GraphComponents out;
  out.input_uniform = nonuniform == 0;
  Buffer<unsigned> ids(ctx, size_t(n));
  Buffer<long long> keys(ctx, size_t(count) + 1), mapping(ctx, size_t(count) + 1);
  Buffer<int> roots(ctx, size_t(std::max(count, 1))), last(ctx, size_t(count) + 1);
  keys.clear(ctx);
  roots.fill_bytes(ctx, 0x7F);
  last.fill_bytes(ctx, 0xFF);
  g_color<T><<<blocks_of(n), 256, 0, s>>>(lab, parent.data(), is_root.data(), rank.data(), has_fg.data(),
                                         compact.data(), n, ids.data(), keys.data());
  g_positions<<<blocks_of(n), 256, 0, s>>>(ids.data(), n, roots.data(), last.data());
  g_mapping<T><<<blocks_of(count + 1), 256, 0, s>>>(lab, last.data(), count, mapping.data());
  BROOK_LAUNCHED();
  Components &c = out.cc;
  c.count = count;
  c.cc_bytes = id_bytes(count);
  c.labels = Buffer<uint8_t>(ctx, size_t(n) * size_t(c.cc_bytes));
  if (c.cc_bytes == 1) g_pack<uint8_t><<<blocks_of(n), 256, 0, s>>>(ids.data(), n, c.labels.data());
  else if (c.cc_bytes == 2) g_pack<uint16_t><<<blocks_of(n), 256, 0, s>>>(ids.data(), n, reinterpret_cast<uint16_t *>(c.labels.data()));
  else g_pack<uint32_t><<<blocks_of(n), 256, 0, s>>>(ids.data(), n, reinterpret_cast<uint32_t *>(c.labels.data()));
  BROOK_LAUNCHED();
  const std::vector<long long> k = keys.download(ctx), m = mapping.download(ctx);
  const std::vector<int> r = roots.download(ctx, size_t(count));
  out.keys.assign(k.begin(), k.end());
  c.mapping.assign(m.begin(), m.end());
  c.roots.assign(r.begin(), r.end());
  if (in.is_signed && sizeof(T) < 8)
    for (auto &v : c.mapping) {
      const int shift = 64 - 8 * int(sizeof(T));
      v = int64_t(uint64_t(v) << shift) >> shift;  // sign-extend the label's width
    }
  return out;
}

// ---- EDT on the doubled grid ---------------------------------------------------------------------
// The binary mask of graph_edt.cu:14-33: position (X, Y, Z) of the doubled grid belongs to voxel
// (X/2, Y/2, Z/2); a foreground voxel's +x, +y, +z half-steps are foreground when its word permits
// the step (bits 0, 2, 4), its other positions always; with a black border the last position on
// each axis is background.
template <class CC>
__global__ void __launch_bounds__(256) g_double(const CC *cc, const unsigned *graph, Dims d, bool black_border,
                                                uint8_t *mask) {
  const long long sx2 = 2LL * d.sx, sy2 = 2LL * d.sy, sz2 = 2LL * d.sz;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= sx2 * sy2 * sz2) return;
  const long long X = i % sx2, Y = (i / sx2) % sy2, Z = i / (sx2 * sy2);
  uint8_t m = 0;
  if (!(black_border && (X == sx2 - 1 || Y == sy2 - 1 || Z == sz2 - 1))) {
    const long long src = (X >> 1) + d.sx * ((Y >> 1) + (long long)d.sy * (Z >> 1));
    if (cc[src] != CC(0)) {
      const int sub = int((X & 1) | ((Y & 1) << 1) | ((Z & 1) << 2));
      const unsigned w = graph[src];
      m = sub == 1 ? (w & 1u) != 0 : sub == 2 ? (w & 4u) != 0 : sub == 4 ? (w & 16u) != 0 : 1;
    }
  }
  mask[i] = m;
}

// The voxel centres of the doubled EDT; per-id maxima (float bits) when asked.
template <class CC>
__global__ void __launch_bounds__(256) g_sample(const CC *cc, const float *doubled, Dims d, float *out,
                                                unsigned *label_max) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long n = (long long)d.sx * d.sy * d.sz;
  if (i >= n) return;
  const CC l = cc[i];
  if (l == CC(0)) { out[i] = 0.0f; return; }
  const long long x = i % d.sx, y = (i / d.sx) % d.sy, z = i / ((long long)d.sx * d.sy);
  const float v = doubled[2 * x + 2LL * d.sx * (2 * y + 2LL * d.sy * 2 * z)];
  out[i] = v;
  if (label_max && v < kInf) atomicMax(label_max + (unsigned long long)l, __float_as_uint(v));
}

template <class CC>
void run_graph_edt(Context &ctx, const CC *cc, Dims d, const unsigned *graph, const float aniso[3], bool bb, float *out,
                   unsigned *label_max) {
  const Dims d2{2 * d.sx, 2 * d.sy, 2 * d.sz};
  const long long n2 = (long long)d2.sx * d2.sy * d2.sz;
  Buffer<uint8_t> mask(ctx, size_t(n2));
  Buffer<float> doubled(ctx, size_t(n2));
  g_double<CC><<<blocks_of(n2), 256, 0, ctx.stream()>>>(cc, graph, d, bb, mask.data());
  BROOK_LAUNCHED();
  const float half[3] = {aniso[0] * 0.5f, aniso[1] * 0.5f, aniso[2] * 0.5f};
  edt(ctx, mask.data(), 1, d2, half, bb, 3, doubled.data());
  g_sample<CC><<<blocks_of((long long)d.sx * d.sy * d.sz), 256, 0, ctx.stream()>>>(cc, doubled.data(), d, out, label_max);
  BROOK_LAUNCHED();
}

// ---- gates of the neighbour masks -------------------------------------------------------------
// meta's same-label mask keeps the steps the voxel's own word permits (pushes: floods, railroads,
// joins, invalidation offers); in[v] bit k: neighbour u = v + N[k] is of v's id and u's word permits
// the step u -> v (pulls: backtraces, band-boundary pulls, parents).
__global__ void __launch_bounds__(256) g_gates(VoxelRec *rec, const unsigned *graph, Dims d, long long n,
                                               unsigned *in) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const unsigned meta = rec[i].meta, same = meta & kNbrBits;
  unsigned pull = 0;
  for_each_nb([&](auto K) {
    constexpr int k = decltype(K)::value;
    if ((same >> k) & 1u) pull |= ((graph[i + nb_offset<k>(d)] >> kGraphBit[kOppositeOf[k]]) & 1u) << k;
  });
  in[i] = pull;
  rec[i].meta = (meta & ~kNbrBits) | (same & step_mask(graph[i], false));
}

}  // namespace

GraphComponents graph_components(Context &ctx, const LabelVolume &in, const unsigned *graph, bool compact_2d) {
  if (in.size() >= (int64_t(1) << 31)) throw Unsupported("volume exceeds the int32 voxel index space");
  if (in.size() == 0) {
    GraphComponents g;
    g.cc.mapping = {0};
    g.keys = {0};
    return g;
  }
  switch (in.bytes) {
    case 1: return run_components<uint8_t>(ctx, in, graph, compact_2d);
    case 2: return run_components<uint16_t>(ctx, in, graph, compact_2d);
    case 4: return run_components<uint32_t>(ctx, in, graph, compact_2d);
    case 8: return run_components<unsigned long long>(ctx, in, graph, compact_2d);
    default: throw std::invalid_argument("voxel graph labels must be integers");
  }
}

void graph_edt(Context &ctx, const void *cc, int cc_bytes, Dims d, const unsigned *graph, const float aniso[3],
               bool black_border, float *out, unsigned *label_max) {
  for (int a = 0; a < 3; ++a)
    if (!(aniso[a] > 0) || !std::isfinite(aniso[a])) throw std::invalid_argument("anisotropy must be finite and positive");
  const long long n = (long long)d.sx * d.sy * d.sz;
  if (n == 0) return;
  if (8 * n >= (1LL << 31)) throw Unsupported("the voxel-graph EDT's doubled grid exceeds 2^31 positions (volume above 2^28 voxels)");
  if (cc_bytes == 1) run_graph_edt<uint8_t>(ctx, static_cast<const uint8_t *>(cc), d, graph, aniso, black_border, out, label_max);
  else if (cc_bytes == 2) run_graph_edt<uint16_t>(ctx, static_cast<const uint16_t *>(cc), d, graph, aniso, black_border, out, label_max);
  else if (cc_bytes == 4) run_graph_edt<uint32_t>(ctx, static_cast<const uint32_t *>(cc), d, graph, aniso, black_border, out, label_max);
  else if (!label_max) run_graph_edt<unsigned long long>(ctx, static_cast<const unsigned long long *>(cc), d, graph, aniso, black_border, out, nullptr);
  else throw std::invalid_argument("per-id maxima need ids of at most 32 bits");
}

void widen_words(Context &ctx, const uint8_t *in, long long n, unsigned *out) {
  if (n <= 0) return;
  g_widen<<<blocks_of(n), 256, 0, ctx.stream()>>>(in, n, out);
  BROOK_LAUNCHED();
}

void graph_gates(Context &ctx, VoxelRec *rec, const unsigned *graph, Dims d, unsigned *in_masks) {
  const long long n = (long long)d.sx * d.sy * d.sz;
  if (n == 0) return;
  g_gates<<<blocks_of(n), 256, 0, ctx.stream()>>>(rec, graph, d, n, in_masks);
  BROOK_LAUNCHED();
}

}  // namespace brook
