// Canonical postprocess of packed skeletons on the device (post.cu:1-651, post_kernels.cuh:1-407:
// Kimimaro 5.8.1's remove_dust, remove_loops, join_close_components and remove_ticks with the
// reference's tie rules).
//
// Whole-result passes are wide: union-find components over every edge, per-edge lengths, keep
// masks and prefix sums, radix sorts of edge keys and of join candidates, a 27-cell grid search
// for join candidates. The rules that are sequential by nature (one component's cycle loop, one
// tree's tick loop, one label's greedy fusion) run one lane per component or label over scratch
// slices sized from the component: rows, neighbour lists (doubly linked in insertion order, with
// a free list), search stack and path for the loops; a CSR neighbour table, chain records, end
// slots and a binary heap keyed (length, lower end, higher end) for the ticks.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "common.hpp"
#include "post.hpp"

#pragma clang fp contract(off)

namespace brook {
namespace {

constexpr int kThreads = 256;
constexpr unsigned long long kNone = ~0ULL;
unsigned grid(long long n) { return unsigned(std::max<long long>(1, (n + kThreads - 1) / kThreads)); }

template <class Op> void with_scratch(Context &ctx, Buffer<unsigned char> &scratch, Op &&op) {
  size_t bytes = 0;
  BROOK_HIP(op(nullptr, bytes));
  if (bytes > scratch.size()) scratch = Buffer<unsigned char>(ctx, std::max<size_t>(bytes, 1));
  BROOK_HIP(op(scratch.data(), bytes));
}

template <class T> T read_one(Context &ctx, const T *p) {
  T v{};
  BROOK_HIP(hipMemcpyAsync(&v, p, sizeof(T), hipMemcpyDeviceToHost, ctx.stream()));
  ctx.synchronize();
  return v;
}

// ---------------------------------------------------------------------------------------------
// graph: owners, global edges, degrees, components (root = least vertex index) (post.cu:45-67)

// The last group whose offset is <= i (empty groups share offsets with the next one).
__device__ __forceinline__ int owner_of(const long long *off, int groups, long long i) {
  int lo = 0, hi = groups;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int uf_root(int *up, int v) {
  int cur = ld_i(up + v);
  if (cur != v) {
    int prev = v, next;
    while (cur > (next = ld_i(up + cur))) {
      st_i(up + prev, next);  // path halving: a pointer only ever moves closer to its root
      prev = cur;
      cur = next;
    }
  }
  return cur;
}

__device__ __forceinline__ void uf_join(int *up, int a, int b) {
  int ra = uf_root(up, a), rb = uf_root(up, b);
  while (ra != rb) {
    const int lo = min(ra, rb), hi = max(ra, rb);
    const int seen = atomicCAS(up + hi, hi, lo);
    if (seen == hi) return;
    if (hi == ra) ra = uf_root(up, seen);
    else rb = uf_root(up, seen);
  }
}

__global__ void graph_vertices(const long long *voff, int k, int n, int *vlabel, int *root, int *degree) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  vlabel[i] = owner_of(voff, k, i);
  root[i] = i;
  degree[i] = 0;
}

__global__ void graph_edges(const unsigned *local, int m, const long long *voff, const long long *eoff, int k, int *elabel,
                            int2 *e, int *root, int *degree, int *bad) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= m) return;
  const int l = owner_of(eoff, k, i);
  elabel[i] = l;
  const long long size = voff[l + 1] - voff[l];
  const unsigned a = local[2LL * i], b = local[2LL * i + 1];
  if ((long long)a >= size || (long long)b >= size) {
    atomicOr(bad, 1);
    e[i] = make_int2(0, 0);
    return;
  }
  const int ga = int(voff[l] + a), gb = int(voff[l] + b);
  e[i] = make_int2(ga, gb);
  atomicAdd(degree + ga, 1);
  atomicAdd(degree + gb, 1);
  uf_join(root, ga, gb);
}

__global__ void graph_flatten(int *root, int n) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  int r = ld_i(root + i), next;
  while (r != (next = ld_i(root + r))) r = next;  // read-only walk: a racing writer only stores roots
  st_i(root + i, r);
}

struct Graph {
  int n = 0, m = 0, k = 0;
  Buffer<long long> voff, eoff;
  Buffer<int> vlabel, elabel, root, degree;
  Buffer<int2> e;
};

Graph build_graph(Context &ctx, const DevicePacked &p) {
  const size_t k = p.labels.size();
  const bool ok = p.voff.size() == k + 1 && p.eoff.size() == k + 1 && p.voff.front() == 0 && p.eoff.front() == 0 &&
                  std::is_sorted(p.voff.begin(), p.voff.end()) && std::is_sorted(p.eoff.begin(), p.eoff.end()) &&
                  p.vertices.size() >= size_t(3 * p.n()) && p.radii.size() >= size_t(p.n()) &&
                  p.edges.size() >= size_t(2 * p.m());
  if (!ok) throw std::invalid_argument("invalid packed graph");
  if (p.n() > INT32_MAX || p.m() > INT32_MAX || k > size_t(INT32_MAX))
    throw std::invalid_argument("packed graph exceeds int32 work capacity");
  Graph g;
  g.n = int(p.n());
  g.m = int(p.m());
  g.k = int(k);
  g.voff = Buffer<long long>(ctx, k + 1);
  g.eoff = Buffer<long long>(ctx, k + 1);
  g.voff.upload(ctx, reinterpret_cast<const long long *>(p.voff.data()), k + 1);
  g.eoff.upload(ctx, reinterpret_cast<const long long *>(p.eoff.data()), k + 1);
  g.vlabel = Buffer<int>(ctx, std::max(g.n, 1));
  g.root = Buffer<int>(ctx, std::max(g.n, 1));
  g.degree = Buffer<int>(ctx, std::max(g.n, 1));
  g.elabel = Buffer<int>(ctx, std::max(g.m, 1));
  g.e = Buffer<int2>(ctx, std::max(g.m, 1));
  Buffer<int> bad(ctx, 1);
  bad.clear(ctx);
  auto s = ctx.stream();
  if (g.n) graph_vertices<<<grid(g.n), kThreads, 0, s>>>(g.voff.data(), g.k, g.n, g.vlabel.data(), g.root.data(), g.degree.data());
  if (g.m)
    graph_edges<<<grid(g.m), kThreads, 0, s>>>(p.edges.data(), g.m, g.voff.data(), g.eoff.data(), g.k, g.elabel.data(),
                                               g.e.data(), g.root.data(), g.degree.data(), bad.data());
  if (g.n) graph_flatten<<<grid(g.n), kThreads, 0, s>>>(g.root.data(), g.n);
  BROOK_LAUNCHED();
  if (read_one(ctx, bad.data())) throw std::invalid_argument("packed edge is out of bounds");
  return g;
}

// ---------------------------------------------------------------------------------------------
// selections (post.cu:70-117, 388-404)

__global__ void edge_both_kept(const int2 *e, int m, const int *keep, int *out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < m) out[i] = keep[e[i].x] && keep[e[i].y];
}

__global__ void gather_offsets(const int *pos, const long long *off, int k, long long *out) {
  const int g = blockIdx.x * kThreads + threadIdx.x;
  if (g <= k) out[g] = pos[off[g]];
}

__global__ void scatter_vertices(const float *v, const float *r, const int *keep, const int *pos, int n, float *vout,
                                 float *rout) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n || !keep[i]) return;
  const int j = pos[i];
  vout[3LL * j] = v[3LL * i];
  vout[3LL * j + 1] = v[3LL * i + 1];
  vout[3LL * j + 2] = v[3LL * i + 2];
  rout[j] = r[i];
}

__global__ void scatter_edges(const int2 *e, const int *elabel, const int *ekeep, const int *epos, const int *vpos,
                              const long long *new_voff, int m, unsigned *out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= m || !ekeep[i]) return;
  const long long base = new_voff[elabel[i]];
  const int j = epos[i];
  out[2LL * j] = unsigned(vpos[e[i].x] - base);
  out[2LL * j + 1] = unsigned(vpos[e[i].y] - base);
}

__global__ void copy_edge_rows(const unsigned *in, const int *keep, const int *pos, int m, unsigned *out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= m || !keep[i]) return;
  out[2LL * pos[i]] = in[2LL * i];
  out[2LL * pos[i] + 1] = in[2LL * i + 1];
}

// Exclusive positions of `flags` (n + 1 entries, the last the total) and the per-group offsets
// they give at the group boundaries `off`.
std::vector<int64_t> positions(Context &ctx, const Buffer<int> &flags, int n, const Buffer<long long> &off, int k,
                               Buffer<int> &pos, Buffer<long long> &new_off) {
  Buffer<int> padded(ctx, size_t(n) + 1);
  if (n) BROOK_HIP(hipMemcpyAsync(padded.data(), flags.data(), size_t(n) * 4, hipMemcpyDeviceToDevice, ctx.stream()));
  BROOK_HIP(hipMemsetAsync(padded.data() + n, 0, 4, ctx.stream()));
  pos = Buffer<int>(ctx, size_t(n) + 1);
  exclusive_sum(ctx, padded.data(), pos.data(), n + 1);
  new_off = Buffer<long long>(ctx, size_t(k) + 1);
  gather_offsets<<<grid(k + 1), kThreads, 0, ctx.stream()>>>(pos.data(), off.data(), k, new_off.data());
  BROOK_LAUNCHED();
  std::vector<long long> h = new_off.download(ctx);
  return std::vector<int64_t>(h.begin(), h.end());
}

// The kept vertices in order and the edges with both ends kept, in order (post.cu:86-109).
DevicePacked select_vertices(Context &ctx, const DevicePacked &p, const Buffer<int> &keep, const Graph &g) {
  DevicePacked out;
  out.labels = p.labels;
  out.anisotropy = p.anisotropy;
  Buffer<int> vpos, epos, ekeep(ctx, std::max(g.m, 1));
  Buffer<long long> new_voff, new_eoff;
  out.voff = positions(ctx, keep, g.n, g.voff, g.k, vpos, new_voff);
  if (g.m) edge_both_kept<<<grid(g.m), kThreads, 0, ctx.stream()>>>(g.e.data(), g.m, keep.data(), ekeep.data());
  BROOK_LAUNCHED();
  out.eoff = positions(ctx, ekeep, g.m, g.eoff, g.k, epos, new_eoff);
  out.vertices = Buffer<float>(ctx, size_t(std::max<int64_t>(3 * out.n(), 1)));
  out.radii = Buffer<float>(ctx, size_t(std::max<int64_t>(out.n(), 1)));
  out.edges = Buffer<uint32_t>(ctx, size_t(std::max<int64_t>(2 * out.m(), 1)));
  if (g.n)
    scatter_vertices<<<grid(g.n), kThreads, 0, ctx.stream()>>>(p.vertices.data(), p.radii.data(), keep.data(), vpos.data(),
                                                               g.n, out.vertices.data(), out.radii.data());
  if (g.m)
    scatter_edges<<<grid(g.m), kThreads, 0, ctx.stream()>>>(g.e.data(), g.elabel.data(), ekeep.data(), epos.data(),
                                                            vpos.data(), new_voff.data(), g.m, out.edges.data());
  BROOK_LAUNCHED();
  return out;
}

// The edges with keep[i] in order; vertices unchanged (post.cu:388-404).
DevicePacked select_edges(Context &ctx, DevicePacked p, const Buffer<int> &keep, const Graph &g) {
  Buffer<int> epos;
  Buffer<long long> new_eoff;
  p.eoff = positions(ctx, keep, g.m, g.eoff, g.k, epos, new_eoff);
  Buffer<uint32_t> edges(ctx, size_t(std::max<int64_t>(2 * p.m(), 1)));
  if (g.m) copy_edge_rows<<<grid(g.m), kThreads, 0, ctx.stream()>>>(p.edges.data(), keep.data(), epos.data(), g.m, edges.data());
  BROOK_LAUNCHED();
  p.edges = std::move(edges);
  return p;
}

// ---------------------------------------------------------------------------------------------
// sorted unique edge keys -> local edges (post.cu:273-292)

__global__ void unique_heads(const unsigned long long *keys, int n, int *flags) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) flags[i] = keys[i] != kNone && (i == 0 || keys[i] != keys[i - 1]);
}

__global__ void compact_keys(const unsigned long long *keys, const int *flags, const int *pos, int n,
                             unsigned long long *out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n && flags[i]) out[pos[i]] = keys[i];
}

// Per group: the first unique key whose lower end is at or past the group's first vertex.
__global__ void key_offsets(const unsigned long long *keys, int n, const long long *voff, int k, long long *eoff) {
  const int g = blockIdx.x * kThreads + threadIdx.x;
  if (g > k) return;
  const unsigned long long target = (unsigned long long)voff[g] << 32;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  eoff[g] = lo;
}

__global__ void local_pairs(const unsigned long long *keys, int n, const int *vlabel, const long long *voff, unsigned *out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const unsigned a = unsigned(keys[i] >> 32), b = unsigned(keys[i]);
  const long long base = voff[vlabel[a]];
  out[2LL * i] = unsigned(a - base);
  out[2LL * i + 1] = unsigned(b - base);
}

DevicePacked edges_from_keys(Context &ctx, DevicePacked p, const Graph &g, Buffer<unsigned long long> &keys, int count) {
  Buffer<unsigned char> scratch;
  Buffer<unsigned long long> sorted(ctx, std::max(count, 1));
  if (count)
    with_scratch(ctx, scratch, [&](void *t, size_t &b) {
      return hipcub::DeviceRadixSort::SortKeys(t, b, keys.data(), sorted.data(), count, 0, 64, ctx.stream());
    });
  Buffer<int> flags(ctx, std::max(count, 1)), pos(ctx, std::max(count, 1));
  int total = 0;
  Buffer<unsigned long long> unique(ctx, std::max(count, 1));
  if (count) {
    unique_heads<<<grid(count), kThreads, 0, ctx.stream()>>>(sorted.data(), count, flags.data());
    exclusive_sum(ctx, flags.data(), pos.data(), count);
    compact_keys<<<grid(count), kThreads, 0, ctx.stream()>>>(sorted.data(), flags.data(), pos.data(), count, unique.data());
    BROOK_LAUNCHED();
    total = read_one(ctx, pos.data() + count - 1) + read_one(ctx, flags.data() + count - 1);
  }
  Buffer<long long> eoff(ctx, size_t(g.k) + 1);
  key_offsets<<<grid(g.k + 1), kThreads, 0, ctx.stream()>>>(unique.data(), total, g.voff.data(), g.k, eoff.data());
  Buffer<uint32_t> edges(ctx, size_t(std::max(2 * total, 1)));
  if (total) local_pairs<<<grid(total), kThreads, 0, ctx.stream()>>>(unique.data(), total, g.vlabel.data(), g.voff.data(), edges.data());
  BROOK_LAUNCHED();
  std::vector<long long> h = eoff.download(ctx);
  p.eoff.assign(h.begin(), h.end());
  p.edges = std::move(edges);
  return p;
}

// ---------------------------------------------------------------------------------------------
// prune_unused (post.cu:433-442)

__global__ void nonzero(const int *degree, int n, int *keep, int *zero) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  keep[i] = degree[i] != 0;
  if (!degree[i]) atomicOr(zero, 1);
}

DevicePacked prune_unused(Context &ctx, DevicePacked p) {
  Graph g = build_graph(ctx, p);
  if (!g.n) return p;
  Buffer<int> keep(ctx, g.n), zero(ctx, 1);
  zero.clear(ctx);
  nonzero<<<grid(g.n), kThreads, 0, ctx.stream()>>>(g.degree.data(), g.n, keep.data(), zero.data());
  BROOK_LAUNCHED();
  if (!read_one(ctx, zero.data())) return p;
  return select_vertices(ctx, p, keep, g);
}

// ---------------------------------------------------------------------------------------------
// remove_dust (post.cu:115-137, 443-475)

__global__ void edge_lengths(const float *v, const int2 *e, const int *root, int m, double *total, int *count) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= m) return;
  const int a = e[i].x, b = e[i].y;
  const double x = double(v[3LL * b] - v[3LL * a]), y = double(v[3LL * b + 1] - v[3LL * a + 1]),
               z = double(v[3LL * b + 2] - v[3LL * a + 2]);
  const int r = root[a];
  atomicAdd(total + r, sqrt((x * x + y * y) + z * z));
  atomicAdd(count + r, 1);
}

// Keep a root whose double total clears the threshold; a root within count * 2^-23 * threshold
// of it is marked undecided (2) for the float32 pairwise rule.
__global__ void dust_decide(const double *total, const int *count, int n, double threshold, int *keep_root) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= n) return;
  const bool near = count[r] > 0 && fabs(total[r] - threshold) <= count[r] * (1.0 / 8388608.0) * threshold;
  keep_root[r] = near ? 2 : (total[r] > threshold ? 1 : 0);
}

__device__ __forceinline__ float length32(const float *v, int a, int b) {
  float x = v[3LL * b] - v[3LL * a], y = v[3LL * b + 1] - v[3LL * a + 1], z = v[3LL * b + 2] - v[3LL * a + 2];
  x = x * x;
  y = y * y;
  z = z * z;
  return sqrtf((x + y) + z);
}

// NumPy's pairwise float32 sum (numpy loops_utils pairwise_sum, block 128, eight partial sums)
// over the lengths of `count` edges listed at `list`, without recursion: the split tree is
// walked depth first, left before right, with partial sums on an explicit stack.
__device__ float pairwise_lengths(const float *v, const int2 *e, const int *list, int count) {
  struct Frame { int start, n, state; float left; };
  Frame st[40];
  int sp = 0;
  st[sp++] = Frame{0, count, 0, 0.0f};
  float result = 0.0f;
  while (sp) {
    Frame &f = st[sp - 1];
    if (f.n <= 128) {
      float s;
      if (f.n < 8) {
        s = -0.0f;
        for (int i = 0; i < f.n; ++i) s += length32(v, e[list[f.start + i]].x, e[list[f.start + i]].y);
      } else {
        float r[8];
        for (int j = 0; j < 8; ++j) r[j] = length32(v, e[list[f.start + j]].x, e[list[f.start + j]].y);
        int i = 8;
        for (; i + 8 <= f.n; i += 8)
          for (int j = 0; j < 8; ++j) r[j] += length32(v, e[list[f.start + i + j]].x, e[list[f.start + i + j]].y);
        s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < f.n; ++i) s += length32(v, e[list[f.start + i]].x, e[list[f.start + i]].y);
      }
      --sp;
      if (!sp) { result = s; break; }
      Frame &up = st[sp - 1];  // a finished half returns to its parent
      if (up.state == 1) up.left = s, up.state = 2;
      else up.left = up.left + s, up.state = 3;
      continue;
    }
    int half = f.n / 2;
    half -= half % 8;
    if (f.state == 0) {
      f.state = 1;
      st[sp++] = Frame{f.start, half, 0, 0.0f};
    } else if (f.state == 2) {
      st[sp++] = Frame{f.start + half, f.n - half, 0, 0.0f};
    } else {  // both halves done
      const float s = f.left;
      --sp;
      if (!sp) { result = s; break; }
      Frame &up = st[sp - 1];
      if (up.state == 1) up.left = s, up.state = 2;
      else up.left = up.left + s, up.state = 3;
    }
  }
  return result;
}

// One lane per undecided root: its edges in edge order (list segment) under the float32 rule.
__global__ void dust_pairwise(const float *v, const int2 *e, const int *roots, const int *list, const long long *seg, int count,
                              double threshold, int *keep_root) {
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c >= count) return;
  const float s = pairwise_lengths(v, e, list + seg[c], int(seg[c + 1] - seg[c]));
  keep_root[roots[c]] = double(s) > threshold ? 1 : 0;
}

__global__ void undecided_keys(const int2 *e, const int *root, const int *slot, int m, int sentinel, int *key, int *index) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= m) return;
  const int s = slot[root[e[i].x]];
  key[i] = s >= 0 ? s : sentinel;
  index[i] = i;
}

__global__ void undecided_slots(const int *keep_root, const int *pos, int n, int *slot, int *roots) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= n) return;
  slot[r] = keep_root[r] == 2 ? pos[r] : -1;
  if (keep_root[r] == 2) roots[pos[r]] = r;
}

__global__ void is_two(const int *keep_root, int n, int *out) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r < n) out[r] = keep_root[r] == 2;
}

__global__ void broadcast_root(const int *keep_root, const int *root, int n, int *keep) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) keep[i] = keep_root[root[i]];
}

__global__ void count_per_slot(const int *key, int m, int count, int *per) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < m && key[i] < count) atomicAdd(per + key[i], 1);
}

DevicePacked remove_dust(Context &ctx, DevicePacked p, double threshold) {
  if (threshold == 0 || p.n() == 0) return p;
  Graph g = build_graph(ctx, p);
  auto s = ctx.stream();
  Buffer<double> total(ctx, g.n);
  Buffer<int> count(ctx, g.n), keep_root(ctx, g.n);
  total.clear(ctx);
  count.clear(ctx);
  if (g.m) edge_lengths<<<grid(g.m), kThreads, 0, s>>>(p.vertices.data(), g.e.data(), g.root.data(), g.m, total.data(), count.data());
  dust_decide<<<grid(g.n), kThreads, 0, s>>>(total.data(), count.data(), g.n, threshold, keep_root.data());
  BROOK_LAUNCHED();
  // undecided roots: their edges grouped by root in edge order (a stable sort by slot)
  Buffer<int> two(ctx, g.n + 1), pos(ctx, g.n + 1);
  is_two<<<grid(g.n), kThreads, 0, s>>>(keep_root.data(), g.n, two.data());
  BROOK_HIP(hipMemsetAsync(two.data() + g.n, 0, 4, s));
  exclusive_sum(ctx, two.data(), pos.data(), g.n + 1);
  const int undecided = read_one(ctx, pos.data() + g.n);
  if (undecided) {
    Buffer<int> slot(ctx, g.n), roots(ctx, undecided);
    undecided_slots<<<grid(g.n), kThreads, 0, s>>>(keep_root.data(), pos.data(), g.n, slot.data(), roots.data());
    Buffer<int> key(ctx, g.m), key_sorted(ctx, g.m), index(ctx, g.m), list(ctx, g.m);
    undecided_keys<<<grid(g.m), kThreads, 0, s>>>(g.e.data(), g.root.data(), slot.data(), g.m, undecided, key.data(), index.data());
    BROOK_LAUNCHED();
    Buffer<unsigned char> scratch;
    int bits = 1;
    while ((1LL << bits) <= undecided) ++bits;
    with_scratch(ctx, scratch, [&](void *t, size_t &b) {
      return hipcub::DeviceRadixSort::SortPairs(t, b, key.data(), key_sorted.data(), index.data(), list.data(), g.m, 0, bits, s);
    });
    Buffer<int> per(ctx, undecided + 1);
    per.clear(ctx);
    count_per_slot<<<grid(g.m), kThreads, 0, s>>>(key.data(), g.m, undecided, per.data());
    BROOK_LAUNCHED();
    std::vector<int> h = per.download(ctx);
    std::vector<long long> seg(size_t(undecided) + 1, 0);
    for (int c = 0; c < undecided; ++c) seg[c + 1] = seg[c] + h[c];
    Buffer<long long> seg_dev(ctx, seg.size());
    seg_dev.upload(ctx, seg.data(), seg.size());
    dust_pairwise<<<grid(undecided), kThreads, 0, s>>>(p.vertices.data(), g.e.data(), roots.data(), list.data(), seg_dev.data(),
                                                       undecided, threshold, keep_root.data());
    BROOK_LAUNCHED();
  }
  Buffer<int> keep(ctx, g.n);
  broadcast_root<<<grid(g.n), kThreads, 0, s>>>(keep_root.data(), g.root.data(), g.n, keep.data());
  BROOK_LAUNCHED();
  return select_vertices(ctx, p, keep, g);
}

// ---------------------------------------------------------------------------------------------
// Components taking part in a sequential rule: their vertices (ascending index) and edges (edge
// order) as contiguous segments in slot order, local ids = position in the segment.

__global__ void component_sizes(const int *root, const int *degree, int n, const int2 *e, int m, int *nv, int *ne,
                                int *branch) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    atomicAdd(nv + root[i], 1);
    if (degree[i] >= 3) branch[root[i]] = 1;
  }
  if (i < m) atomicAdd(ne + root[e[i].x], 1);
}

__global__ void pick_components(const int *nv, const int *ne, const int *branch, int n, bool trees, int *flag) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= n) return;
  flag[r] = trees ? (branch[r] && ne[r] == nv[r] - 1) : (ne[r] > 0 && ne[r] - nv[r] + 1 > 0);
}

__global__ void component_slots(const int *flag, const int *pos, const int *nv, const int *ne, int n, int *slot,
                                int *slot_nv, int *slot_ne) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= n) return;
  slot[r] = flag[r] ? pos[r] : -1;
  if (flag[r]) {
    slot_nv[pos[r]] = nv[r];
    slot_ne[pos[r]] = ne[r];
  }
}

__global__ void slot_keys_vertices(const int *root, const int *slot, int n, int sentinel, int *key, int *index) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int s = slot[root[i]];
  key[i] = s >= 0 ? s : sentinel;
  index[i] = i;
}

__global__ void slot_keys_edges(const int2 *e, const int *root, const int *slot, int m, int sentinel, int *key, int *index) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= m) return;
  const int s = slot[root[e[i].x]];
  key[i] = s >= 0 ? s : sentinel;
  index[i] = i;
}

__global__ void local_ids(const int *vlist, const int *root, const int *slot, const long long *vs, int count, int *local) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= count) return;
  const int v = vlist[j];
  local[v] = int(j - vs[slot[root[v]]]);
}

__global__ void local_rows(const int *elist, const int2 *e, const int *local, int count, int2 *rows) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= count) return;
  const int2 ge = e[elist[j]];
  rows[j] = make_int2(local[ge.x], local[ge.y]);
}

struct Pieces {
  int count = 0, nv = 0, ne = 0;  // components, their vertices and edges
  Buffer<int> slot;               // per root: slot or -1
  Buffer<int> vlist, elist;       // global ids in slot order
  Buffer<long long> vs, es;       // per slot: segment starts (count + 1)
  std::vector<long long> vs_host, es_host;
  Buffer<int> local;              // per vertex of a piece: local id
  Buffer<int2> rows;              // per edge of a piece: local (x, y) as stored
};

Pieces pieces_of(Context &ctx, const Graph &g, bool trees) {
  Pieces pc;
  auto s = ctx.stream();
  Buffer<int> nv(ctx, g.n), ne(ctx, g.n), branch(ctx, g.n), flag(ctx, g.n + 1), pos(ctx, g.n + 1);
  nv.clear(ctx);
  ne.clear(ctx);
  branch.clear(ctx);
  component_sizes<<<grid(std::max(g.n, g.m)), kThreads, 0, s>>>(g.root.data(), g.degree.data(), g.n, g.e.data(), g.m, nv.data(),
                                                               ne.data(), branch.data());
  pick_components<<<grid(g.n), kThreads, 0, s>>>(nv.data(), ne.data(), branch.data(), g.n, trees, flag.data());
  BROOK_LAUNCHED();
  BROOK_HIP(hipMemsetAsync(flag.data() + g.n, 0, 4, s));
  exclusive_sum(ctx, flag.data(), pos.data(), g.n + 1);
  pc.count = read_one(ctx, pos.data() + g.n);
  if (!pc.count) return pc;
  pc.slot = Buffer<int>(ctx, g.n);
  Buffer<int> snv(ctx, pc.count), sne(ctx, pc.count);
  component_slots<<<grid(g.n), kThreads, 0, s>>>(flag.data(), pos.data(), nv.data(), ne.data(), g.n, pc.slot.data(), snv.data(),
                                                 sne.data());
  BROOK_LAUNCHED();
  const std::vector<int> hv = snv.download(ctx), he = sne.download(ctx);
  pc.vs_host.assign(1, 0);
  pc.es_host.assign(1, 0);
  for (int c = 0; c < pc.count; ++c) {
    pc.vs_host.push_back(pc.vs_host.back() + hv[c]);
    pc.es_host.push_back(pc.es_host.back() + he[c]);
  }
  pc.nv = int(pc.vs_host.back());
  pc.ne = int(pc.es_host.back());
  pc.vs = Buffer<long long>(ctx, pc.vs_host.size());
  pc.es = Buffer<long long>(ctx, pc.es_host.size());
  pc.vs.upload(ctx, pc.vs_host.data(), pc.vs_host.size());
  pc.es.upload(ctx, pc.es_host.data(), pc.es_host.size());
  int bits = 1;
  while ((1LL << bits) <= pc.count) ++bits;
  Buffer<unsigned char> scratch;
  {  // vertices by (slot, index): a stable sort by slot; the other vertices sort last
    Buffer<int> key(ctx, g.n), key_sorted(ctx, g.n), index(ctx, g.n);
    pc.vlist = Buffer<int>(ctx, g.n);
    slot_keys_vertices<<<grid(g.n), kThreads, 0, s>>>(g.root.data(), pc.slot.data(), g.n, pc.count, key.data(), index.data());
    BROOK_LAUNCHED();
    with_scratch(ctx, scratch, [&](void *t, size_t &b) {
      return hipcub::DeviceRadixSort::SortPairs(t, b, key.data(), key_sorted.data(), index.data(), pc.vlist.data(), g.n, 0, bits, s);
    });
  }
  {
    Buffer<int> key(ctx, g.m), key_sorted(ctx, g.m), index(ctx, g.m);
    pc.elist = Buffer<int>(ctx, g.m);
    slot_keys_edges<<<grid(g.m), kThreads, 0, s>>>(g.e.data(), g.root.data(), pc.slot.data(), g.m, pc.count, key.data(),
                                                   index.data());
    BROOK_LAUNCHED();
    with_scratch(ctx, scratch, [&](void *t, size_t &b) {
      return hipcub::DeviceRadixSort::SortPairs(t, b, key.data(), key_sorted.data(), index.data(), pc.elist.data(), g.m, 0, bits, s);
    });
  }
  pc.local = Buffer<int>(ctx, g.n);
  pc.rows = Buffer<int2>(ctx, std::max(pc.ne, 1));
  local_ids<<<grid(pc.nv), kThreads, 0, s>>>(pc.vlist.data(), g.root.data(), pc.slot.data(), pc.vs.data(), pc.nv, pc.local.data());
  if (pc.ne) local_rows<<<grid(pc.ne), kThreads, 0, s>>>(pc.elist.data(), g.e.data(), pc.local.data(), pc.ne, pc.rows.data());
  BROOK_LAUNCHED();
  return pc;
}

// ---------------------------------------------------------------------------------------------
// remove_loops: one lane per cyclic component (post_kernels.cuh:22-199)

struct LoopArena {
  int2 *rows;                            // 2 * E + 2 per component: (x, y), x = -1 once dead
  int *nbr, *nxt, *prv, *rid, *mul;      // 2 * E + 4 entries per component
  int *head, *tail, *cnt, *seen, *order;  // nv + 1 per component
  int *branch, *path;                    // nv + 1 per component
  unsigned char *drop;                   // nv + 1 per component
  int3 *stack;                           // 2 * E + 2 per component: (node, parent, depth)
};

struct LoopLocal {
  int2 *rows;
  int *nbr, *nxt, *prv, *rid, *mul, *head, *tail, *cnt;
  int freelist, bump, entry_cap;
  int nrows;
  long long nlive;
  bool overflow;  // the entry pool ran out (only rows repeated in both orientations can do that)

  __device__ int find(int a, int b) const {
    for (int e = head[a]; e >= 0; e = nxt[e])
      if (nbr[e] == b) return e;
    return -1;
  }
  __device__ int entry() {
    if (freelist >= 0) { const int e = freelist; freelist = nxt[e]; return e; }
    if (bump == entry_cap) { overflow = true; return entry_cap - 1; }  // in bounds; the result is discarded
    return bump++;
  }
  __device__ void link(int a, int b, int r) {
    const int e = entry();
    nbr[e] = b;
    rid[e] = r;
    mul[e] = 1;
    nxt[e] = -1;
    prv[e] = tail[a];
    if (tail[a] >= 0) nxt[tail[a]] = e;
    else head[a] = e;
    tail[a] = e;
  }
  __device__ void unlink(int a, int e) {
    if (prv[e] >= 0) nxt[prv[e]] = nxt[e];
    else head[a] = nxt[e];
    if (nxt[e] >= 0) prv[nxt[e]] = prv[e];
    else tail[a] = prv[e];
    nxt[e] = freelist;
    freelist = e;
  }
  __device__ void add_row(int x, int y) {
    const int r = nrows++;
    rows[r] = make_int2(x, y);
    nlive++;
    cnt[x]++;
    cnt[y]++;
    const int f = find(x, y);
    if (f >= 0) {  // a repeated pair: the first match of each list counts it
      mul[f]++;
      mul[find(y, x)]++;
    } else {
      link(x, y, r);
      link(y, x, r);
    }
  }
};

__device__ __forceinline__ float sq_dist(const float *P, int v, float x, float y, float z) {
  const float dx = P[3 * v] - x, dy = P[3 * v + 1] - y, dz = P[3 * v + 2] - z;
  return (dx * dx + dy * dy) + dz * dz;
}

__global__ void loops_kernel(int count, const long long *vs, const long long *es, const int2 *rows_in, const int *vlist,
                             const float *vertices, const float *radii, LoopArena ar, int2 *rows_out, int *ne_out, int *failed) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= count) return;
  const long long v0 = vs[c], e0 = es[c];
  const int nv = int(vs[c + 1] - v0), E = int(es[c + 1] - e0);
  const int row_cap = 2 * E + 2;
  const long long rbase = 2 * e0 + 2LL * c, ebase = 2 * e0 + 4LL * c, nbase = v0 + c;
  LoopLocal s;
  s.rows = ar.rows + rbase;
  s.nbr = ar.nbr + ebase;
  s.nxt = ar.nxt + ebase;
  s.prv = ar.prv + ebase;
  s.rid = ar.rid + ebase;
  s.mul = ar.mul + ebase;
  s.head = ar.head + nbase;
  s.tail = ar.tail + nbase;
  s.cnt = ar.cnt + nbase;
  int *seen = ar.seen + nbase, *order = ar.order + nbase, *branch = ar.branch + nbase, *path = ar.path + nbase;
  unsigned char *drop = ar.drop + nbase;
  int3 *stack = ar.stack + rbase;
  const int *vl = vlist + v0;
  auto P = [&](int v) { return vertices + 3LL * vl[v]; };
  s.freelist = -1;
  s.bump = 0;
  s.entry_cap = 2 * E + 4;
  s.nrows = 0;
  s.nlive = 0;
  s.overflow = false;
  for (int v = 0; v < nv; ++v) { s.head[v] = s.tail[v] = -1; s.cnt[v] = 0; seen[v] = 0; }
  for (int i = 0; i < E; ++i) s.add_row(rows_in[e0 + i].x, rows_in[e0 + i].y);
  int app0 = s.nrows, app1 = s.nrows;  // rows appended by the previous iteration
  const long long budget = 64LL * (E + nv) + 64;
  for (long long iteration = 0;; ++iteration) {
    if (iteration > budget) { atomicOr(failed, 1); break; }
    if (s.overflow) { atomicOr(failed, 2); break; }
    if (s.nlive == 0) break;
    int first = 0;
    while (s.rows[first].x < 0) first++;
    // depth-first search from the first live row's first end
    int sp = 0, plen = 0, node = -1, nseen = 0;
    bool found = false;
    stack[sp++] = make_int3(s.rows[first].x, -1, 0);
    while (sp) {
      const int3 f = stack[--sp];
      node = f.x;
      if (plen > f.z) plen = f.z;
      path[plen++] = node;
      if (seen[node]) { found = true; break; }
      seen[node] = 1;
      order[nseen++] = node;
      for (int e = s.head[node]; e >= 0; e = s.nxt[e])
        if (s.nbr[e] != f.y) stack[sp++] = make_int3(s.nbr[e], node, f.z + 1);
    }
    for (int j = 0; j < nseen; ++j) seen[order[j]] = 0;
    if (!found || plen <= 1) break;
    int i0 = 0;
    while (i0 < plen - 1 && path[i0] != node) i0++;
    if (plen - i0 < 3) break;
    const int *cy = path + i0;
    const int cl = plen - i0;
    int nb = 0;
    for (int j = 0; j + 1 < cl; ++j)
      if (s.cnt[cy[j]] >= 3) branch[nb++] = cy[j];
    for (int i = 1; i < nb; ++i) {  // ascending (few)
      const int x = branch[i];
      int j = i - 1;
      while (j >= 0 && branch[j] > x) { branch[j + 1] = branch[j]; --j; }
      branch[j + 1] = x;
    }
    for (int j = 0; j < cl; ++j) drop[j] = 0;
    int add_a = -1, add_b = -1, hub = -1;
    bool star = false;
    if (nb == 0) {
      for (int j = 0; j + 1 < cl; ++j) drop[j] = 1;
    } else if (nb == 1) {
      const int bn = branch[0];
      int end = -1;
      float best = -1.0f;
      for (int j = 0; j + 1 < cl; ++j) {
        const int v = cy[j];
        const float d = sq_dist(P(v), 0, P(bn)[0], P(bn)[1], P(bn)[2]);
        if (end < 0 || d > best || (d == best && v < end)) { best = d; end = v; }
      }
      for (int j = 0; j + 1 < cl; ++j) drop[j] = 1;
      add_a = bn;
      add_b = end;
    } else if (nb == 2) {
      const int L = cl - 1;
      int p0 = -1, p1 = -1;
      for (int j = 0; j < L; ++j)
        if (s.cnt[cy[j + 1]] >= 3) {
          if (p0 < 0) p0 = j;
          else p1 = j;
        }
      const bool inner = double(p1 - p0) < double(L) / 2.0;
      for (int j = 0; j + 1 < cl; ++j) {
        const bool on_inner = j - 1 >= 0 && j - 1 >= p0 && j <= p1;
        if (!(inner ? on_inner : !on_inner)) drop[j] = 1;
      }
    } else {
      float cx = 0.0f, cyy = 0.0f, cz = 0.0f;
      for (int i = 0; i < nb; ++i) { cx += P(branch[i])[0]; cyy += P(branch[i])[1]; cz += P(branch[i])[2]; }
      cx /= float(nb);
      cyy /= float(nb);
      cz /= float(nb);
      float best = 0.0f;
      for (int v = 0; v < nv; ++v) {
        const float d = sq_dist(P(v), 0, cx, cyy, cz);
        if (hub < 0 || d < best) { best = d; hub = v; }
      }
      float far = 0.0f;
      for (int i = 0; i < nb; ++i) {
        const float d = sq_dist(P(branch[i]), 0, P(hub)[0], P(hub)[1], P(hub)[2]);
        if (i == 0 || d > far) far = d;
      }
      if (sqrtf(far) > radii[vl[hub]]) {
        drop[0] = 1;
      } else {
        for (int j = 0; j + 1 < cl; ++j) drop[j] = 1;
        star = true;
      }
    }
    for (int r = app0; r < app1; ++r)
      if (s.rows[r].x > s.rows[r].y) s.rows[r] = make_int2(s.rows[r].y, s.rows[r].x);
    for (int j = 0; j + 1 < cl; ++j) {
      if (!drop[j]) continue;
      const int a = cy[j], b = cy[j + 1];
      const int e = s.find(a, b);
      if (e < 0) continue;
      const int m = s.mul[e];
      if (m == 1) {
        s.rows[s.rid[e]].x = -1;
      } else {
        const int lo = min(a, b), hi = max(a, b);
        for (int r = first; r < s.nrows; ++r)
          if (s.rows[r].x == lo && s.rows[r].y == hi) s.rows[r].x = -1;
      }
      s.nlive -= m;
      s.cnt[a] -= m;
      s.cnt[b] -= m;
      const int back = s.find(b, a);
      s.unlink(a, e);
      s.unlink(b, back);
    }
    const int adds = add_b >= 0 ? 1 : (star ? nb : 0);
    if (s.nrows + adds > row_cap) {  // squeeze out the dead rows, keeping the order of the live ones
      int w = 0;
      for (int r = 0; r < s.nrows; ++r)
        if (s.rows[r].x >= 0) s.rows[w++] = s.rows[r];
      s.nrows = w;
      for (int v = 0; v < nv; ++v)
        for (int e = s.head[v]; e >= 0; e = s.nxt[e]) s.rid[e] = -1;
      for (int r = 0; r < s.nrows; ++r) {
        const int f = s.find(s.rows[r].x, s.rows[r].y);
        if (f >= 0 && s.rid[f] < 0) {  // (a row whose pair lost its entry has none)
          s.rid[f] = r;
          s.rid[s.find(s.rows[r].y, s.rows[r].x)] = r;
        }
      }
    }
    if (s.nrows + adds > row_cap) { atomicOr(failed, 2); break; }  // live rows outgrew the input's count
    app0 = s.nrows;
    if (add_b >= 0) s.add_row(add_a, add_b);
    if (star)
      for (int i = 0; i < nb; ++i)
        if (branch[i] != hub) s.add_row(branch[i], hub);
    app1 = s.nrows;
  }
  int w = 0;
  for (int r = 0; r < s.nrows; ++r) w += s.rows[r].x >= 0;
  if (w > E) {  // more live rows than the component had (rows repeated in both orientations)
    atomicOr(failed, 2);
    w = 0;
  }
  for (int r = 0, k = 0; r < s.nrows && k < w; ++r)
    if (s.rows[r].x >= 0) rows_out[e0 + k++] = s.rows[r];
  ne_out[c] = w;
}

__global__ void loop_keys_untouched(const int2 *e, const int *root, const int *slot, int m, unsigned long long *keys) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= m) return;
  const int2 x = e[i];
  keys[i] = slot[root[x.x]] < 0 && x.x != x.y
                ? ((unsigned long long)min(x.x, x.y) << 32) | (unsigned long long)max(x.x, x.y)
                : kNone;
}

__global__ void loop_keys_kept(const int2 *rows, const int *ne_out, const long long *es, const long long *vs, const int *vlist,
                               const int *elist, const int *comp_of_edge, int count_edges, unsigned long long *keys) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= count_edges) return;
  const int c = comp_of_edge[j];
  if (j - es[c] >= ne_out[c]) return;
  const unsigned a = unsigned(vlist[vs[c] + rows[j].x]), b = unsigned(vlist[vs[c] + rows[j].y]);
  if (a != b) keys[elist[j]] = ((unsigned long long)min(a, b) << 32) | (unsigned long long)max(a, b);
}

__global__ void edge_component(const long long *es, int count, int *comp_of_edge) {
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c >= count) return;
  for (long long j = es[c]; j < es[c + 1]; ++j) comp_of_edge[j] = c;
}

__global__ void keep_unemptied(const int *root, const int *slot, const int *ne_out, int n, int *keep) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int c = slot[root[i]];
  keep[i] = c < 0 || ne_out[c] > 0;
}

DevicePacked remove_loops(Context &ctx, DevicePacked p) {
  if (!p.n() || !p.m()) return p;
  Graph g = build_graph(ctx, p);
  Pieces pc = pieces_of(ctx, g, false);
  if (!pc.count) return p;
  auto s = ctx.stream();
  const long long A = pc.count, N = pc.nv, M = pc.ne;
  Buffer<int2> rows(ctx, size_t(2 * M + 2 * A)), rows_out(ctx, size_t(std::max<long long>(M, 1)));
  Buffer<int3> frames(ctx, size_t(2 * M + 2 * A));
  const size_t entries = size_t(2 * M + 4 * A), nodes = size_t(N + A);
  Buffer<int> nbr(ctx, entries), nxt(ctx, entries), prv(ctx, entries), rid(ctx, entries), mul(ctx, entries);
  Buffer<int> head(ctx, nodes), tail(ctx, nodes), cnt(ctx, nodes), seen(ctx, nodes), order(ctx, nodes), branch(ctx, nodes),
      path(ctx, nodes);
  Buffer<unsigned char> drop(ctx, nodes);
  Buffer<int> ne_out(ctx, A), failed(ctx, 1);
  failed.clear(ctx);
  LoopArena ar{rows.data(), nbr.data(), nxt.data(), prv.data(), rid.data(), mul.data(), head.data(), tail.data(), cnt.data(),
               seen.data(), order.data(), branch.data(), path.data(), drop.data(), frames.data()};
  loops_kernel<<<unsigned((A + 63) / 64), 64, 0, s>>>(pc.count, pc.vs.data(), pc.es.data(), pc.rows.data(), pc.vlist.data(),
                                                     p.vertices.data(), p.radii.data(), ar, rows_out.data(), ne_out.data(),
                                                     failed.data());
  BROOK_LAUNCHED();
  const int failure = read_one(ctx, failed.data());
  if (failure & 1) throw std::runtime_error("remove_loops did not converge");
  if (failure & 2) throw std::runtime_error("remove_loops row capacity exceeded (rows repeated in both orientations)");
  Buffer<unsigned long long> keys(ctx, size_t(std::max(g.m, 1)));
  Buffer<int> comp_of_edge(ctx, size_t(std::max<long long>(M, 1)));
  loop_keys_untouched<<<grid(g.m), kThreads, 0, s>>>(g.e.data(), g.root.data(), pc.slot.data(), g.m, keys.data());
  edge_component<<<grid(A), kThreads, 0, s>>>(pc.es.data(), pc.count, comp_of_edge.data());
  if (M)
    loop_keys_kept<<<grid(M), kThreads, 0, s>>>(rows_out.data(), ne_out.data(), pc.es.data(), pc.vs.data(), pc.vlist.data(),
                                                pc.elist.data(), comp_of_edge.data(), int(M), keys.data());
  BROOK_LAUNCHED();
  const std::vector<int> kept = ne_out.download(ctx);
  DevicePacked out = edges_from_keys(ctx, std::move(p), g, keys, g.m);
  if (std::find(kept.begin(), kept.end(), 0) == kept.end()) return out;
  Buffer<int> keep(ctx, g.n);
  keep_unemptied<<<grid(g.n), kThreads, 0, s>>>(g.root.data(), pc.slot.data(), ne_out.data(), g.n, keep.data());
  BROOK_LAUNCHED();
  Graph next = build_graph(ctx, out);
  return select_vertices(ctx, out, keep, next);
}

// ---------------------------------------------------------------------------------------------
// remove_ticks: one lane per tree with a branch point (post_kernels.cuh:309-405)

// CSR neighbours of the pieces' vertices: a node's rows that start at it, then those that end at
// it, each in row order (a stable sort of both directions by source).
__global__ void csr_entries(const int2 *rows, const long long *vs, const long long *es, const int *comp_of_edge, int M,
                            unsigned *source, unsigned long long *value) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= M) return;
  const int c = comp_of_edge[j];
  const unsigned r = unsigned(j - es[c]);
  const int2 x = rows[j];
  source[j] = unsigned(vs[c] + x.x);
  value[j] = ((unsigned long long)unsigned(x.y) << 32) | r;
  source[M + j] = unsigned(vs[c] + x.y);
  value[M + j] = ((unsigned long long)unsigned(x.x) << 32) | r;
}

__global__ void csr_degree(const unsigned *source_sorted, int total, int *deg) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j < total) atomicAdd(deg + source_sorted[j], 1);
}

struct TickArena {
  int *bc, *slot_head;               // nv per component
  int4 *stack;                       // (node, parent, root, chain), nv + 1 per component
  float *stack_dist;
  int *su, *sv, *ver, *chain_first, *chain_last, *chain_next;  // E + 1 per component (chain ids)
  double *len;
  unsigned char *alive, *eligible, *dead;
  int *slot_next, *slot_owner;       // 2 (E + 1) per component
  double *heap_len;                  // 2 (E + 1) + 2 per component
  int *heap_lo, *heap_hi, *heap_id, *heap_ver;
  int *chain_of;                     // E per component (row -> chain)
};

struct Heap {
  double *len;
  int *lo, *hi, *id, *ver;
  int n = 0;
  __device__ bool less(int i, int j) const {
    return len[i] < len[j] || (len[i] == len[j] && (lo[i] < lo[j] || (lo[i] == lo[j] && hi[i] < hi[j])));
  }
  __device__ void swap(int i, int j) {
    const double l = len[i]; len[i] = len[j]; len[j] = l;
    int t = lo[i]; lo[i] = lo[j]; lo[j] = t;
    t = hi[i]; hi[i] = hi[j]; hi[j] = t;
    t = id[i]; id[i] = id[j]; id[j] = t;
    t = ver[i]; ver[i] = ver[j]; ver[j] = t;
  }
  __device__ void push(double l, int a, int b, int s, int v) {
    int i = n++;
    len[i] = l; lo[i] = min(a, b); hi[i] = max(a, b); id[i] = s; ver[i] = v;
    while (i > 0) {
      const int up = (i - 1) >> 1;
      if (!less(i, up)) break;
      swap(i, up);
      i = up;
    }
  }
  __device__ void pop() {
    if (--n <= 0) return;
    swap(0, n);
    int i = 0;
    for (;;) {
      const int l = 2 * i + 1, r = l + 1;
      int m = i;
      if (l < n && less(l, m)) m = l;
      if (r < n && less(r, m)) m = r;
      if (m == i) break;
      swap(i, m);
      i = m;
    }
  }
};

__global__ void ticks_kernel(int count, const long long *vs, const long long *es, const int *vlist, const float *vertices,
                             const long long *csr_off, const unsigned long long *csr, double threshold, TickArena ar,
                             unsigned char *keep_rows) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= count) return;
  const long long v0 = vs[c], e0 = es[c];
  const int nv = int(vs[c + 1] - v0), E = int(es[c + 1] - e0);
  const long long cbase = e0 + c, sbase = 2 * (e0 + c), hbase = 2 * (e0 + c) + 2LL * c;
  int *bc = ar.bc + v0, *slot_head = ar.slot_head + v0;
  int4 *stack = ar.stack + v0 + c;
  float *sdist = ar.stack_dist + v0 + c;
  int *su = ar.su + cbase, *sv = ar.sv + cbase, *ver = ar.ver + cbase;
  int *cf = ar.chain_first + cbase, *cla = ar.chain_last + cbase, *cn = ar.chain_next + cbase;
  double *len = ar.len + cbase;
  unsigned char *alive = ar.alive + cbase, *eligible = ar.eligible + cbase, *dead = ar.dead + cbase;
  int *snext = ar.slot_next + sbase, *sowner = ar.slot_owner + sbase;
  int *chain_of = ar.chain_of + e0;
  Heap heap{ar.heap_len + hbase, ar.heap_lo + hbase, ar.heap_hi + hbase, ar.heap_id + hbase, ar.heap_ver + hbase, 0};
  const long long *off = csr_off + v0;
  const int *vl = vlist + v0;
  auto deg = [&](int v) { return int(off[v + 1] - off[v]); };
  int start = -1;
  for (int v = 0; v < nv; ++v) {
    const int d = deg(v);
    bc[v] = d >= 3 ? d : 0;
    slot_head[v] = -1;
    if (start < 0 && d == 1) start = v;
  }
  for (int r = 0; r < E; ++r) keep_rows[e0 + r] = 1;
  if (start < 0) return;
  int chains = 0, sp = 0;
  stack[sp] = make_int4(start, -1, start, -1);
  sdist[sp++] = 0.0f;
  while (sp) {
    --sp;
    const int node = stack[sp].x, parent = stack[sp].y;
    int root = stack[sp].z, chain = stack[sp].w;
    float dist = sdist[sp];
    const int d = deg(node);
    if ((d == 1 || d >= 3) && node != root) {  // a chain reaches a critical node: a super-edge
      su[chain] = root;
      sv[chain] = node;
      len[chain] = double(dist);
      ver[chain] = 0;
      alive[chain] = 1;
      eligible[chain] = deg(root) == 1 || d == 1;
      snext[2 * chain] = slot_head[root];
      slot_head[root] = 2 * chain;
      snext[2 * chain + 1] = slot_head[node];
      slot_head[node] = 2 * chain + 1;
      sowner[2 * chain] = sowner[2 * chain + 1] = chain;
      if (eligible[chain]) heap.push(len[chain], root, node, chain, 0);
      dist = 0.0f;
      root = node;
    }
    const float *pn = vertices + 3LL * vl[node];
    for (long long q = off[node]; q < off[node + 1]; ++q) {
      const int child = int(csr[q] >> 32), row = int(unsigned(csr[q]));
      if (child == parent) continue;
      const float *pc = vertices + 3LL * vl[child];
      float dx = pn[0] - pc[0], dy = pn[1] - pc[1], dz = pn[2] - pc[2];
      dx *= dx;
      dy *= dy;
      dz *= dz;
      int ch = chain;
      if (node == root) {  // a chain leaves a critical node
        ch = chains++;
        cf[ch] = cla[ch] = ch;
        cn[ch] = -1;
        dead[ch] = 0;
        alive[ch] = 0;
        eligible[ch] = 0;
      }
      chain_of[row] = ch;
      stack[sp] = make_int4(child, node, root, ch);
      sdist[sp++] = dist + sqrtf((dx + dy) + dz);
    }
  }
  int nalive = 0;
  for (int k = 0; k < chains; ++k) nalive += alive[k];
  while (nalive > 1) {
    while (heap.n > 0 && (!alive[heap.id[0]] || ver[heap.id[0]] != heap.ver[0])) heap.pop();
    if (heap.n == 0) break;
    const int best = heap.id[0];
    const int e1 = su[best], e2 = sv[best];
    if (bc[e1] == 1 && bc[e2] == 1) break;
    if (len[best] >= threshold) break;
    heap.pop();
    for (int k = cf[best]; k >= 0; k = cn[k]) dead[k] = 1;
    alive[best] = 0;
    nalive--;
    bc[e1] -= 1;
    bc[e2] -= 1;
    for (int t = 0; t < 2; ++t) {
      const int v = t ? e2 : e1;
      if (bc[v] != 2) continue;
      int a = -1, b = -1;
      for (int k = slot_head[v]; k >= 0; k = snext[k]) {
        const int o = sowner[k];
        if (!alive[o] || (su[o] != v && sv[o] != v) || o == a) continue;
        if (a < 0) a = o;
        else { b = o; break; }
      }
      if (a < 0 || b < 0) { bc[v] = 0; continue; }
      const int ea = su[a] == v ? sv[a] : su[a], eb = su[b] == v ? sv[b] : su[b];
      len[a] = (0.0 + len[a]) + len[b];
      su[a] = ea;
      sv[a] = eb;
      eligible[a] = 1;
      cn[cla[a]] = cf[b];
      cla[a] = cla[b];
      alive[b] = 0;
      nalive--;
      for (int k = slot_head[eb]; k >= 0; k = snext[k])
        if (sowner[k] == b) sowner[k] = a;
      ver[a] += 1;
      heap.push(len[a], ea, eb, a, ver[a]);
      bc[v] = 0;
    }
  }
  for (int r = 0; r < E; ++r) keep_rows[e0 + r] = !dead[chain_of[r]];
}

__global__ void tick_keep(const int *elist, const unsigned char *keep_rows, int M, int *keep) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j < M) keep[elist[j]] = keep_rows[j];
}

__global__ void fill_ones(int *out, int n) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) out[i] = 1;
}

DevicePacked remove_ticks(Context &ctx, DevicePacked p, double threshold) {
  if (threshold == 0 || !p.n() || !p.m()) return p;
  Graph g = build_graph(ctx, p);
  Pieces pc = pieces_of(ctx, g, true);
  if (!pc.count) return p;
  auto s = ctx.stream();
  const long long A = pc.count, N = pc.nv, M = pc.ne;
  if (M > INT32_MAX / 2) throw std::invalid_argument("tick adjacency exceeds int32 capacity");
  Buffer<int> comp_of_edge(ctx, size_t(M));
  edge_component<<<grid(A), kThreads, 0, s>>>(pc.es.data(), pc.count, comp_of_edge.data());
  Buffer<unsigned> source(ctx, size_t(2 * M)), source_sorted(ctx, size_t(2 * M));
  Buffer<unsigned long long> value(ctx, size_t(2 * M)), csr(ctx, size_t(2 * M));
  csr_entries<<<grid(M), kThreads, 0, s>>>(pc.rows.data(), pc.vs.data(), pc.es.data(), comp_of_edge.data(), int(M),
                                           source.data(), value.data());
  BROOK_LAUNCHED();
  Buffer<unsigned char> scratch;
  with_scratch(ctx, scratch, [&](void *t, size_t &b) {
    return hipcub::DeviceRadixSort::SortPairs(t, b, source.data(), source_sorted.data(), value.data(), csr.data(), int(2 * M), 0,
                                              32, s);
  });
  Buffer<int> deg(ctx, size_t(N) + 1);
  deg.clear(ctx);
  csr_degree<<<grid(2 * M), kThreads, 0, s>>>(source_sorted.data(), int(2 * M), deg.data());
  BROOK_LAUNCHED();
  Buffer<long long> off(ctx, size_t(N) + 1);
  {
    std::vector<int> h = deg.download(ctx);
    std::vector<long long> o(size_t(N) + 1, 0);
    for (long long v = 0; v < N; ++v) o[v + 1] = o[v] + h[v];
    off.upload(ctx, o.data(), o.size());
  }
  const size_t nodes = size_t(N), nodes1 = size_t(N + A), chains = size_t(M + A), slots = size_t(2 * (M + A)),
               heapsz = size_t(2 * (M + A) + 2 * A);
  Buffer<int> bc(ctx, nodes), slot_head(ctx, nodes), su(ctx, chains), sv(ctx, chains), ver(ctx, chains), cf(ctx, chains),
      cl(ctx, chains), cn(ctx, chains), snext(ctx, slots), sowner(ctx, slots), hlo(ctx, heapsz), hhi(ctx, heapsz),
      hid(ctx, heapsz), hver(ctx, heapsz), chain_of(ctx, size_t(M));
  Buffer<int4> stack(ctx, nodes1);
  Buffer<float> sdist(ctx, nodes1);
  Buffer<double> len(ctx, chains), hlen(ctx, heapsz);
  Buffer<unsigned char> alive(ctx, chains), eligible(ctx, chains), dead(ctx, chains), keep_rows(ctx, size_t(M));
  TickArena ar{bc.data(), slot_head.data(), stack.data(), sdist.data(), su.data(), sv.data(), ver.data(), cf.data(), cl.data(),
               cn.data(), len.data(), alive.data(), eligible.data(), dead.data(), snext.data(), sowner.data(), hlen.data(),
               hlo.data(), hhi.data(), hid.data(), hver.data(), chain_of.data()};
  ticks_kernel<<<unsigned((A + 63) / 64), 64, 0, s>>>(pc.count, pc.vs.data(), pc.es.data(), pc.vlist.data(), p.vertices.data(),
                                                     off.data(), csr.data(), threshold, ar, keep_rows.data());
  Buffer<int> keep(ctx, size_t(g.m));
  fill_ones<<<grid(g.m), kThreads, 0, s>>>(keep.data(), g.m);
  tick_keep<<<grid(M), kThreads, 0, s>>>(pc.elist.data(), keep_rows.data(), int(M), keep.data());
  BROOK_LAUNCHED();
  return select_edges(ctx, std::move(p), keep, g);
}

// ---------------------------------------------------------------------------------------------
// join_components (post.cu:521-648, post_kernels.cuh:210-306)

__global__ void join_statistics(const int *root, const int *vlabel, const float *radii, int n, int *ncomp, unsigned *rmax) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int l = vlabel[i];
  if (root[i] == i) atomicAdd(ncomp + l, 1);
  const float r = radii[i];
  if (r > 0) atomicMax(rmax + l, __float_as_uint(r));
}

// Cell of a vertex of a label with several components: (rank, floor(coordinate / bound)).
__global__ void join_cells(const int *vlabel, const int *rank, const double *bound, const float *v, int n, long long *kl,
                           long long *kx, long long *ky, long long *kz, int *index, int *member) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int r = rank[vlabel[i]];
  member[i] = r >= 0;
  index[i] = i;
  if (r < 0) {
    kl[i] = LLONG_MAX;
    kx[i] = ky[i] = kz[i] = 0;
    return;
  }
  const double b = bound[r];
  kl[i] = r;
  kx[i] = (long long)floor(double(v[3LL * i]) / b);
  ky[i] = (long long)floor(double(v[3LL * i + 1]) / b);
  kz[i] = (long long)floor(double(v[3LL * i + 2]) / b);
}

__global__ void permute_key(const long long *key, const int *order, int n, long long *out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) out[i] = key[order[i]];
}

struct Cells {
  const int *order;  // sorted position -> vertex
  const long long *kl, *kx, *ky, *kz;  // per vertex
  int n;
  __device__ int cmp(int pos, long long l, long long x, long long y, long long z) const {
    const int v = order[pos];
    if (kl[v] != l) return kl[v] < l ? -1 : 1;
    if (kx[v] != x) return kx[v] < x ? -1 : 1;
    if (ky[v] != y) return ky[v] < y ? -1 : 1;
    if (kz[v] != z) return kz[v] < z ? -1 : 1;
    return 0;
  }
};

constexpr int kJoinTable = 16;

// Per vertex i of a joinable label: for each component above its own within reach, the nearest
// vertex (ties by the lower index); a vertex meeting more components than its table holds sends
// the excess straight out. mode 0 counts, mode 1 writes from offs[i].
__global__ void join_candidates(int n, int mode, Cells cells, const int *member, const int *vlabel, const int *rank,
                                const double *bound, const float *v, const int *root, long long *counts, const long long *offs,
                                int *oa, int *ob, double *od, int *ov, int *ou) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n || !member[i]) return;
  const int r = rank[vlabel[i]];
  const double bnd = bound[r];
  float b32 = float(bnd);
  b32 = b32 * b32 * 1.0001f + 1e-30f;
  const float px = v[3LL * i], py = v[3LL * i + 1], pz = v[3LL * i + 2];
  const long long cx = cells.kx[i], cy = cells.ky[i], cz = cells.kz[i];
  int tb[kJoinTable], tu[kJoinTable];
  double td[kJoinTable];
  int nt = 0;
  long long emitted = 0;
  const long long base = mode ? offs[i] : 0;
  const int ci = root[i];
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const long long x = cx + dx, y = cy + dy, z = cz + dz;
        int lo = 0, hi = cells.n;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (cells.cmp(mid, r, x, y, z) < 0) lo = mid + 1;
          else hi = mid;
        }
        for (int q = lo; q < cells.n && cells.cmp(q, r, x, y, z) == 0; ++q) {
          const int j = cells.order[q];
          const int cj = root[j];
          if (cj <= ci) continue;
          const float fx = v[3LL * j] - px, fy = v[3LL * j + 1] - py, fz = v[3LL * j + 2] - pz;
          if ((fx * fx + fy * fy) + fz * fz > b32) continue;
          const double ex = double(v[3LL * j]) - double(px), ey = double(v[3LL * j + 1]) - double(py),
                       ez = double(v[3LL * j + 2]) - double(pz);
          const double d = sqrt((ex * ex + ey * ey) + ez * ez);
          if (!(d < bnd)) continue;
          int t = -1;
          for (int k = 0; k < nt; ++k)
            if (tb[k] == cj) { t = k; break; }
          if (t >= 0) {
            if (d < td[t] || (d == td[t] && j < tu[t])) { td[t] = d; tu[t] = j; }
          } else if (nt < kJoinTable) {
            tb[nt] = cj; td[nt] = d; tu[nt] = j; nt++;
          } else {
            if (mode) { const long long o = base + emitted; oa[o] = ci; ob[o] = cj; od[o] = d; ov[o] = i; ou[o] = j; }
            emitted++;
          }
        }
      }
  for (int k = 0; k < nt; ++k) {
    if (mode) { const long long o = base + emitted; oa[o] = ci; ob[o] = tb[k]; od[o] = td[k]; ov[o] = i; ou[o] = tu[k]; }
    emitted++;
  }
  if (!mode) counts[i] = emitted;
}

__global__ void candidate_key(const int *order, int n, const int *a, const int *b, const double *d, const int *va, const int *vb,
                              int pass, unsigned long long *key) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int p = order[i];
  if (pass == 0) key[i] = ((unsigned long long)min(va[p], vb[p]) << 32) | (unsigned long long)max(va[p], vb[p]);
  else if (pass == 1) key[i] = (unsigned long long)__double_as_longlong(d[p]);
  else key[i] = ((unsigned long long)unsigned(a[p]) << 32) | (unsigned long long)unsigned(b[p]);
}

__global__ void first_of_pair(const int *order, int n, const int *a, const int *b, int *flag) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  flag[i] = i == 0 || a[order[i]] != a[order[i - 1]] || b[order[i]] != b[order[i - 1]];
}

__global__ void root_flags(const int *root, const int *vlabel, const int *rank, int n, int *flag) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) flag[i] = root[i] == i && rank[vlabel[i]] >= 0;
}

// A component's position among its label's components (ascending root).
__global__ void root_positions(const int *flag, const int *pos, const int *vlabel, const int *rank, const long long *cs, int n,
                               int *local) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n && flag[i]) local[i] = int(pos[i] - cs[rank[vlabel[i]]]);
}

__global__ void compact_pairs(const int *order, const int *flag, const int *fpos, int n, const int *a, const int *b,
                              const double *d, const int *va, const int *vb, const int *local, int *cx, int *cy, double *dd,
                              int *pa, int *pb) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const int p = order[i], j = fpos[i];
  cx[j] = local[a[p]];
  cy[j] = local[b[p]];
  dd[j] = d[p];
  pa[j] = va[p];
  pb[j] = vb[p];
}

// One lane per label: the greedy fusion over the label's candidate list [ps, pe).
__global__ void join_fuse(int labels, const long long *ps, const long long *cs, const double *reach, int *cx, int *cy,
                          const double *dd, const int *va, const int *vb, const float *radii, unsigned char *alive, int *pos,
                          int *slot, int *out_a, int *out_b, int *made) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= labels) return;
  const long long p0 = ps[l], p1 = ps[l + 1], c0 = cs[l];
  const int nc = int(cs[l + 1] - c0);
  int *POS = pos + c0, *SLOT = slot + c0;
  for (int k = 0; k < nc; ++k) { POS[k] = k; SLOT[k] = -1; }
  for (long long q = p0; q < p1; ++q) alive[q] = 1;
  int fusions = 0, count = 0;
  for (;;) {
    long long best = -1;
    float bd = 0.0f;
    int bi = 0, bj = 0;
    for (long long q = p0; q < p1; ++q) {
      if (!alive[q]) continue;
      const double d = dd[q];
      if (d > double(radii[va[q]] + radii[vb[q]])) continue;  // blocked by the radii
      const float d32 = float(d);
      int ia = POS[cx[q]], ib = POS[cy[q]];
      if (ia > ib) { const int t = ia; ia = ib; ib = t; }
      if (best < 0 || d32 < bd || (d32 == bd && (ia < bi || (ia == bi && ib < bj)))) { best = q; bd = d32; bi = ia; bj = ib; }
    }
    if (best < 0 || double(bd) > reach[l]) break;
    const int X = cx[best], Y = cy[best];
    out_a[c0 + count] = va[best];
    out_b[c0 + count] = vb[best];
    count++;
    alive[best] = 0;
    fusions++;
    for (long long q = p0; q < p1; ++q) {  // Y joins X; per other cluster the nearer candidate stays
      if (!alive[q]) continue;
      if (cx[q] == Y) cx[q] = X;
      if (cy[q] == Y) cy[q] = X;
      if (cx[q] == cy[q]) { alive[q] = 0; continue; }
      if (cx[q] != X && cy[q] != X) continue;
      const int k = cx[q] == X ? cy[q] : cx[q];
      if (SLOT[k] < 0) { SLOT[k] = int(q - p0); continue; }
      const long long o = p0 + SLOT[k];
      const int lq = min(va[q], vb[q]), hq = max(va[q], vb[q]), lo = min(va[o], vb[o]), ho = max(va[o], vb[o]);
      const bool q_wins = dd[q] < dd[o] || (dd[q] == dd[o] && (lq < lo || (lq == lo && hq < ho)));
      if (q_wins) { alive[o] = 0; SLOT[k] = int(q - p0); }
      else alive[q] = 0;
    }
    for (long long q = p0; q < p1; ++q)
      if (alive[q]) { SLOT[cx[q]] = -1; SLOT[cy[q]] = -1; }
    POS[X] = -fusions;  // the newest fusion heads the list
  }
  made[l] = count;
}

__global__ void pair_label_offsets(const int *pa, int np, const int *vlabel, const int *rank, int labels, long long *ps) {
  const int l = blockIdx.x * kThreads + threadIdx.x;
  if (l > labels) return;
  int lo = 0, hi = np;  // first candidate whose label rank is >= l
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (rank[vlabel[pa[mid]]] < l) lo = mid + 1;
    else hi = mid;
  }
  ps[l] = lo;
}

__global__ void join_keys(const int2 *e, int m, const int *out_a, const int *out_b, const long long *cs, const int *made,
                          int labels, int roots, unsigned long long *keys) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < m) keys[i] = ((unsigned long long)min(e[i].x, e[i].y) << 32) | (unsigned long long)max(e[i].x, e[i].y);
  if (i < roots) {
    int lo = 0, hi = labels;  // the label whose component range holds i
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (cs[mid] <= i) lo = mid;
      else hi = mid;
    }
    unsigned long long key = kNone;
    if (i - cs[lo] < made[lo])
      key = ((unsigned long long)min(out_a[i], out_b[i]) << 32) | (unsigned long long)max(out_a[i], out_b[i]);
    keys[m + i] = key;
  }
}

DevicePacked join_components(Context &ctx, DevicePacked input) {
  DevicePacked p = prune_unused(ctx, std::move(input));
  if (!p.n() || !p.m()) return p;
  Graph g = build_graph(ctx, p);
  auto s = ctx.stream();
  Buffer<int> ncomp(ctx, g.k);
  Buffer<unsigned> rmax(ctx, g.k);
  ncomp.clear(ctx);
  rmax.clear(ctx);
  join_statistics<<<grid(g.n), kThreads, 0, s>>>(g.root.data(), g.vlabel.data(), p.radii.data(), g.n, ncomp.data(), rmax.data());
  BROOK_LAUNCHED();
  const std::vector<int> nc = ncomp.download(ctx);
  const std::vector<unsigned> bits = rmax.download(ctx);
  std::vector<int> rank(g.k, -1);
  std::vector<double> reach, bound;
  std::vector<long long> cs{0};
  for (int l = 0; l < g.k; ++l)
    if (nc[l] > 1) {
      rank[l] = int(reach.size());
      float r;
      std::memcpy(&r, &bits[l], 4);
      const double x = std::max(2.0 * double(r), 0.0);
      reach.push_back(x);
      bound.push_back(x + 0.000001);
      cs.push_back(cs.back() + nc[l]);
    }
  const int labels = int(reach.size()), roots = int(cs.back());
  if (!labels) return p;
  Buffer<int> rank_dev(ctx, g.k);
  rank_dev.upload(ctx, rank.data(), g.k);
  Buffer<double> reach_dev(ctx, labels), bound_dev(ctx, labels);
  reach_dev.upload(ctx, reach.data(), labels);
  bound_dev.upload(ctx, bound.data(), labels);
  Buffer<long long> cs_dev(ctx, cs.size());
  cs_dev.upload(ctx, cs.data(), cs.size());
  // cells: a stable multi-key sort by (rank, x, y, z) cell, least significant key first
  Buffer<long long> kl(ctx, g.n), kx(ctx, g.n), ky(ctx, g.n), kz(ctx, g.n), key(ctx, g.n), key_sorted(ctx, g.n);
  Buffer<int> order(ctx, g.n), order_sorted(ctx, g.n), member(ctx, g.n);
  join_cells<<<grid(g.n), kThreads, 0, s>>>(g.vlabel.data(), rank_dev.data(), bound_dev.data(), p.vertices.data(), g.n, kl.data(),
                                            kx.data(), ky.data(), kz.data(), order.data(), member.data());
  BROOK_LAUNCHED();
  Buffer<unsigned char> scratch;
  for (const Buffer<long long> *k : {&kz, &ky, &kx, &kl}) {
    permute_key<<<grid(g.n), kThreads, 0, s>>>(k->data(), order.data(), g.n, key.data());
    BROOK_LAUNCHED();
    with_scratch(ctx, scratch, [&](void *t, size_t &b) {
      return hipcub::DeviceRadixSort::SortPairs(t, b, key.data(), key_sorted.data(), order.data(), order_sorted.data(), g.n, 0, 64, s);
    });
    std::swap(order, order_sorted);
  }
  Cells cells{order.data(), kl.data(), kx.data(), ky.data(), kz.data(), g.n};
  Buffer<long long> counts(ctx, g.n), offs(ctx, size_t(g.n) + 1);
  counts.clear(ctx);
  join_candidates<<<grid(g.n), kThreads, 0, s>>>(g.n, 0, cells, member.data(), g.vlabel.data(), rank_dev.data(), bound_dev.data(),
                                                 p.vertices.data(), g.root.data(), counts.data(), nullptr, nullptr, nullptr,
                                                 nullptr, nullptr, nullptr);
  BROOK_LAUNCHED();
  {
    Buffer<long long> padded(ctx, size_t(g.n) + 1);
    BROOK_HIP(hipMemcpyAsync(padded.data(), counts.data(), size_t(g.n) * 8, hipMemcpyDeviceToDevice, s));
    BROOK_HIP(hipMemsetAsync(padded.data() + g.n, 0, 8, s));
    exclusive_sum(ctx, padded.data(), offs.data(), g.n + 1);
  }
  const long long total = read_one(ctx, offs.data() + g.n);
  if (!total) return p;
  if (total > INT32_MAX) throw std::invalid_argument("join candidates exceed int32 capacity");
  const int nt = int(total);
  Buffer<int> ca(ctx, nt), cb(ctx, nt), cv(ctx, nt), cu(ctx, nt);
  Buffer<double> cd(ctx, nt);
  join_candidates<<<grid(g.n), kThreads, 0, s>>>(g.n, 1, cells, member.data(), g.vlabel.data(), rank_dev.data(), bound_dev.data(),
                                                 p.vertices.data(), g.root.data(), counts.data(), offs.data(), ca.data(),
                                                 cb.data(), cd.data(), cv.data(), cu.data());
  BROOK_LAUNCHED();
  // candidates by (A, B, distance, lower vertex, higher vertex); the first of each (A, B) stays
  Buffer<int> porder(ctx, nt), porder_sorted(ctx, nt);
  Buffer<unsigned long long> pkey(ctx, nt), pkey_sorted(ctx, nt);
  {
    std::vector<int> iota(nt);
    for (int i = 0; i < nt; ++i) iota[i] = i;
    porder.upload(ctx, iota.data(), nt);
  }
  for (int pass = 0; pass < 3; ++pass) {
    candidate_key<<<grid(nt), kThreads, 0, s>>>(porder.data(), nt, ca.data(), cb.data(), cd.data(), cv.data(), cu.data(), pass,
                                                pkey.data());
    BROOK_LAUNCHED();
    with_scratch(ctx, scratch, [&](void *t, size_t &b) {
      return hipcub::DeviceRadixSort::SortPairs(t, b, pkey.data(), pkey_sorted.data(), porder.data(), porder_sorted.data(), nt, 0,
                                                64, s);
    });
    std::swap(porder, porder_sorted);
  }
  Buffer<int> first(ctx, nt + 1), fpos(ctx, nt + 1);
  first_of_pair<<<grid(nt), kThreads, 0, s>>>(porder.data(), nt, ca.data(), cb.data(), first.data());
  BROOK_HIP(hipMemsetAsync(first.data() + nt, 0, 4, s));
  exclusive_sum(ctx, first.data(), fpos.data(), nt + 1);
  const int np = read_one(ctx, fpos.data() + nt);
  Buffer<int> rflag(ctx, g.n + 1), rpos(ctx, g.n + 1), local(ctx, g.n);
  root_flags<<<grid(g.n), kThreads, 0, s>>>(g.root.data(), g.vlabel.data(), rank_dev.data(), g.n, rflag.data());
  BROOK_HIP(hipMemsetAsync(rflag.data() + g.n, 0, 4, s));
  exclusive_sum(ctx, rflag.data(), rpos.data(), g.n + 1);
  root_positions<<<grid(g.n), kThreads, 0, s>>>(rflag.data(), rpos.data(), g.vlabel.data(), rank_dev.data(), cs_dev.data(), g.n,
                                                local.data());
  Buffer<int> px(ctx, np), py(ctx, np), pa(ctx, np), pb(ctx, np);
  Buffer<double> pd(ctx, np);
  compact_pairs<<<grid(nt), kThreads, 0, s>>>(porder.data(), first.data(), fpos.data(), nt, ca.data(), cb.data(), cd.data(), cv.data(),
                                              cu.data(), local.data(), px.data(), py.data(), pd.data(), pa.data(), pb.data());
  Buffer<long long> ps(ctx, size_t(labels) + 1);
  pair_label_offsets<<<grid(labels + 1), kThreads, 0, s>>>(pa.data(), np, g.vlabel.data(), rank_dev.data(), labels, ps.data());
  BROOK_LAUNCHED();
  Buffer<unsigned char> alive(ctx, np);
  Buffer<int> pos(ctx, roots), slot(ctx, roots), out_a(ctx, roots), out_b(ctx, roots), made(ctx, labels);
  join_fuse<<<unsigned((labels + 63) / 64), 64, 0, s>>>(labels, ps.data(), cs_dev.data(), reach_dev.data(), px.data(), py.data(),
                                                       pd.data(), pa.data(), pb.data(), p.radii.data(), alive.data(), pos.data(),
                                                       slot.data(), out_a.data(), out_b.data(), made.data());
  BROOK_LAUNCHED();
  const std::vector<int> mh = made.download(ctx);
  if (*std::max_element(mh.begin(), mh.end()) == 0) return p;  // no join: the edges keep their order
  const int count = g.m + roots;
  Buffer<unsigned long long> keys(ctx, count);
  join_keys<<<grid(std::max(g.m, roots)), kThreads, 0, s>>>(g.e.data(), g.m, out_a.data(), out_b.data(), cs_dev.data(), made.data(),
                                                           labels, roots, keys.data());
  BROOK_LAUNCHED();
  return edges_from_keys(ctx, std::move(p), g, keys, count);
}

}  // namespace

DevicePacked upload_packed(Context &ctx, const PackedView &view) {
  DevicePacked p;
  p.labels.assign(view.labels, view.labels + view.skeletons);
  p.voff.assign(view.v_off, view.v_off + view.skeletons + 1);
  p.eoff.assign(view.e_off, view.e_off + view.skeletons + 1);
  p.anisotropy = view.anisotropy;
  const size_t n = size_t(std::max<int64_t>(p.n(), 0)), m = size_t(std::max<int64_t>(p.m(), 0));
  p.vertices = Buffer<float>(ctx, std::max<size_t>(3 * n, 1));
  p.radii = Buffer<float>(ctx, std::max<size_t>(n, 1));
  p.edges = Buffer<uint32_t>(ctx, std::max<size_t>(2 * m, 1));
  auto s = ctx.stream();
  if (n) {
    BROOK_HIP(hipMemcpyAsync(p.vertices.data(), view.vertices, 12 * n, hipMemcpyDefault, s));
    BROOK_HIP(hipMemcpyAsync(p.radii.data(), view.radii, 4 * n, hipMemcpyDefault, s));
  }
  if (m) BROOK_HIP(hipMemcpyAsync(p.edges.data(), view.edges, 8 * m, hipMemcpyDefault, s));
  ctx.synchronize();
  return p;
}

Packed download_packed(Context &ctx, const DevicePacked &p) {
  Packed out;
  out.labels = p.labels;
  out.v_off.assign(p.voff.begin(), p.voff.end());
  out.e_off.assign(p.eoff.begin(), p.eoff.end());
  out.anisotropy = p.anisotropy;
  out.vertices = p.vertices.download(ctx, size_t(3 * p.n()));
  out.radii = p.radii.download(ctx, size_t(p.n()));
  out.edges = p.edges.download(ctx, size_t(2 * p.m()));
  return out;
}

DevicePacked postprocess(Context &ctx, DevicePacked p, PostStep step, double a, double b) {
  switch (step) {
    case PostStep::dust: return remove_dust(ctx, std::move(p), a);
    case PostStep::loops: return remove_loops(ctx, std::move(p));
    case PostStep::join: return join_components(ctx, std::move(p));
    case PostStep::ticks: return remove_ticks(ctx, std::move(p), a);
    case PostStep::prune: return prune_unused(ctx, std::move(p));
    case PostStep::all: break;
  }
  p = remove_dust(ctx, std::move(p), a);
  p = remove_loops(ctx, std::move(p));
  p = join_components(ctx, std::move(p));
  p = remove_ticks(ctx, std::move(p), b);
  return prune_unused(ctx, std::move(p));
}

std::vector<int64_t> skeleton_components(Context &ctx, const DevicePacked &p) {
  Graph g = build_graph(ctx, p);
  std::vector<int> r = g.root.download(ctx, size_t(g.n));
  return std::vector<int64_t>(r.begin(), r.end());
}

}  // namespace brook
