// Merging packed skeletons (packed_merge.cu:88-166, merge_fragments packed.cpp:78-86): the
// skeletons of several packed results, each fragment's vertices offset by its origin, become one
// packed result. One vertex per (label, x, y, z) -- the first occurrence in input order keeps its
// radius --, edges remapped, self-loops and duplicates dropped, vertices on no edge dropped;
// labels ascending, a label whose skeletons have vertices but no edge left is kept with none.
//
// Device plan, with one host synchronisation before the result is read back:
//  1. pack_vertices: the shifted position of every vertex and a 128-bit key {label rank, x, y, z},
//     the coordinates as order-preserving unsigned images (-0 with +0, every NaN as ~0);
//  2. one stable radix sort of (key, input index) over the key's used bits;
//  3. run heads (a NaN coordinate equals nothing), an inclusive scan: merged vertex ids;
//  4. edge keys (min id, max id), self-loops as ~0, sorted and made unique;
//  5. vertices some edge uses, compacted in key order, per-label counts from wave-aggregated runs,
//     and edge endpoints local to their label.
// Every size after step 1 is bounded by the input counts, so the buffers take those bounds and
// the kernels read the data-dependent counts from device memory.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <map>

#include "common.hpp"
#include "merge.hpp"

namespace brook {
namespace {

constexpr int kThreads = 256;
unsigned blocks_for(long long n) { return unsigned(std::max<long long>((n + kThreads - 1) / kThreads, 1)); }

// Skeleton parts (one label's skeleton in one fragment, with vertices), in input order; vertex i
// belongs to the last part whose vertex start is <= i, edge j likewise.
struct Parts {
  const long long *vertex_start, *edge_start;
  const int *vertex_count, *rank;
  const float *origin;  // 3 per part
  int count;
};

__device__ int part_of(const long long *start, int count, long long i) {
  int lo = 0, hi = count;  // start[lo] <= i < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Order-preserving unsigned image of a float: negative values flip every bit, the others set the
// sign bit; -0 maps with +0 and every NaN to ~0 (above +inf, whose image is 0xFF800000).
__device__ unsigned ordered_bits(float v) {
  if (v != v) return ~0u;
  const unsigned b = __float_as_uint(v + 0.0f);  // -0 + 0 = +0
  return (b >> 31) ? ~b : (b | 0x80000000u);
}

struct VertexKey {
  unsigned rank, x, y, z;
};
// Significance for the radix sort: label rank, then x, y, z.
struct KeyFields {
  __host__ __device__ ::rocprim::tuple<unsigned &, unsigned &, unsigned &, unsigned &> operator()(VertexKey &k) const {
    return ::rocprim::tuple<unsigned &, unsigned &, unsigned &, unsigned &>(k.rank, k.x, k.y, k.z);
  }
};

__global__ void pack_vertices(Parts parts, long long nv, const float *in, float *shifted, VertexKey *keys, int *index) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nv) return;
  const int p = part_of(parts.vertex_start, parts.count, i);
  float c[3];
  for (int a = 0; a < 3; ++a) {
    c[a] = in[3 * i + a] + parts.origin[3 * p + a];
    shifted[3 * i + a] = c[a];
  }
  keys[i] = VertexKey{unsigned(parts.rank[p]), ordered_bits(c[0]), ordered_bits(c[1]), ordered_bits(c[2])};
  index[i] = int(i);
}

__global__ void run_heads(const VertexKey *keys, long long nv, int *head) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nv) return;
  const VertexKey k = keys[i];
  int fresh = i == 0 || k.x == ~0u || k.y == ~0u || k.z == ~0u;
  if (!fresh) {
    const VertexKey q = keys[i - 1];
    fresh = q.rank != k.rank || q.x != k.x || q.y != k.y || q.z != k.z;
  }
  head[i] = fresh;
}

// merged[v] = merged id of input vertex v; first[id] = its first input vertex; rank_of[id].
__global__ void number_vertices(const VertexKey *keys, const int *index, const int *head, const int *ids, long long nv,
                                int *merged, int *first, int *rank_of) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nv) return;
  const int id = ids[i] - 1, v = index[i];
  merged[v] = id;
  if (head[i]) {
    first[id] = v;
    rank_of[id] = int(keys[i].rank);
  }
}

// Edge j as (min id << 32 | max id); self-loops become ~0 (they sort last), endpoints outside
// their skeleton raise the error flag and drop the edge.
__global__ void edge_keys(Parts parts, long long ne, const unsigned *edges, const int *merged, unsigned long long *keys,
                          int *error) {
  const long long j = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (j >= ne) return;
  const int p = part_of(parts.edge_start, parts.count, j);
  const unsigned u = edges[2 * j], w = edges[2 * j + 1], n = unsigned(parts.vertex_count[p]);
  unsigned long long key = ~0ULL;
  if (u >= n || w >= n) {
    g_st(error, 1);
  } else {
    const unsigned a = unsigned(merged[parts.vertex_start[p] + u]), b = unsigned(merged[parts.vertex_start[p] + w]);
    if (a != b) key = (((unsigned long long)min(a, b)) << 32) | max(a, b);
  }
  keys[j] = key;
}

// The unique edges without the trailing ~0 (at most one after Unique).
__global__ void count_edges(const unsigned long long *unique_keys, const int *selected, int *count) {
  int n = *selected;
  if (n > 0 && unique_keys[n - 1] == ~0ULL) --n;
  *count = n;
}

__global__ void mark_used(const unsigned long long *keys, const int *count, int *used) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= *count) return;
  g_st(used + (keys[j] >> 32), 1);
  g_st(used + unsigned(keys[j]), 1);
}

// Adds a wave's flagged lanes to counts[rank] once per run of equal ranks (ranks are sorted, so a
// wave holds few runs). Every lane of the wave calls it.
__device__ void count_by_rank(int rank, bool flagged, int *counts) {
  const int lane = threadIdx.x & 63;
  const WaveRun run = wave_run(rank, true, 1, lane);
  const unsigned long long mine = __ballot(flagged);
  if (!run.head) return;
  const unsigned long long span = lanes_upto(lane + run.len - 1) & ~(lane ? lanes_upto(lane - 1) : 0ULL);
  const int n = __popcll(mine & span);
  if (n) g_add(counts + rank, n);
}

__global__ void emit_vertices(const float *shifted, const float *radii, const int *first, const int *rank_of, const int *used,
                              const int *slot, long long bound, float *out, float *out_radii, int *counts) {
  const long long id = (long long)blockIdx.x * kThreads + threadIdx.x;
  const bool live = id < bound && used[id];
  const int rank = live ? rank_of[id] : -1 - int(threadIdx.x & 63);  // dead lanes: runs of their own
  if (live) {
    const long long v = first[id], dest = slot[id] - 1;
    for (int a = 0; a < 3; ++a) out[3 * dest + a] = shifted[3 * v + a];
    out_radii[dest] = radii[v];
  }
  count_by_rank(rank, live, counts);
}

__global__ void count_edge_ranks(const unsigned long long *keys, const int *count, const int *rank_of, int *counts) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  const bool live = j < *count;
  const int rank = live ? rank_of[keys[j] >> 32] : -1 - int(threadIdx.x & 63);
  count_by_rank(rank, live, counts);
}

__global__ void emit_edges(const unsigned long long *keys, const int *count, const int *rank_of, const int *slot,
                           const int *vertex_base, unsigned *out) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= *count) return;
  const unsigned a = unsigned(keys[j] >> 32), b = unsigned(keys[j]);
  const int base = vertex_base[rank_of[a]];
  out[2LL * j] = unsigned(slot[a] - 1 - base);
  out[2LL * j + 1] = unsigned(slot[b] - 1 - base);
}

// Runs a hipcub primitive twice: once for its scratch size, once with the (grown) scratch.
template <class Op> void with_scratch(Context &ctx, Buffer<unsigned char> &scratch, Op &&op) {
  size_t bytes = 0;
  BROOK_HIP(op(nullptr, bytes));
  if (bytes > scratch.size()) scratch = Buffer<unsigned char>(ctx, bytes);
  BROOK_HIP(op(scratch.data(), bytes));
}

int bits_for(unsigned long long max_value) {
  int b = 0;
  while (b < 64 && (max_value >> b)) ++b;
  return b;
}

// Copies `count` elements that may live in host or device memory into a device buffer.
template <class T> void gather(Context &ctx, T *dest, const T *src, size_t count) {
  if (count) BROOK_HIP(hipMemcpyAsync(dest, src, count * sizeof(T), hipMemcpyDefault, ctx.stream()));
}

}  // namespace

Packed merge_packed(Context &ctx, const std::vector<PackedView> &fragments, const std::vector<std::array<float, 3>> &origins) {
  if (fragments.size() != origins.size()) throw std::invalid_argument("fragment/origin count mismatch");
  Packed out;
  if (!fragments.empty()) out.anisotropy = fragments.front().anisotropy;
  // Labels of the skeletons with vertices, ascending (packed_merge.cu:91-98).
  std::map<int64_t, int> rank;
  for (const PackedView &f : fragments) {
    if (f.v_off[0] != 0 || f.e_off[0] != 0) throw std::invalid_argument("invalid packed offsets");
    for (size_t k = 0; k < f.skeletons; ++k) {
      const long long nv = f.v_off[k + 1] - f.v_off[k], ne = f.e_off[k + 1] - f.e_off[k];
      if (nv < 0 || ne < 0) throw std::invalid_argument("invalid packed offsets");
      if (!nv && ne) throw std::invalid_argument("edges without vertices");
      if (nv) rank.emplace(f.labels[k], 0);
    }
  }
  for (auto &entry : rank) {
    entry.second = int(out.labels.size());
    out.labels.push_back(entry.first);
  }
  if (out.labels.empty()) return out;
  // The part table; each fragment's arrays are copied whole (skeletons without vertices hold none).
  std::vector<long long> vstart, estart, fragment_v, fragment_e;
  std::vector<int> vcount, prank;
  std::vector<float> porigin;
  long long nv = 0, ne = 0;
  for (size_t fi = 0; fi < fragments.size(); ++fi) {
    const PackedView &f = fragments[fi];
    fragment_v.push_back(nv);
    fragment_e.push_back(ne);
    for (size_t k = 0; k < f.skeletons; ++k) {
      const long long n = f.v_off[k + 1] - f.v_off[k];
      if (!n) continue;
      vstart.push_back(nv + f.v_off[k]);
      estart.push_back(ne + f.e_off[k]);
      vcount.push_back(int(std::min<long long>(n, INT32_MAX)));
      prank.push_back(rank.at(f.labels[k]));
      porigin.insert(porigin.end(), origins[fi].begin(), origins[fi].end());
    }
    nv += f.v_off[f.skeletons];
    ne += f.e_off[f.skeletons];
    if (nv > INT32_MAX - 1 || ne > INT32_MAX - 1) throw std::invalid_argument("packed merge exceeds int32 work capacity");
  }
  const int np = int(vstart.size()), ngroups = int(out.labels.size());
  auto s = ctx.stream();
  Buffer<long long> d_vstart(ctx, size_t(np)), d_estart(ctx, size_t(np));
  Buffer<int> d_vcount(ctx, size_t(np)), d_rank(ctx, size_t(np));
  Buffer<float> d_origin(ctx, size_t(3 * np));
  d_vstart.upload(ctx, vstart.data(), size_t(np));
  d_estart.upload(ctx, estart.data(), size_t(np));
  d_vcount.upload(ctx, vcount.data(), size_t(np));
  d_rank.upload(ctx, prank.data(), size_t(np));
  d_origin.upload(ctx, porigin.data(), size_t(3 * np));
  const Parts parts{d_vstart.data(), d_estart.data(), d_vcount.data(), d_rank.data(), d_origin.data(), np};

  Buffer<float> in(ctx, size_t(3 * nv)), shifted(ctx, size_t(3 * nv)), radii(ctx, size_t(nv));
  Buffer<unsigned> edges(ctx, size_t(std::max<long long>(2 * ne, 1)));
  for (size_t fi = 0; fi < fragments.size(); ++fi) {
    const PackedView &f = fragments[fi];
    const size_t fv = size_t(f.v_off[f.skeletons]), fe = size_t(f.e_off[f.skeletons]);
    gather(ctx, in.data() + 3 * fragment_v[fi], f.vertices, 3 * fv);
    gather(ctx, radii.data() + fragment_v[fi], f.radii, fv);
    gather(ctx, edges.data() + 2 * fragment_e[fi], f.edges, 2 * fe);
  }
  Buffer<unsigned char> scratch(ctx, 1);
  // 1-3: merged vertex ids
  Buffer<VertexKey> keys_in(ctx, size_t(nv)), keys(ctx, size_t(nv));
  Buffer<int> index_in(ctx, size_t(nv)), index(ctx, size_t(nv)), head(ctx, size_t(nv)), ids(ctx, size_t(nv)),
      merged(ctx, size_t(nv)), first(ctx, size_t(nv)), rank_of(ctx, size_t(nv));
  pack_vertices<<<blocks_for(nv), kThreads, 0, s>>>(parts, nv, in.data(), shifted.data(), keys_in.data(), index_in.data());
  BROOK_LAUNCHED();
  const int key_bits = 96 + bits_for(unsigned(ngroups - 1));
  with_scratch(ctx, scratch, [&](void *p, size_t &bytes) {
    return hipcub::DeviceRadixSort::SortPairs(p, bytes, keys_in.data(), keys.data(), index_in.data(), index.data(), int(nv),
                                              KeyFields{}, 0, key_bits, s);
  });
  run_heads<<<blocks_for(nv), kThreads, 0, s>>>(keys.data(), nv, head.data());
  BROOK_LAUNCHED();
  with_scratch(ctx, scratch, [&](void *p, size_t &bytes) {
    return hipcub::DeviceScan::InclusiveSum(p, bytes, head.data(), ids.data(), int(nv), s);
  });
  number_vertices<<<blocks_for(nv), kThreads, 0, s>>>(keys.data(), index.data(), head.data(), ids.data(), nv, merged.data(),
                                                      first.data(), rank_of.data());
  BROOK_LAUNCHED();
  // 4: unique edges
  Buffer<unsigned long long> ekeys_in(ctx, size_t(std::max<long long>(ne, 1))), ekeys(ctx, size_t(std::max<long long>(ne, 1))),
      unique_keys(ctx, size_t(std::max<long long>(ne, 1)));
  Buffer<int> flags(ctx, 3);  // error, selected, edge count
  flags.clear(ctx);
  int *error = flags.data(), *selected = flags.data() + 1, *nedges = flags.data() + 2;
  if (ne) {
    edge_keys<<<blocks_for(ne), kThreads, 0, s>>>(parts, ne, edges.data(), merged.data(), ekeys_in.data(), error);
    BROOK_LAUNCHED();
    const int edge_bits = std::max(32 + bits_for(unsigned(nv - 1)), 33);
    with_scratch(ctx, scratch, [&](void *p, size_t &bytes) {
      return hipcub::DeviceRadixSort::SortKeys(p, bytes, ekeys_in.data(), ekeys.data(), int(ne), 0, edge_bits, s);
    });
    with_scratch(ctx, scratch, [&](void *p, size_t &bytes) {
      return hipcub::DeviceSelect::Unique(p, bytes, ekeys.data(), unique_keys.data(), selected, int(ne), s);
    });
    count_edges<<<1, 1, 0, s>>>(unique_keys.data(), selected, nedges);
    BROOK_LAUNCHED();
  }
  // 5: kept vertices and the output
  Buffer<int> used(ctx, size_t(nv)), slot(ctx, size_t(nv)), counts(ctx, size_t(3 * ngroups));
  used.clear(ctx);
  counts.clear(ctx);
  int *vcounts = counts.data(), *ecounts = counts.data() + ngroups, *vbase = counts.data() + 2 * ngroups;
  if (ne) {
    mark_used<<<blocks_for(ne), kThreads, 0, s>>>(unique_keys.data(), nedges, used.data());
    BROOK_LAUNCHED();
  }
  with_scratch(ctx, scratch, [&](void *p, size_t &bytes) {
    return hipcub::DeviceScan::InclusiveSum(p, bytes, used.data(), slot.data(), int(nv), s);
  });
  Buffer<float> vout(ctx, size_t(3 * nv)), rout(ctx, size_t(nv));
  emit_vertices<<<blocks_for(nv), kThreads, 0, s>>>(shifted.data(), radii.data(), first.data(), rank_of.data(), used.data(),
                                                    slot.data(), nv, vout.data(), rout.data(), vcounts);
  BROOK_LAUNCHED();
  with_scratch(ctx, scratch, [&](void *p, size_t &bytes) {
    return hipcub::DeviceScan::ExclusiveSum(p, bytes, vcounts, vbase, ngroups, s);
  });
  Buffer<unsigned> eout(ctx, size_t(std::max<long long>(2 * ne, 1)));
  if (ne) {
    count_edge_ranks<<<blocks_for(ne), kThreads, 0, s>>>(unique_keys.data(), nedges, rank_of.data(), ecounts);
    BROOK_LAUNCHED();
    emit_edges<<<blocks_for(ne), kThreads, 0, s>>>(unique_keys.data(), nedges, rank_of.data(), slot.data(), vbase, eout.data());
    BROOK_LAUNCHED();
  }
  const std::vector<int> status = flags.download(ctx);  // synchronises
  if (status[0]) throw std::invalid_argument("packed edge is out of bounds");
  const std::vector<int> c = counts.download(ctx, size_t(2 * ngroups));
  long long kept = 0, kept_edges = 0;
  for (int g = 0; g < ngroups; ++g) {
    kept += c[g];
    kept_edges += c[ngroups + g];
    out.v_off.push_back(kept);
    out.e_off.push_back(kept_edges);
  }
  if (kept_edges != status[2]) throw std::logic_error("packed merge: edge counts disagree");
  out.vertices = vout.download(ctx, size_t(3 * kept));
  out.radii = rout.download(ctx, size_t(kept));
  out.edges = eout.download(ctx, size_t(2 * kept_edges));
  return out;
}

}  // namespace brook
