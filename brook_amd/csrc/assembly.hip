// On-device assembly of the traced paths (assembly.cu:20-160 with the same-label merge of
// packed_merge.cu:88-128 folded in): vertices deduplicated per output group and ordered by
// (x, y, z); edges (min, max) between consecutive path vertices, sorted and unique; radius = DBF.
// A group is one input label (its components merged), groups numbered in order of their first
// component id that has a skeleton.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "assembly.hpp"
#include "fields.hpp"

namespace brook {
namespace {

constexpr int kGroupShift = 31;

// One wave per record: vertex keys (group << 31 | C-order voxel index) of records with >= 2 vertices.
template <class CC>
__global__ void vertex_keys(const int *store, const int *rec_start, const int *rec_len, int records, const CC *cc,
                            const int *group_of, Dims d, unsigned long long *keys) {
  int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= records) return;
  int s = rec_start[r], n = rec_len[r];
  for (int j = lane; j < n; j += 64) {
    int v = store[s + j];
    unsigned long long key = kMaxKey;
    if (n >= 2) {
      int x, y, z;
      unravel(v, d, x, y, z);
      long long c = ((long long)x * d.sy + y) * d.sz + z;
      key = ((unsigned long long)group_of[(unsigned long long)cc[v]] << kGroupShift) | (unsigned long long)c;
    }
    keys[s + j] = key;
  }
}

__global__ void unique_flags(const unsigned long long *sorted, long long n, int *flags) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  unsigned long long k = sorted[i];
  flags[i] = (k != kMaxKey && (i == 0 || k != sorted[i - 1])) ? 1 : 0;
}

__global__ void compact(const unsigned long long *sorted, const int *flags, const int *pos, long long n,
                        unsigned long long *out) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n && flags[i]) out[pos[i]] = sorted[i];
}

__device__ __forceinline__ int lower_bound(const unsigned long long *a, int n, unsigned long long key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Edge keys (min rank << 32 | max rank) between consecutive vertices of each record.
template <class CC>
__global__ void edge_keys(const int *store, const int *rec_start, const int *rec_len, int records, const CC *cc,
                          const int *group_of, Dims d, const unsigned long long *verts, int nverts,
                          unsigned long long *edges) {
  int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= records) return;
  int s = rec_start[r], n = rec_len[r];
  for (int j = lane; j < n; j += 64) {
    unsigned long long key = kMaxKey;
    if (j + 1 < n) {
      int rank[2];
      for (int e = 0; e < 2; ++e) {
        int v = store[s + j + e], x, y, z;
        unravel(v, d, x, y, z);
        long long c = ((long long)x * d.sy + y) * d.sz + z;
        unsigned long long vk = ((unsigned long long)group_of[(unsigned long long)cc[v]] << kGroupShift) | (unsigned long long)c;
        rank[e] = lower_bound(verts, nverts, vk);
      }
      if (rank[0] != rank[1])
        key = ((unsigned long long)min(rank[0], rank[1]) << 32) | (unsigned long long)max(rank[0], rank[1]);
    }
    edges[s + j] = key;
  }
}

__global__ void group_offsets(const unsigned long long *verts, int nverts, const unsigned long long *edges, int nedges,
                              int groups, long long *v_off, long long *e_off) {
  int g = blockIdx.x * 256 + threadIdx.x;
  if (g > groups) return;
  int vo = lower_bound(verts, nverts, (unsigned long long)g << kGroupShift);
  v_off[g] = vo;
  e_off[g] = lower_bound(edges, nedges, (unsigned long long)vo << 32);
}

__global__ void emit_vertices(const unsigned long long *verts, int nverts, Dims d, const float *dbf, float ax, float ay,
                              float az, int ox, int oy, int oz, float *vout, float *rout) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nverts) return;
  long long c = (long long)(verts[i] & ((1ULL << kGroupShift) - 1));
  long long syz = (long long)d.sy * d.sz;
  int x = int(c / syz), y = int((c / d.sz) % d.sy), z = int(c % d.sz);
  vout[3LL * i] = float(x + ox) * ax;
  vout[3LL * i + 1] = float(y + oy) * ay;
  vout[3LL * i + 2] = float(z + oz) * az;
  rout[i] = dbf[flat(x, y, z, d)];
}

__global__ void emit_edges(const unsigned long long *edges, int nedges, const unsigned long long *verts,
                           const long long *v_off, unsigned *eout) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nedges) return;
  unsigned a = unsigned(edges[i] >> 32), b = unsigned(edges[i]);
  long long base = v_off[verts[a] >> kGroupShift];
  eout[2LL * i] = unsigned(a - base);
  eout[2LL * i + 1] = unsigned(b - base);
}

unsigned g256(long long n) { return unsigned((n + 255) / 256); }

// sort + unique of `n` keys in place; returns the unique count (unique keys at the front).
int sort_unique(Context &ctx, Buffer<unsigned long long> &keys, long long n, int end_bit) {
  if (n == 0) return 0;
  Buffer<unsigned long long> scratch(ctx, n);
  sort_keys(ctx, keys.data(), scratch.data(), n, end_bit);
  Buffer<int> flags(ctx, n), pos(ctx, n);
  unique_flags<<<g256(n), 256, 0, ctx.stream()>>>(keys.data(), n, flags.data());
  BROOK_LAUNCHED();
  exclusive_sum(ctx, flags.data(), pos.data(), n);
  compact<<<g256(n), 256, 0, ctx.stream()>>>(keys.data(), flags.data(), pos.data(), n, scratch.data());
  BROOK_LAUNCHED();
  int tail[2] = {0, 0};
  BROOK_HIP(hipMemcpyAsync(tail, pos.data() + n - 1, 4, hipMemcpyDeviceToHost, ctx.stream()));
  BROOK_HIP(hipMemcpyAsync(tail + 1, flags.data() + n - 1, 4, hipMemcpyDeviceToHost, ctx.stream()));
  ctx.synchronize();
  int count = tail[0] + tail[1];
  if (count) BROOK_HIP(hipMemcpyAsync(keys.data(), scratch.data(), size_t(count) * 8, hipMemcpyDeviceToDevice, ctx.stream()));
  return count;
}

template <class CC>
Packed run_assembly(Context &ctx, const AssemblyInput &in) {
  Packed out;
  const long long n = in.store_used;
  const int groups = int(in.group_labels.size());
  out.labels = in.group_labels;
  out.anisotropy = {in.aniso[0], in.aniso[1], in.aniso[2]};
  if (n == 0 || groups == 0) {
    out.v_off.assign(groups + 1, 0);
    out.e_off.assign(groups + 1, 0);
    return out;
  }
  Buffer<int> group_of(ctx, in.group_of.size());
  group_of.upload(ctx, in.group_of.data(), in.group_of.size());
  const CC *cc = static_cast<const CC *>(in.cc);
  Buffer<unsigned long long> vkeys(ctx, n), ekeys(ctx, n);
  const unsigned rb = unsigned((in.records + 3) / 4);
  vertex_keys<CC><<<rb, 256, 0, ctx.stream()>>>(in.store, in.rec_start, in.rec_len, in.records, cc, group_of.data(), in.d,
                                                vkeys.data());
  BROOK_LAUNCHED();
  int gbits = 1;
  while ((1LL << gbits) <= groups) ++gbits;
  const int nv = sort_unique(ctx, vkeys, n, kGroupShift + gbits);
  edge_keys<CC><<<rb, 256, 0, ctx.stream()>>>(in.store, in.rec_start, in.rec_len, in.records, cc, group_of.data(), in.d,
                                              vkeys.data(), nv, ekeys.data());
  BROOK_LAUNCHED();
  int vbits = 1;
  while ((1LL << vbits) <= nv) ++vbits;
  const int ne = sort_unique(ctx, ekeys, n, 32 + vbits);
  Buffer<long long> voff(ctx, groups + 1), eoff(ctx, groups + 1);
  group_offsets<<<g256(groups + 1), 256, 0, ctx.stream()>>>(vkeys.data(), nv, ekeys.data(), ne, groups, voff.data(),
                                                            eoff.data());
  BROOK_LAUNCHED();
  Buffer<float> verts(ctx, 3 * size_t(std::max(nv, 1))), radii(ctx, std::max(nv, 1));
  Buffer<unsigned> edges(ctx, 2 * size_t(std::max(ne, 1)));
  emit_vertices<<<g256(nv), 256, 0, ctx.stream()>>>(vkeys.data(), nv, in.d, in.dbf, in.aniso[0], in.aniso[1], in.aniso[2],
                                                    in.origin[0], in.origin[1], in.origin[2], verts.data(), radii.data());
  BROOK_LAUNCHED();
  emit_edges<<<g256(ne), 256, 0, ctx.stream()>>>(ekeys.data(), ne, vkeys.data(), voff.data(), edges.data());
  BROOK_LAUNCHED();
  out.v_off = voff.download(ctx);
  out.e_off = eoff.download(ctx);
  out.vertices = verts.download(ctx, 3 * size_t(nv));
  out.radii = radii.download(ctx, nv);
  out.edges = edges.download(ctx, 2 * size_t(ne));
  return out;
}

}  // namespace

Packed assemble(Context &ctx, int cc_bytes, const AssemblyInput &in) {
  if (cc_bytes == 1) return run_assembly<uint8_t>(ctx, in);
  if (cc_bytes == 2) return run_assembly<uint16_t>(ctx, in);
  return run_assembly<uint32_t>(ctx, in);
}

}  // namespace brook
