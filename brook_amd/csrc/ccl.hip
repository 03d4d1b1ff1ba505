// Label-aware 26-connected components on gfx950 (ids in ascending order of each component's
// minimum flat index, as ccl.cu:24-82 assigns them), plus the per-id bbox/count analysis
// (preamble.cu:10-29) and face extraction for the border planes.
//
// Union-find over int32 parents: every x-run is hooked to its first voxel by one wave ballot
// (no atomics), the 12 backward neighbours in the y-1 / z-1 rows are united with CAS hooking
// (larger root -> smaller), so each root is its component's minimum flat index. Ids come from a
// tiled exclusive scan of the root flags in flat order -- no host sort.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "stages.hpp"

namespace brook {
namespace {

constexpr int kBG = INT32_MIN;  // parent of a background voxel
constexpr int kTile = 4096;     // voxels per root-count tile (256 threads x 16)

template <class T>
__global__ void __launch_bounds__(256) ccl_init(const T *lab, int *parent, Dims d, int n) {
  int i = blockIdx.x * 256 + threadIdx.x;
  int lane = threadIdx.x & 63;
  bool valid = i < n;
  T l = valid ? lab[i] : T(0);
  int x = valid ? i % d.sx : 0;
  T prev = shfl_up_any(l, 1);
  if (lane == 0 && valid && x > 0) prev = lab[i - 1];
  bool start = valid && (x == 0 || prev != l);
  unsigned long long mask = __ballot(start);
  if (!valid) return;
  if (l == T(0)) { parent[i] = kBG; return; }
  unsigned long long below = mask & lanes_upto(lane);
  int base = i - lane;
  parent[i] = below ? base + (63 - __clzll((long long)below)) : base - 1;
}

__device__ __forceinline__ int find_root(int *parent, int v) {
  int cur = ld_i(parent + v);
  if (cur != v) {
    int prev = v, next;
    while (cur > (next = ld_i(parent + cur))) {
      st_i(parent + prev, next);  // racy path halving: only ever points closer to the root
      prev = cur;
      cur = next;
    }
  }
  return cur;
}

__device__ __forceinline__ void unite(int *parent, int a, int b) {
  int va = find_root(parent, a), vb = find_root(parent, b);
  while (va != vb) {
    if (va < vb) {
      int ret = atomicCAS(parent + vb, vb, va);
      if (ret == vb) return;
      vb = find_root(parent, ret);
    } else {
      int ret = atomicCAS(parent + va, va, vb);
      if (ret == va) return;
      va = find_root(parent, ret);
    }
  }
}

// Unions between x-runs of equal label in the four backward rows ((y-1, z) and (y-1..y+1, z-1);
// x-neighbours in the voxel's own row are one run already). Every voxel of a run shares the run's
// root, so one union per pair of touching runs suffices: voxel x of run [c, d] unites with the
// row's run at x where that run begins at x or x == c (centre contacts), and the run ends take
// the diagonal contacts (x-1 at c, x+1 at d when the centre differs).
template <class T>
__global__ void __launch_bounds__(256) ccl_union(const T *lab, int *parent, Dims d, int n) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  T l = lab[i];
  if (l == T(0)) return;
  int x, y, z;
  unravel(i, d, x, y, z);
  const bool first = x == 0 || lab[i - 1] != l, last = x + 1 == d.sx || lab[i + 1] != l;
  const int rows[4][2] = {{-1, 0}, {-1, -1}, {0, -1}, {1, -1}};
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    int yy = y + rows[r][0], zz = z + rows[r][1];
    if (yy < 0 || yy >= d.sy || zz < 0) continue;
    int row = d.sx * (yy + d.sy * zz);
    const bool centre = lab[row + x] == l;
    const bool left = x > 0 && lab[row + x - 1] == l;
    if (centre && (first || !left)) unite(parent, i, row + x);
    if (first && left && !centre) unite(parent, i, row + x - 1);
    if (last && !centre && x + 1 < d.sx && lab[row + x + 1] == l) unite(parent, i, row + x + 1);
  }
}

__global__ void __launch_bounds__(256) ccl_count(const int *parent, int n, int *tile_counts) {
  __shared__ int total;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  int base = blockIdx.x * kTile + threadIdx.x * 16, c = 0;
#pragma unroll 4
  for (int j = 0; j < 16; ++j) {
    int i = base + j;
    if (i < n && parent[i] == i) ++c;
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(&total, c);
  __syncthreads();
  if (threadIdx.x == 0) tile_counts[blockIdx.x] = total;
}

// Exclusive scan of the tile counts in place (one workgroup); total -> *count.
__global__ void __launch_bounds__(1024) scan_tiles(int *tile_counts, int tiles, long long *count) {
  __shared__ int sums[1024];
  int per = (tiles + 1023) / 1024, b = threadIdx.x * per, s = 0;
  for (int j = 0; j < per; ++j) if (b + j < tiles) s += tile_counts[b + j];
  sums[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = threadIdx.x >= off ? sums[threadIdx.x - off] : 0;
    __syncthreads();
    sums[threadIdx.x] += v;
    __syncthreads();
  }
  int run = sums[threadIdx.x] - s;
  for (int j = 0; j < per; ++j)
    if (b + j < tiles) { int c = tile_counts[b + j]; tile_counts[b + j] = run; run += c; }
  if (threadIdx.x == 1023) *count = sums[1023];
}

template <class T>
__global__ void __launch_bounds__(256) ccl_mark(const T *lab, int *parent, int n, const int *tile_offsets,
                                                long long *roots, long long *mapping) {
  __shared__ int wave_sums[4];
  int base = blockIdx.x * kTile + threadIdx.x * 16, c = 0;
  unsigned flags = 0;
  for (int j = 0; j < 16; ++j) {
    int i = base + j;
    if (i < n && parent[i] == i) { flags |= 1u << j; ++c; }
  }
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = c;
  for (int off = 1; off < 64; off <<= 1) {
    int v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  if (lane == 63) wave_sums[wave] = incl;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += wave_sums[w];
  int id = tile_offsets[blockIdx.x] + before + incl - c;  // ids are 1-based: id + 1 below
  for (int j = 0; j < 16; ++j) {
    if (!(flags & (1u << j))) continue;
    int i = base + j;
    ++id;
    parent[i] = -id;
    roots[id - 1] = i;
    mapping[id] = (long long)lab[i];
  }
}

template <class T, class CC>
__global__ void __launch_bounds__(256) ccl_relabel(const int *parent, int n, CC *cc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int p = parent[i];
  if (p == kBG) { cc[i] = CC(0); return; }
  while (p >= 0) p = parent[p];  // roots hold -id now
  cc[i] = CC(-p);
}

// ---- analysis: per-id bbox and count ---------------------------------------------------------
// A workgroup walks kRunChunk consecutive voxels; each x-run within a wave updates its label's
// LDS slot, and the workgroup then sends one count/bbox update per label it saw. (Few labels over
// many runs otherwise serialize on the same global addresses.)
constexpr int kRunChunk = 16384;

template <class CC>
__global__ void __launch_bounds__(256) analyze_runs(const CC *cc, Dims d, int n, int *lo, int *hi,
                                                    unsigned long long *counts) {
  __shared__ LabelSlots slots;
  __shared__ unsigned cnt[kLabelSlots];
  __shared__ int box[6][kLabelSlots];  // lo x,y,z then hi x,y,z
  slots.init();
  for (int s = threadIdx.x; s < kLabelSlots; s += 256) {
    cnt[s] = 0;
    for (int a = 0; a < 3; ++a) { box[a][s] = INT32_MAX; box[3 + a][s] = -1; }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long long begin = (long long)blockIdx.x * kRunChunk;
  for (int t = 0; t < kRunChunk; t += 256) {
    const long long g = begin + t + threadIdx.x;
    const bool valid = g < n;
    const int i = valid ? int(g) : 0;
    const CC l = valid ? cc[i] : CC(0);
    const int x = valid ? i % d.sx : 0;
    const WaveRun r = wave_run(l, valid, x, lane);
    if (!valid || l == CC(0) || !r.head) continue;
    const int y = (i / d.sx) % d.sy, z = i / (d.sx * d.sy), xe = x + r.len - 1;
    const int s = slots.insert(int(l));
    if (s >= 0) {
      atomicAdd(&cnt[s], unsigned(r.len));
      atomicMin(&box[0][s], x);
      atomicMin(&box[1][s], y);
      atomicMin(&box[2][s], z);
      atomicMax(&box[3][s], xe);
      atomicMax(&box[4][s], y);
      atomicMax(&box[5][s], z);
      continue;
    }
    atomicAdd(counts + l, (unsigned long long)r.len);
    int *L = lo + 3 * int(l), *H = hi + 3 * int(l);
    atomicMin(L, x); atomicMin(L + 1, y); atomicMin(L + 2, z);
    atomicMax(H, xe); atomicMax(H + 1, y); atomicMax(H + 2, z);
  }
  __syncthreads();
  for (int s = threadIdx.x; s < kLabelSlots; s += 256) {
    const int l = slots.tag[s];
    if (l < 0) continue;
    atomicAdd(counts + l, (unsigned long long)cnt[s]);
    int *L = lo + 3 * l, *H = hi + 3 * l;
    for (int a = 0; a < 3; ++a) {
      if (ld_i(L + a) > box[a][s]) atomicMin(L + a, box[a][s]);
      if (ld_i(H + a) < box[3 + a][s]) atomicMax(H + a, box[3 + a][s]);
    }
  }
}

template <class CC>
__global__ void __launch_bounds__(256) faces_copy(const CC *cc, Dims d, FaceCanvas fc, CC *canvas) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long long)fc.rows * fc.cols) return;
  const int r = int(i % fc.rows), col = int(i / fc.rows);
  const int dims[3] = {d.sx, d.sy, d.sz};
  CC v = CC(0);
  for (int f = 0; f < 6; ++f) {
    const int x = col - fc.col0[f], y = r - fc.row0[f];
    if (x < 0 || x >= dims[fc.a[f]] || y < 0 || y >= dims[fc.b[f]]) continue;
    int p[3];
    p[fc.a[f]] = x;
    p[fc.b[f]] = y;
    p[fc.c[f]] = f % 2 ? dims[fc.c[f]] - 1 : 0;
    v = cc[flat(p[0], p[1], p[2], d)];
    break;
  }
  canvas[i] = v;
}

template <class CC>
__global__ void __launch_bounds__(256) faces_gather(const CC *ids, const float *dist, Dims d, FaceCanvas fc, int f0,
                                                   int f1, long long total, CC *face_ids, float *face_dist) {
  const long long j = (long long)blockIdx.x * 256 + threadIdx.x;
  if (j >= total) return;
  const int dims[3] = {d.sx, d.sy, d.sz};
  long long k = j;
  int f = f0;
  for (; f < f1 - 1; ++f) {
    const long long pixels = (long long)dims[fc.a[f]] * dims[fc.b[f]];
    if (k < pixels) break;
    k -= pixels;
  }
  const int fy = dims[fc.b[f]];
  const long long x = k / fy, y = k % fy;
  const long long at = ((long long)fc.col0[f] + x) * fc.rows + fc.row0[f] + y;
  face_ids[j] = ids[at];
  face_dist[j] = dist[at];
}

unsigned blocks_for(int64_t n, int per) { return unsigned((n + per - 1) / per); }

template <class T> Components run_ccl(Context &ctx, const LabelVolume &in) {
  Components out;
  const int n = int(in.size());
  const Dims d = in.dims;
  const T *lab = static_cast<const T *>(in.data);
  Buffer<int> parent(ctx, n);
  auto s = ctx.stream();
  {
    KernelTimer kt(ctx, "ccl_init");
    ccl_init<T><<<blocks_for(n, 256), 256, 0, s>>>(lab, parent.data(), d, n);
    BROOK_LAUNCHED();
  }
  {
    KernelTimer kt(ctx, "ccl_union");
    ccl_union<T><<<blocks_for(n, 256), 256, 0, s>>>(lab, parent.data(), d, n);
    BROOK_LAUNCHED();
  }
  const int tiles = int(blocks_for(n, kTile));
  Buffer<int> tile(ctx, tiles);
  Buffer<long long> total(ctx, 1);
  {
    KernelTimer kt(ctx, "ccl_count");
    ccl_count<<<tiles, 256, 0, s>>>(parent.data(), n, tile.data());
    BROOK_LAUNCHED();
  }
  scan_tiles<<<1, 1024, 0, s>>>(tile.data(), tiles, total.data());
  BROOK_LAUNCHED();
  const int64_t count = total.download(ctx, 1)[0];
  out.count = count;
  Buffer<long long> roots(ctx, std::max<int64_t>(count, 1)), mapping(ctx, count + 1);
  mapping.clear(ctx);
  {
    KernelTimer kt(ctx, "ccl_mark");
    ccl_mark<T><<<tiles, 256, 0, s>>>(lab, parent.data(), n, tile.data(), roots.data(), mapping.data());
    BROOK_LAUNCHED();
  }
  out.cc_bytes = id_bytes(count);
  out.labels = Buffer<uint8_t>(ctx, size_t(n) * out.cc_bytes);
  {
    KernelTimer kt(ctx, "ccl_relabel");
    if (out.cc_bytes == 1)
      ccl_relabel<T, uint8_t><<<blocks_for(n, 256), 256, 0, s>>>(parent.data(), n, out.labels.data());
    else if (out.cc_bytes == 2)
      ccl_relabel<T, uint16_t><<<blocks_for(n, 256), 256, 0, s>>>(parent.data(), n,
                                                                   reinterpret_cast<uint16_t *>(out.labels.data()));
    else
      ccl_relabel<T, uint32_t><<<blocks_for(n, 256), 256, 0, s>>>(parent.data(), n,
                                                                   reinterpret_cast<uint32_t *>(out.labels.data()));
    BROOK_LAUNCHED();
  }
  auto r = roots.download(ctx, count);
  auto m = mapping.download(ctx, count + 1);
  out.roots.assign(r.begin(), r.end());
  out.mapping.assign(m.begin(), m.end());
  if (in.is_signed && sizeof(T) < 8)
    for (auto &v : out.mapping) {
      const int shift = 64 - 8 * int(sizeof(T));
      v = int64_t(uint64_t(v) << shift) >> shift;  // sign-extend the label's width
    }
  return out;
}

template <class CC> void run_analyze(Context &ctx, const Components &cc, Dims d, std::vector<Box> &boxes) {
  const int n = d.sx * d.sy * d.sz;
  size_t k = size_t(cc.count) + 1;
  Buffer<int> lo(ctx, 3 * k), hi(ctx, 3 * k);
  Buffer<unsigned long long> counts(ctx, k);
  std::vector<int> l0(3 * k, INT32_MAX), h0(3 * k, -1);
  lo.upload(ctx, l0.data(), l0.size());
  hi.upload(ctx, h0.data(), h0.size());
  counts.clear(ctx);
  analyze_runs<CC><<<blocks_for(n, kRunChunk), 256, 0, ctx.stream()>>>(reinterpret_cast<const CC *>(cc.labels.data()), d, n,
                                                                  lo.data(), hi.data(), counts.data());
  BROOK_LAUNCHED();
  auto L = lo.download(ctx), H = hi.download(ctx);
  auto C = counts.download(ctx);
  boxes.assign(k, Box{});
  int64_t fg = 0;
  for (size_t i = 1; i < k; ++i) {
    for (int a = 0; a < 3; ++a) { boxes[i].lo[a] = L[3 * i + a]; boxes[i].hi[a] = H[3 * i + a]; }
    boxes[i].count = int64_t(C[i]);
    fg += boxes[i].count;
  }
  boxes[0].count = int64_t(n) - fg;
}

}  // namespace

// ---- connected components without device allocation ------------------------------------------
// The working memory of connected_components_noalloc: one static array per device (the loader
// places it in each device's memory with the code object). Host code never passes `scratch` itself
// to a kernel (on the host the name is a stand-in, not a device address): each device's address is
// looked up once, with that device current, by noalloc_scratch_init, and kept in g_scratch.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock[kMaxDevices];  // one call at a time per device: calls share the array
void *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_init;
size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

template <class T>
int64_t run_ccl_noalloc(Context &ctx, const LabelVolume &in, uint32_t *ids, int64_t *mapping, int64_t *roots,
                        int64_t capacity, unsigned char *base) {
  const int n = int(in.size());
  const Dims d = in.dims;
  const T *lab = static_cast<const T *>(in.data);
  const int tiles = int(blocks_for(n, kTile));
  int *parent = reinterpret_cast<int *>(base);
  size_t at = align256(size_t(n) * 4);
  int *tile = reinterpret_cast<int *>(base + at);
  at += align256(size_t(tiles) * 4);
  long long *total = reinterpret_cast<long long *>(base + at);
  at += 256;
  long long *roots_dev = reinterpret_cast<long long *>(base + at);
  at += align256(size_t(capacity) * 8);
  long long *mapping_dev = reinterpret_cast<long long *>(base + at);
  auto s = ctx.stream();
  ccl_init<T><<<blocks_for(n, 256), 256, 0, s>>>(lab, parent, d, n);
  BROOK_LAUNCHED();
  ccl_union<T><<<blocks_for(n, 256), 256, 0, s>>>(lab, parent, d, n);
  BROOK_LAUNCHED();
  ccl_count<<<tiles, 256, 0, s>>>(parent, n, tile);
  BROOK_LAUNCHED();
  scan_tiles<<<1, 1024, 0, s>>>(tile, tiles, total);
  BROOK_LAUNCHED();
  long long count = 0;
  BROOK_HIP(hipMemcpyAsync(&count, total, 8, hipMemcpyDeviceToHost, s));
  ctx.synchronize();
  if (count > capacity) return count;  // the caller reports it; no table is written
  ccl_mark<T><<<tiles, 256, 0, s>>>(lab, parent, n, tile, roots_dev, mapping_dev);
  BROOK_LAUNCHED();
  ccl_relabel<T, uint32_t><<<blocks_for(n, 256), 256, 0, s>>>(parent, n, ids);
  BROOK_LAUNCHED();
  if (count) {
    BROOK_HIP(hipMemcpyAsync(roots, roots_dev, size_t(count) * 8, hipMemcpyDeviceToHost, s));
    BROOK_HIP(hipMemcpyAsync(mapping + 1, mapping_dev + 1, size_t(count) * 8, hipMemcpyDeviceToHost, s));
  }
  ctx.synchronize();
  mapping[0] = 0;
  if (in.is_signed && sizeof(T) < 8)
    for (long long i = 1; i <= count; ++i) {
      const int shift = 64 - 8 * int(sizeof(T));
      mapping[i] = int64_t(uint64_t(mapping[i]) << shift) >> shift;  // sign-extend the label's width
    }
  return count;
}
}  // namespace

void noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lock(g_scratch_init);
  if (g_scratch[device]) return;
  void *p = nullptr;
  BROOK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // this device's copy (it is current)
  if (!p) throw DeviceError(hipErrorInvalidSymbol, "no device address for the no-allocation scratch");
  g_scratch[device] = p;
}

size_t noalloc_scratch_need(int64_t voxels, int64_t capacity) {
  return align256(size_t(voxels) * 4) + align256(size_t(blocks_for(voxels, kTile)) * 4) + 256 +
         align256(size_t(capacity) * 8) + size_t(capacity + 1) * 8;
}

int64_t connected_components_noalloc(Context &ctx, const LabelVolume &input, uint32_t *ids, int64_t *mapping,
                                     int64_t *roots, int64_t capacity) {
  const int device = ctx.device();
  if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index beyond the no-allocation table");
  if (!g_scratch[device]) noalloc_scratch_init(device);
  if (input.size() == 0) {
    mapping[0] = 0;
    return 0;
  }
  if (noalloc_scratch_need(input.size(), capacity) > kNoAllocScratchBytes)
    throw std::invalid_argument("volume and table capacity exceed the no-allocation scratch (" +
                                std::to_string(kNoAllocScratchBytes >> 20) +
                                " MiB: 4 bytes per voxel plus 16 per table entry)");
  std::lock_guard<std::mutex> lock(g_scratch_lock[device]);
  unsigned char *base = static_cast<unsigned char *>(g_scratch[device]);
  switch (input.bytes) {
    case 1: return run_ccl_noalloc<uint8_t>(ctx, input, ids, mapping, roots, capacity, base);
    case 2: return run_ccl_noalloc<uint16_t>(ctx, input, ids, mapping, roots, capacity, base);
    case 4: return run_ccl_noalloc<uint32_t>(ctx, input, ids, mapping, roots, capacity, base);
    case 8: return run_ccl_noalloc<unsigned long long>(ctx, input, ids, mapping, roots, capacity, base);
    default: throw std::invalid_argument("connected components requires integer labels");
  }
}

Components connected_components(Context &ctx, const LabelVolume &input) {
  if (input.size() >= (int64_t(1) << 31) - kTile) throw Unsupported("volume exceeds the int32 voxel index space");
  if (input.size() == 0) {
    Components c;
    c.mapping = {0};
    return c;
  }
  switch (input.bytes) {
    case 1: return run_ccl<uint8_t>(ctx, input);
    case 2: return run_ccl<uint16_t>(ctx, input);
    case 4: return run_ccl<uint32_t>(ctx, input);
    case 8: return run_ccl<unsigned long long>(ctx, input);
    default: throw std::invalid_argument("connected components requires integer labels");
  }
}

std::vector<Box> analyze(Context &ctx, const Components &cc, Dims d) {
  std::vector<Box> boxes;
  if (int64_t(d.sx) * d.sy * d.sz == 0) { boxes.assign(1, Box{}); return boxes; }
  switch (cc.cc_bytes) {
    case 1: run_analyze<uint8_t>(ctx, cc, d, boxes); break;
    case 2: run_analyze<uint16_t>(ctx, cc, d, boxes); break;
    default: run_analyze<uint32_t>(ctx, cc, d, boxes); break;
  }
  return boxes;
}

void gather_faces(Context &ctx, const void *ids, int id_bytes, const float *dist, Dims d, const FaceCanvas &fc,
                  int f0, int f1, void *face_ids, float *face_dist) {
  if (f0 < 0 || f1 > 6 || f0 >= f1) throw std::invalid_argument("gather_faces: face range");
  const int dims[3] = {d.sx, d.sy, d.sz};
  long long total = 0;
  for (int f = f0; f < f1; ++f) total += (long long)dims[fc.a[f]] * dims[fc.b[f]];
  const unsigned g = blocks_for(total, 256);
  if (id_bytes == 1)
    faces_gather<uint8_t><<<g, 256, 0, ctx.stream()>>>(static_cast<const uint8_t *>(ids), dist, d, fc, f0, f1, total,
                                                       static_cast<uint8_t *>(face_ids), face_dist);
  else if (id_bytes == 2)
    faces_gather<uint16_t><<<g, 256, 0, ctx.stream()>>>(static_cast<const uint16_t *>(ids), dist, d, fc, f0, f1, total,
                                                        static_cast<uint16_t *>(face_ids), face_dist);
  else
    faces_gather<uint32_t><<<g, 256, 0, ctx.stream()>>>(static_cast<const uint32_t *>(ids), dist, d, fc, f0, f1, total,
                                                        static_cast<uint32_t *>(face_ids), face_dist);
  BROOK_LAUNCHED();
}

void extract_faces(Context &ctx, const void *cc, int cc_bytes, Dims d, const FaceCanvas &fc, void *canvas) {
  const unsigned g = blocks_for(int64_t(fc.rows) * fc.cols, 256);
  if (cc_bytes == 1)
    faces_copy<uint8_t><<<g, 256, 0, ctx.stream()>>>(static_cast<const uint8_t *>(cc), d, fc, static_cast<uint8_t *>(canvas));
  else if (cc_bytes == 2)
    faces_copy<uint16_t><<<g, 256, 0, ctx.stream()>>>(static_cast<const uint16_t *>(cc), d, fc,
                                                      static_cast<uint16_t *>(canvas));
  else
    faces_copy<uint32_t><<<g, 256, 0, ctx.stream()>>>(static_cast<const uint32_t *>(cc), d, fc,
                                                      static_cast<uint32_t *>(canvas));
  BROOK_LAUNCHED();
}

}  // namespace brook
