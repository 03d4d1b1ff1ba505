// Shared device helpers for the gfx950 kernels: neighbour tables, coherent loads, float tricks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace brook {

// 26-neighbour offsets in the reference's fixed order (faces, xy/yz/xz edges, corners;
// common.cuh:33-47). Every flood, walk and invalidation visits neighbours in this order.
__constant__ static const int kNX[26] = {-1, 1, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, -1, -1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1};
__constant__ static const int kNY[26] = {0, 0, -1, 1, 0, 0, -1, 1, -1, 1, -1, -1, 1, 1, 0, 0, 0, 0, -1, -1, 1, 1, -1, -1, 1, 1};
__constant__ static const int kNZ[26] = {0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1};
// Rank of neighbour k in dijkstra3d's order and its inverse (common.cuh:74-81); opposite offsets.
__constant__ static const int kD3Rank[26] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 21, 20, 24, 19, 23, 22, 25};
__constant__ static const int kD3Inv[26] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 22, 20, 19, 24, 23, 21, 25};
__constant__ static const int kOpp[26] = {1, 0, 3, 2, 5, 4, 9, 8, 7, 6, 13, 12, 11, 10, 17, 16, 15, 14, 25, 24, 23, 22, 21, 20, 19, 18};

// The same table as compile-time constants, for fully unrolled neighbour loops.
constexpr int kDX[26] = {-1, 1, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, -1, -1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1};
constexpr int kDY[26] = {0, 0, -1, 1, 0, 0, -1, 1, -1, 1, -1, -1, 1, 1, 0, 0, 0, 0, -1, -1, 1, 1, -1, -1, 1, 1};
constexpr int kDZ[26] = {0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1};
// The same deltas packed two bits per neighbour (delta + 1), for loops whose k is a run-time value:
// shifts of an immediate instead of table loads.
constexpr unsigned long long pack_deltas(const int (&t)[26]) {
  unsigned long long p = 0;
  for (int k = 0; k < 26; ++k) p |= (unsigned long long)(t[k] + 1) << (2 * k);
  return p;
}
constexpr unsigned long long kPackDX = pack_deltas(kDX), kPackDY = pack_deltas(kDY), kPackDZ = pack_deltas(kDZ);
constexpr int unpack_delta(unsigned long long p, int k) { return int((p >> (2 * k)) & 3ULL) - 1; }
constexpr bool deltas_match() {
  for (int k = 0; k < 26; ++k)
    if (unpack_delta(kPackDX, k) != kDX[k] || unpack_delta(kPackDY, k) != kDY[k] || unpack_delta(kPackDZ, k) != kDZ[k])
      return false;
  return true;
}
static_assert(deltas_match(), "packed neighbour deltas");
__device__ __forceinline__ int nb_dx(int k) { return unpack_delta(kPackDX, k); }
__device__ __forceinline__ int nb_dy(int k) { return unpack_delta(kPackDY, k); }
__device__ __forceinline__ int nb_dz(int k) { return unpack_delta(kPackDZ, k); }

constexpr float kInf = __builtin_huge_valf();
constexpr unsigned long long kMaxKey = ~0ULL;

struct Dims {
  int sx, sy, sz;
};

__device__ __forceinline__ void unravel(int v, const Dims &d, int &x, int &y, int &z) {
  int sxy = d.sx * d.sy;
  z = v / sxy;
  int r = v - z * sxy;
  y = r / d.sx;
  x = r - y * d.sx;
}
__device__ __forceinline__ bool inside(int x, int y, int z, const Dims &d) {
  return (unsigned)x < (unsigned)d.sx && (unsigned)y < (unsigned)d.sy && (unsigned)z < (unsigned)d.sz;
}
__device__ __forceinline__ int flat(int x, int y, int z, const Dims &d) { return x + d.sx * (y + d.sy * z); }
// Flat offset of neighbour k (only meaningful where its mask bit says it is inside the volume).
template <int K> __device__ __forceinline__ int nb_offset(const Dims &d) {
  return kDX[K] + d.sx * (kDY[K] + d.sy * kDZ[K]);
}
__device__ __forceinline__ int nb_offset(int k, const Dims &d) { return kNX[k] + d.sx * (kNY[k] + d.sy * kNZ[k]); }

// Calls f(integral_constant k) for k = B..E-1 with k known at compile time.
template <int B, int E, class F> __device__ __forceinline__ void for_nb(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    for_nb<B + 1, E>(f);
  }
}
template <class F> __device__ __forceinline__ void for_each_nb(F &&f) { for_nb<0, 26>(f); }

// Calls f(integral_constant B, integral_constant E) for each part [B, E) of the 26 neighbours that
// has a set bit in m; Batch = neighbours per part (26, 13 or 9).
template <int Batch, class F> __device__ __forceinline__ void for_parts(unsigned m, F &&f) {
  using std::integral_constant;
  if constexpr (Batch >= 26) {
    if (m) f(integral_constant<int, 0>{}, integral_constant<int, 26>{});
  } else if constexpr (Batch >= 13) {
    if (m & 0x1FFFu) f(integral_constant<int, 0>{}, integral_constant<int, 13>{});
    if (m >> 13) f(integral_constant<int, 13>{}, integral_constant<int, 26>{});
  } else {
    if (m & 0x1FFu) f(integral_constant<int, 0>{}, integral_constant<int, 9>{});
    if ((m >> 9) & 0x1FFu) f(integral_constant<int, 9>{}, integral_constant<int, 18>{});
    if (m >> 18) f(integral_constant<int, 18>{}, integral_constant<int, 26>{});
  }
}

// ---- device-memory access in the global address space -------------------------------------------
// Pointers that reach a kernel inside a struct are generic, and generic accesses compile to FLAT
// instructions, which also count against the LDS counter (so every LDS wait waits for them too)
// and cannot use a scalar base. The kernels access device memory through the forms below, which
// name the global address space: plain loads and stores, and relaxed agent-scope atomics (the
// semantics of HIP's atomicX functions).
template <class T> __device__ __forceinline__ __attribute__((address_space(1))) T *gp(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}
template <class T> struct same_type { using type = T; };
template <class T> using same_t = typename same_type<T>::type;
template <class T> __device__ __forceinline__ T g_ld(const T *p) { return *gp(p); }
template <class T> __device__ __forceinline__ void g_st(T *p, same_t<T> v) { *gp(p) = v; }
template <class T> __device__ __forceinline__ T g_min(T *p, same_t<T> v) {
  return __hip_atomic_fetch_min(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T g_max(T *p, same_t<T> v) {
  return __hip_atomic_fetch_max(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T g_add(T *p, same_t<T> v) {
  return __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T g_sub(T *p, same_t<T> v) {
  return __hip_atomic_fetch_sub(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T g_or(T *p, same_t<T> v) {
  return __hip_atomic_fetch_or(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T g_and(T *p, same_t<T> v) {
  return __hip_atomic_fetch_and(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T g_xchg(T *p, same_t<T> v) {
  return __hip_atomic_exchange(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Compare-and-swap; returns the previous value.
template <class T> __device__ __forceinline__ T g_cas(T *p, same_t<T> expected, same_t<T> desired) {
  __hip_atomic_compare_exchange_strong(gp(p), &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expected;
}

// Loads of memory that other threads change inside the same kernel: relaxed, agent scope, so the
// access is served by L2 (the CU's vector L1 may hold a stale line otherwise).
__device__ __forceinline__ float ld_f(const float *p) { return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ int ld_i(const int *p) { return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ unsigned ld_u(const unsigned *p) { return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ unsigned long long ld_u64(const unsigned long long *p) {
  return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned char ld_u8(const unsigned char *p) {
  return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_f(float *p, float v) { __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st_i(int *p, int v) { __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st_u8(unsigned char *p, unsigned char v) {
  __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_u64(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Loads and stores of per-voxel data that, during one kernel, only the workgroup owning the voxel's
// component label touches (floods, tracer): workgroup scope is all the memory model needs, and the
// accesses may then be served by the CU's vector L1 instead of always going to L2.
#ifndef BROOK_LABEL_SCOPE
#define BROOK_LABEL_SCOPE __HIP_MEMORY_SCOPE_WORKGROUP
#endif
__device__ __forceinline__ float lp_f(const float *p) { return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, BROOK_LABEL_SCOPE); }
__device__ __forceinline__ int lp_i(const int *p) { return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, BROOK_LABEL_SCOPE); }
__device__ __forceinline__ unsigned lp_u(const unsigned *p) { return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, BROOK_LABEL_SCOPE); }
__device__ __forceinline__ unsigned long long lp_u64(const unsigned long long *p) {
  return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, BROOK_LABEL_SCOPE);
}
__device__ __forceinline__ void sp_f(float *p, float v) { __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, BROOK_LABEL_SCOPE); }
__device__ __forceinline__ void sp_u64(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, BROOK_LABEL_SCOPE);
}

// One vertex-weighted step fl(d + f) that never returns d itself for a positive cost: one ulp up
// instead (the rule of the reference's brook_vw_step, common.cuh:123-127), which keeps distances
// strictly increasing along every step so predecessor walks terminate.
__device__ __forceinline__ float vw_step(float d, float f) {
  float c = d + f;
  if (c == d && f > 0.0f && d < kInf) c = __int_as_float(__float_as_int(d) + 1);
  return c;
}

// atomicMin of a non-negative float through its int bits (monotone for floats >= 0, +inf included);
// returns the previous value. A relaxed pre-check skips the atomic when it cannot improve.
__device__ __forceinline__ float atomic_min_pos(float *addr, float val) {
  float cur = ld_f(addr);
  if (cur <= val) return cur;
  return __int_as_float(g_min(reinterpret_cast<int *>(addr), __float_as_int(val)));
}

// Slot in a list whose length lives at *counter (any address space), for every lane that calls
// this at the same point: one atomic per wave, lanes numbered in lane order.
__device__ __forceinline__ int wave_slot(int *counter) {
  const unsigned long long act = __ballot(1);
  const int lane = __lane_id();
  const int leader = __ffsll((long long)act) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(act));
  base = __shfl(base, leader, 64);
  return base + __popcll(act & ((1ULL << lane) - 1ULL));
}

// Reserves n slots per lane in a list whose length lives at *counter; returns the lane's first
// slot. All 64 lanes must be active: one inclusive scan and one atomic per wave.
__device__ __forceinline__ int wave_reserve(int *counter, int n) {
  const int lane = __lane_id();
  int x = n;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  const int total = __shfl(x, 63, 64);
  int base = 0;
  if (lane == 63 && total) base = atomicAdd(counter, total);
  base = __shfl(base, 63, 64);
  return base + x - n;
}

// Appends u + off[k] for every set bit k of mask (all 64 lanes active; off: neighbour offsets).
__device__ __forceinline__ void emit_neighbours(int *list, int *counter, unsigned mask, int u, const int *off) {
  int pos = wave_reserve(counter, __popc(mask));
  while (mask) {
    const int k = __ffs(mask) - 1;
    mask &= mask - 1;
    g_st(list + pos++, u + off[k]);
  }
}

// Minimum over the wave (all 64 lanes active).
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

// Everything the floods and the tracer touch per voxel, in one 16-byte record (one cache line holds
// eight voxels of a row): a neighbour costs one line instead of one per array.
//   pdrf  the penalty field (rails 0)
//   meta  bits [0,26) the same-label neighbour mask, then kAlive and kWaiting (invalidation: waiting
//         for a later band)
//   word  a distance word (see word_dist_bits; the railroad also keeps bit 1 = "far listed") or an
//         invalidation claim (bottleneck, source, queued bit); ~0 between uses
struct alignas(16) VoxelRec {
  float pdrf;
  unsigned meta;
  unsigned long long word;
};
constexpr unsigned kNbrBits = (1u << 26) - 1u;
constexpr unsigned kAlive = 1u << 26, kWaiting = 1u << 29;

__device__ __forceinline__ float rec_pdrf(const VoxelRec *r, int v) { return lp_f(&r[v].pdrf); }
__device__ __forceinline__ unsigned rec_meta(const VoxelRec *r, int v) { return lp_u(&r[v].meta); }
// The whole record in one 16-byte load (x = pdrf bits, y = meta, z = word low, w = word high), for
// walks that need two of its fields: one scattered access instead of two. During the flood and
// trace kernels only the workgroup that owns the voxel's label touches its record, as for the
// workgroup-scope loads above.
typedef unsigned RecWords __attribute__((ext_vector_type(4)));
__device__ __forceinline__ RecWords rec_load(const VoxelRec *r, int v) {
  return *gp(reinterpret_cast<const RecWords *>(r + v));
}
// The high half of the word (distance bits or claim bottleneck) without the 64-bit load.
__device__ __forceinline__ unsigned rec_word_hi(const VoxelRec *r, int v) {
  return lp_u(reinterpret_cast<const unsigned *>(&r[v].word) + 1);
}

// A 64-bit voxel word of the distance floods: high half = the distance's float bits, low half bit 0
// = "not queued". ~0 (the shared sentinel) reads as +inf, not queued.
__device__ __forceinline__ unsigned word_dist_bits(unsigned long long w) {
  const unsigned hi = (unsigned)(w >> 32);
  return hi == 0xFFFFFFFFu ? 0x7f800000u : hi;
}

// Shuffle of any integer width up by `delta` lanes (wave64).
template <class T> __device__ __forceinline__ T shfl_up_any(T v, unsigned delta) {
  if constexpr (sizeof(T) == 8) return T(__shfl_up((unsigned long long)v, delta, 64));
  else return T(__shfl_up((unsigned)v, delta, 64));
}

// A group of lanes per front entry, for fronts of at most a quarter of the workgroup: each entry
// gets G lanes (the largest power of two up to 32 that fits every entry into the workgroup, at
// least 4); lane g of a group handles neighbours [g*S, g*S + S), S = ceil(26 / G) <= kSpan, so all
// of an entry's loads, then all of its atomics, are in flight together instead of part after part.
// One code path serves every G (a span is kSpan slots, the tail predicated off), which keeps a
// kernel's register pressure at the level of its one-thread-per-entry path. (Pairs of lanes for
// fronts up to half a workgroup, two spans each, measured no faster on the 512^3 workload.)
constexpr int kSpan = 7;
struct Group {
  int G, i, g, lead, kb, cnt;  // lanes per entry, entry, lane in group, group's first lane, span
};
// The group of thread `tid` for a front of n entries in a workgroup of `block` threads (host-callable,
// so the CPU tests check the layout for every n, block and thread).
__host__ __device__ constexpr Group group_at(int n, int block, int tid) {
  Group q{};
  q.G = 32;
  while (q.G > 4 && n * q.G > block) q.G >>= 1;
  int shift = 0;
  while ((1 << shift) < q.G) ++shift;
  const int span = (26 + q.G - 1) / q.G;
  q.i = tid >> shift;
  q.g = tid & (q.G - 1);
  q.lead = (tid & 63) & ~(q.G - 1);  // lane of the group's first thread in its wave
  q.kb = q.g * span;
  q.cnt = span < 26 - q.kb ? span : 26 - q.kb;  // may be <= 0: nothing to do
  return q;
}
__device__ __forceinline__ Group group_of(int n, int block) { return group_at(n, block, int(threadIdx.x)); }

// Mask of lanes 0..lane.
__device__ __forceinline__ unsigned long long lanes_upto(int lane) {
  return lane == 63 ? ~0ULL : ((2ULL << lane) - 1ULL);
}

// Runs of one label among a wave's 64 consecutive voxels of one x-row: every lane learns the lane
// its run starts at; a run's first lane (head) also learns the run's length. Invalid lanes are
// runs of their own. Every lane of the wave must call it.
struct WaveRun {
  bool head;
  int start_lane, len;
};
template <class T> __device__ __forceinline__ WaveRun wave_run(T l, bool valid, int x, int lane) {
  const T prev = shfl_up_any(l, 1);
  const bool start = !valid || lane == 0 || x == 0 || prev != l;
  const unsigned long long starts = __ballot(start);
  const unsigned long long after = starts & ~lanes_upto(lane);
  WaveRun r;
  r.head = start;
  r.start_lane = 63 - __clzll((long long)(starts & lanes_upto(lane)));
  r.len = (after ? __ffsll((long long)after) - 2 : 63) - lane + 1;
  return r;
}

// A workgroup's LDS table of labels (open addressing, bounded probing): run updates of one label
// fold into its slot, and the workgroup sends one global update per label it saw instead of one
// per run. A label whose probe chain is full gets -1 for good (slots are never freed), and its
// runs go to global memory directly.
constexpr int kLabelSlots = 256, kLabelProbes = 8;
struct LabelSlots {
  int tag[kLabelSlots];  // label, or -1
  __device__ void init() {
    for (int s = threadIdx.x; s < kLabelSlots; s += blockDim.x) tag[s] = -1;
  }
  static __device__ __forceinline__ int home(int l) { return int((unsigned(l) * 0x9E3779B1u) >> 24); }
  __device__ int insert(int l) {
    for (int p = 0, s = home(l); p < kLabelProbes; ++p, s = (s + 1) & (kLabelSlots - 1)) {
      const int old = atomicCAS(&tag[s], -1, l);
      if (old == -1 || old == l) return s;
    }
    return -1;
  }
  __device__ int find(int l) const {
    for (int p = 0, s = home(l); p < kLabelProbes; ++p, s = (s + 1) & (kLabelSlots - 1))
      if (tag[s] == l) return s;
    return -1;
  }
};
static_assert(kLabelSlots == 256, "home() keeps the top 8 bits of the hash");

}  // namespace brook
