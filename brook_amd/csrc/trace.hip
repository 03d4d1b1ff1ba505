// The TEASAR trace loop on gfx950: one workgroup per component label, labels pulled from a
// device work queue (largest first), every label's whole path sequence traced inside its
// workgroup with workgroup barriers only -- no grid-wide barrier, no host round trip per path.
//
// Per label, exactly the semantics of the reference's lockstep tracer (labels never interact, so
// one label at a time is exact; lockstep.cu:91-622, lockstep_coop.inc, lockstep_regular.inc):
//   target     manual targets first (in pop order), then the first alive voxel of the label's
//              descending-DAF order at or after its cursor          (lockstep_coop.inc:343-416)
//   railroad   vertex-weighted SSSP from the target with the one-ulp step, rails absorb, near-far
//              bands bounded by the best rail; the stop rail is min (distance, flat index)
//                                                                   (lockstep_coop.inc:71-305)
//   backtrace  one wave: lane k probes neighbour k, two wave-min reductions give the lowest k
//              among the smallest distances; the join takes u's first rail in dijkstra3d order
//                                                                   (lockstep_regular.inc:61-154)
//   rails      path voxels get PDRF 0                               (lockstep_regular.inc:246-251)
//   invalidate rolling ball, bands of the bottleneck M on a fixed grid, synchronous rounds from
//              a snapshot inside a band, exact pulls at band boundaries
//                                                         (invalidation_kernels.cuh:105-299)
// Neighbour walks read the per-voxel same-label mask (one word instead of 26 ids).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "trace.hpp"
#include "tuned.hpp"

namespace brook {
namespace {

constexpr int kBlock = BROOK_TRACE_BLOCK;  // threads per label workgroup
constexpr unsigned long long kMinfarNone = 0x7f8000007f800000ULL;
constexpr unsigned long long kClaimed = 0xFFFFFFFFFFFFFFFEULL;
constexpr int kProf = kTraceProfileSlots;
// Neighbours whose memory operations one thread keeps in flight together (26: all; 13: two parts;
// 9: three parts). Fewer in flight, fewer registers.
constexpr int kNbBatch = BROOK_NB_BATCH;


__device__ __forceinline__ int wave_min(int v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return v;
}

// Front lists (the near rounds of the railroad and the invalidation rounds): entry j lives in LDS for
// j < kLdsFront and in the label's global list beyond, so a round's appends and the next round's
// reads of a small front never wait for device memory.
constexpr int kLdsFront = 4096;
constexpr int kRadiusSlots = 4096;

struct Shared {
  // Front-list lengths rotate over three counters: round r reads cnt[r], appends to cnt[r+1] and
  // clears cnt[r+2] (read in round r-1), which saves the barrier that a swap of two would need.
  int cnt[3];
  int n_far, n_farnxt, n_touched, n_keep;
  int done, target, path_len, found, band;
  long long cursor, seg_end;
  int pcount, ordered, after_used;  // ordered: the current target came from the DAF order
  unsigned lowest;
  float T;
  unsigned long long rail_best, minfar;
  int store_pos;
  // The railroad's far lists hold 2 * count entries; a list longer than count after a round is
  // compacted (see far_compact). `overflow`: an append found the list full (never expected).
  int far_cap, overflow;
  int off[26];  // flat offsets of the 26 neighbours
  // per-label profile: [0] total [1] target [2] railroad [3] backtrace [4] invalidation (wall-clock
  // ticks), [5] railroad rounds, [6] invalidation rounds, [7] paths, [8] railroad phases,
  // [9] far entries scanned, [10] near entries expanded, [11] invalidation entries expanded,
  // [12] waiting entries pulled, [13] bands
  unsigned long long prof[kProf];
  long long tick;
  int front[2][kLdsFront];
  // invalidation: ball radii of the path's voxels by source (see source_radius)
  unsigned long long radius[kRadiusSlots];
};

__device__ __forceinline__ int rot_next(int r) { return r == 2 ? 0 : r + 1; }
__device__ __forceinline__ int rot_prev(int r) { return r == 0 ? 2 : r - 1; }

struct Front {
  int *g;  // the label's list in device memory (entries >= kLdsFront)
  int *s;  // LDS (entries < kLdsFront)
};
__device__ __forceinline__ int front_get(const Front &f, int j) { return j < kLdsFront ? f.s[j] : lp_i(f.g + j); }
__device__ __forceinline__ void front_put(const Front &f, int j, int v) {
  if (j < kLdsFront) f.s[j] = v;
  else g_st(f.g + j, v);
}
// emit_neighbours into a front list.
__device__ __forceinline__ void emit_front(const Front &f, int *counter, unsigned mask, int u, const int *off) {
  int pos = wave_reserve(counter, __popc(mask));
  while (mask) {
    const int k = __ffs(mask) - 1;
    mask &= mask - 1;
    front_put(f, pos++, u + off[k]);
  }
}

// emit_neighbours into a list of `cap` entries: a wave whose appends would pass the end writes
// nothing and raises `overflow` instead.
__device__ __forceinline__ void emit_capped(int *list, int *counter, unsigned mask, int u, const int *off, int cap,
                                            int *overflow) {
  const int n = __popc(mask);
  int pos = wave_reserve(counter, n);
  if (pos + n > cap) {
    if (n) *overflow = 1;
    return;
  }
  while (mask) {
    const int k = __ffs(mask) - 1;
    mask &= mask - 1;
    g_st(list + pos++, u + off[k]);
  }
}

// Rounds (and phases) one railroad or invalidation may take before the label is given up as not
// converging: far above what the label-correcting rounds need (a few per voxel of the label at
// most), and finite, so a logic error ends as a host-visible error instead of a hang.
__device__ __forceinline__ long long round_budget(int count) { return 4LL * count + 4096; }

// Statuses of one railroad: 1 nearest rail settled, 2 no rail reachable, 3 did not converge
// within the round budget, 4 a far list overflowed.
constexpr int kRailFound = 1, kNoRail = 2, kStalled = 3, kOverflowed = 4;
// counters[8] bits: 1 = a railroad or invalidation did not converge, 2 = a far list overflowed.
__device__ __forceinline__ void report_failure(const TraceArgs &a, int status) {
  g_or(a.counters + 8, status == kStalled ? 1 : 2);
}

__device__ __forceinline__ void lap(Shared &sh, int slot) {
  if (threadIdx.x == 0) {
    long long now = wall_clock64();
    sh.prof[slot] += (unsigned long long)(now - sh.tick);
    sh.tick = now;
  }
}

// The neighbours v may pull from (see TraceArgs::inmask).
__device__ __forceinline__ unsigned in_mask(const TraceArgs &a, int v) {
  return a.inmask ? g_ld(a.inmask + v) : rec_meta(a.rec, v) & kNbrBits;
}

// ---- railroad -------------------------------------------------------------------------------
// Distances live in the records' packed words: high half the distance bits, bit 0 "not
// queued", bit 1 "far listed". A front voxel is dequeued by one atomicOr that also returns its
// distance; an improvement is one atomicMin whose old value says whether the voxel still has to
// be queued or far-listed. Near candidates store 00 (queued), rails 01, far candidates 11; a near
// improvement of a far-listed voxel clears its far bit, so its far-list entry goes stale.
// Per-thread minima of the reached rails and far distances are folded into LDS once per round.
constexpr unsigned long long kRoadNear = 0, kRoadRail = 1, kRoadFar = 3;
// A word whose far bit is set by a far candidate (the unreached sentinel ~0 is not far-listed).
__device__ __forceinline__ bool far_listed(unsigned long long w) {
  return (w & 2ULL) && (unsigned)(w >> 32) != 0xFFFFFFFFu;
}

struct RoadMins {
  unsigned long long rail = kMaxKey, far = kMinfarNone;
};

// Neighbours (bit k = neighbour k) to append to the touched, near and far lists.
struct Pushes {
  unsigned touched = 0, near = 0, far = 0;
};

__device__ __forceinline__ void road_fold(Shared &sh, RoadMins &mins) {
  const unsigned long long r = wave_min_u64(mins.rail), f = wave_min_u64(mins.far);
  if (__lane_id() == 0) {
    if (r != kMaxKey) atomicMin(&sh.rail_best, r);
    if (f != kMinfarNone) atomicMin(&sh.minfar, f);
  }
  mins = RoadMins{};
}

// Masked neighbours B..E-1 of u at once: every load, then every atomic, is issued before any result
// is waited for; the outcome is a set of bit masks (list appends happen once per loop iteration).
// A rail is recognised by cand == du (vw_step adds nothing only for a zero cost).
template <int B, int E>
__device__ __forceinline__ void relax_part(const TraceArgs &a, RoadMins &mins, int u, unsigned m, float du, float bound,
                                           Pushes &p) {
  float c[E - B];
  unsigned seen[E - B];
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    const bool on = (m >> k) & 1u;
    const RecWords r = on ? rec_load(a.rec, u + nb_offset<k>(a.d)) : RecWords{0x7f800000u, 0u, 0u, 0u};
    c[i] = __uint_as_float(r.x);
    seen[i] = r.w;
  });
  unsigned long long old[E - B];
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    const float cc = vw_step(du, c[i]);
    c[i] = cc;
    const unsigned cb = __float_as_uint(cc);
    const bool rail = cc == du, near = !rail && cc <= bound;
    old[i] = 0;  // "nothing changed" unless the atomic says otherwise
    if (cb < seen[i] && cc < kInf)
      old[i] = g_min(&a.rec[u + nb_offset<k>(a.d)].word,
                         ((unsigned long long)cb << 32) | (near ? kRoadNear : rail ? kRoadRail : kRoadFar));
  });
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    const float cc = c[i];
    const unsigned cb = __float_as_uint(cc);
    const bool rail = cc == du, near = !rail && cc <= bound;
    const unsigned long long nv = ((unsigned long long)cb << 32) | (near ? kRoadNear : rail ? kRoadRail : kRoadFar);
    if (!(old[i] > nv)) return;
    if ((unsigned)(old[i] >> 32) == 0xFFFFFFFFu) p.touched |= 1u << k;
    if (rail) {
      const unsigned long long key = ((unsigned long long)cb << 32) | (unsigned)(u + nb_offset<k>(a.d));
      mins.rail = key < mins.rail ? key : mins.rail;
    } else if (near) {
      if (old[i] & 1ULL) p.near |= 1u << k;
    } else {
      const unsigned long long key = ((unsigned long long)cb << 32) | __float_as_uint(cc - du);
      mins.far = key < mins.far ? key : mins.far;
      if (!far_listed(old[i])) p.far |= 1u << k;
    }
  });
}

__device__ __forceinline__ void relax_all(const TraceArgs &a, RoadMins &mins, int u, unsigned m, float du, float bound,
                                          Pushes &p) {
  for_parts<kNbBatch>(m, [&](auto B, auto E) { relax_part<decltype(B)::value, decltype(E)::value>(a, mins, u, m, du, bound, p); });
}

// Dequeues front voxel u: marks it "not queued" and returns its distance, in one atomic.
__device__ __forceinline__ float dequeue(const TraceArgs &a, int u) {
  return __uint_as_float((unsigned)(g_or(&a.rec[u].word, 1ULL) >> 32));
}

// ---- fronts of up to a quarter workgroup: a group of lanes per entry (group_of) ----------------
constexpr int kGroupFront = kBlock / 4;
__device__ __forceinline__ Group group_of(int n) { return brook::group_of(n, kBlock); }

// Neighbours kb..kb+cnt-1 (those set in m; cnt <= kSpan) of u: relax_part with a run-time span.
__device__ __forceinline__ void relax_span(const TraceArgs &a, RoadMins &mins, int u, unsigned m, int kb, int cnt, float du,
                                           float bound, const int *off, Pushes &p) {
  constexpr int S = kSpan;
  float c[S];
  unsigned seen[S];
  int v[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int k = kb + j;
    const bool on = j < cnt && ((m >> k) & 1u);
    v[j] = u + off[on ? k : 0];
    const RecWords r = on ? rec_load(a.rec, v[j]) : RecWords{0x7f800000u, 0u, 0u, 0u};
    c[j] = __uint_as_float(r.x);
    seen[j] = r.w;
  }
  unsigned long long old[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const float cc = vw_step(du, c[j]);
    c[j] = cc;
    const unsigned cb = __float_as_uint(cc);
    const bool rail = cc == du, near = !rail && cc <= bound;
    old[j] = 0;  // "nothing changed" unless the atomic says otherwise
    if (cb < seen[j] && cc < kInf)
      old[j] = g_min(&a.rec[v[j]].word, ((unsigned long long)cb << 32) | (near ? kRoadNear : rail ? kRoadRail : kRoadFar));
  }
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int k = kb + j;
    const float cc = c[j];
    const unsigned cb = __float_as_uint(cc);
    const bool rail = cc == du, near = !rail && cc <= bound;
    const unsigned long long nv = ((unsigned long long)cb << 32) | (near ? kRoadNear : rail ? kRoadRail : kRoadFar);
    if (!(old[j] > nv)) continue;
    if ((unsigned)(old[j] >> 32) == 0xFFFFFFFFu) p.touched |= 1u << k;
    if (rail) {
      const unsigned long long key = ((unsigned long long)cb << 32) | (unsigned)v[j];
      mins.rail = key < mins.rail ? key : mins.rail;
    } else if (near) {
      if (old[j] & 1ULL) p.near |= 1u << k;
    } else {
      const unsigned long long key = ((unsigned long long)cb << 32) | __float_as_uint(cc - du);
      mins.far = key < mins.far ? key : mins.far;
      if (!far_listed(old[j])) p.far |= 1u << k;
    }
  }
}

// One railroad round over a front of at most kGroupFront entries, a group of lanes per entry: the
// group's first lane dequeues the entry and shares it.
__device__ __forceinline__ void relax_groups(const TraceArgs &a, Shared &sh, RoadMins &mins, const Front &cur, int nc,
                                             float bound, const Front &nxt, int *n_nxt, int *far, int *touched) {
  const Group q = group_of(nc);
  int u = 0;
  unsigned m = 0;
  float du = kInf;
  if (q.g == 0 && q.i < nc) {
    u = front_get(cur, q.i);
    m = rec_meta(a.rec, u) & kNbrBits;
    du = dequeue(a, u);
  }
  u = __shfl(u, q.lead, 64);
  m = __shfl(m, q.lead, 64);
  du = __shfl(du, q.lead, 64);
  Pushes p;
  if (q.i < nc && du <= bound) relax_span(a, mins, u, m, q.kb, q.cnt, du, bound, sh.off, p);
  emit_neighbours(touched, &sh.n_touched, p.touched, u, sh.off);
  emit_front(nxt, n_nxt, p.near, u, sh.off);
  emit_capped(far, &sh.n_far, p.far, u, sh.off, sh.far_cap, &sh.overflow);
}

// One railroad round over a larger front: a thread per entry, its neighbours in parts.
__device__ __forceinline__ void relax_front(const TraceArgs &a, Shared &sh, RoadMins &mins, const Front &cur, int nc,
                                            float bound, const Front &nxt, int *n_nxt, int *far, int *touched) {
  for (int base = 0; base < nc; base += blockDim.x) {  // uniform: every lane reaches the appends
    const int i = base + threadIdx.x;
    Pushes p;
    int u = 0;
    if (i < nc) {
      u = front_get(cur, i);
      const unsigned m = (rec_meta(a.rec, u) & kNbrBits);
      const float du = dequeue(a, u);
      if (du <= bound) relax_all(a, mins, u, m, du, bound, p);
    }
    emit_neighbours(touched, &sh.n_touched, p.touched, u, sh.off);
    emit_front(nxt, n_nxt, p.near, u, sh.off);
    emit_capped(far, &sh.n_far, p.far, u, sh.off, sh.far_cap, &sh.overflow);
  }
}

// Far-list compaction between rounds. A voxel can hold several entries: its entry goes stale when
// a near improvement clears its far bit, and it is far-listed anew when a later, smaller bound
// (a closer rail) leaves its next improvement beyond the band. Within one round the bound is fixed,
// so a round appends at most one entry per voxel (at most `count`); a list that holds more than
// `count` entries after a round keeps one entry per far-listed voxel (the one whose atomic clears
// the far bit, restored after a barrier), so 2 * count entries always suffice.
__device__ __forceinline__ void far_compact(const TraceArgs &a, Shared &sh, int *&fcur, int *&fnxt) {
  const int nf = sh.n_far;
  for (int i = threadIdx.x; i < nf; i += blockDim.x) {
    const int w = lp_i(fcur + i);
    if (far_listed(g_and(&a.rec[w].word, ~2ULL))) g_st(fnxt + wave_slot(&sh.n_farnxt), w);
  }
  __syncthreads();
  const int kept = sh.n_farnxt;
  for (int i = threadIdx.x; i < kept; i += blockDim.x) g_or(&a.rec[lp_i(fnxt + i)].word, 2ULL);
  __syncthreads();
  if (threadIdx.x == 0) { sh.n_far = kept; sh.n_farnxt = 0; }
  int *tmp = fcur; fcur = fnxt; fnxt = tmp;
  __syncthreads();
}

__device__ __forceinline__ float rail_bound(const Shared &sh) {
  float b = sh.T;
  unsigned long long rb = sh.rail_best;
  if (rb != kMaxKey) b = fminf(b, __uint_as_float((unsigned)(rb >> 32)));
  return b;
}

// Returns kRailFound when the nearest rail is settled (sh.rail_best), kNoRail when no rail is
// reachable, kStalled / kOverflowed when the run is given up (see round_budget, far_compact). A
// `source` run never takes t itself for a rail: with no rails in the label it settles every voxel
// reachable from t (the vertex-weighted distances of the fix_branching=False field). The far lists
// hold 2 * count entries.
__device__ __forceinline__ int railroad(const TraceArgs &a, Shared &sh, int t, int count, const Front &near_a,
                                        const Front &near_b, int *far_a, int *far_b, int *touched, bool source = false) {
  if (threadIdx.x == 0) {
    g_st(touched, t);
    sh.n_touched = 1;
    sh.cnt[0] = sh.cnt[1] = sh.cnt[2] = sh.n_far = sh.n_farnxt = 0;
    sh.far_cap = 2 * count;
    sh.overflow = 0;
    sh.minfar = kMinfarNone;
    float ft = rec_pdrf(a.rec, t);
    if (ft == 0.0f && !source) {
      sp_u64(&a.rec[t].word, 1ULL);  // distance 0, not queued
      sh.rail_best = (unsigned)t;
      sh.done = kRailFound;
    } else {
      sp_u64(&a.rec[t].word, 0ULL);  // distance 0, queued
      sh.rail_best = kMaxKey;
      sh.done = 0;
      front_put(near_a, 0, t);
      sh.cnt[0] = 1;
      sh.T = fmaxf(2.0f * ft, 1e-30f);
    }
  }
  __syncthreads();
  RoadMins mins;
  Front cur = near_a, nxt = near_b;
  int *fcur = far_a, *fnxt = far_b;
  int r = 0;  // the near list's length is sh.cnt[r]
  long long rounds = 0;  // rounds and phases so far (the same in every thread)
  const long long budget = round_budget(count);
  while (sh.done == 0) {
    bool stop = false;
    for (;;) {  // settle the near band
      const int nc = sh.cnt[r];
      if (nc == 0) break;
      const int rn = rot_next(r);
      if (threadIdx.x == 0) { sh.cnt[rot_prev(r)] = 0; sh.prof[5] += 1; sh.prof[10] += nc; }
      const float bound = rail_bound(sh);
      if (nc <= kGroupFront) relax_groups(a, sh, mins, cur, nc, bound, nxt, &sh.cnt[rn], fcur, touched);
      else relax_front(a, sh, mins, cur, nc, bound, nxt, &sh.cnt[rn], fcur, touched);
      road_fold(sh, mins);
      __syncthreads();
      r = rn;
      const Front tmp = cur; cur = nxt; nxt = tmp;
      if (sh.overflow || ++rounds > budget) { stop = true; break; }
      if (sh.n_far > count) far_compact(a, sh, fcur, fnxt);
    }
    if (stop) {  // uniform: every thread read the same flags after the round's barrier
      __syncthreads();
      if (threadIdx.x == 0) sh.done = sh.overflow ? kOverflowed : kStalled;
      __syncthreads();
      break;
    }
    if (threadIdx.x == 0) {  // phase end: nearest rail final? else advance the band
      unsigned long long rb = sh.rail_best, mf = sh.minfar;
      if (rb != kMaxKey && (__uint_as_float((unsigned)(rb >> 32)) <= sh.T || mf == kMinfarNone)) sh.done = kRailFound;
      else if (mf == kMinfarNone) sh.done = kNoRail;
      else if (rounds + 1 > budget) sh.done = kStalled;
      else {
        float dmin = __uint_as_float((unsigned)(mf >> 32)), f = __uint_as_float((unsigned)mf);
        float w = 2.0f * f, lo = 5e-4f * dmin;
        if (!(w > lo)) w = lo;
        sh.T = dmin + w;
        sh.minfar = kMinfarNone;
      }
      sh.prof[8] += 1;
      sh.prof[9] += sh.n_far;
    }
    ++rounds;
    __syncthreads();
    if (sh.done) break;
    const int nf = sh.n_far;
    const float bound = rail_bound(sh);
    // Promote far voxels inside the new band. Entries whose far bit is clear are stale (the voxel
    // went near); of several entries of one voxel, the one whose atomic clears the far bit acts.
    // A far-listed word is never queued, so the claim also queues it; a voxel still beyond the
    // band gets both bits back. Each voxel lands in at most one list (at most count entries).
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {
      const int w = lp_i(fcur + i);
      if (!far_listed(lp_u64(&a.rec[w].word))) continue;
      const unsigned long long was = g_and(&a.rec[w].word, ~3ULL);
      if (!far_listed(was)) continue;
      const float dw = __uint_as_float((unsigned)(was >> 32));
      if (dw <= bound) {
        front_put(cur, wave_slot(&sh.cnt[r]), w);
      } else {
        g_or(&a.rec[w].word, kRoadFar);
        const unsigned long long key = ((unsigned long long)__float_as_uint(dw) << 32) | __float_as_uint(rec_pdrf(a.rec, w));
        mins.far = key < mins.far ? key : mins.far;
        g_st(fnxt + wave_slot(&sh.n_farnxt), w);
      }
    }
    road_fold(sh, mins);
    __syncthreads();
    if (threadIdx.x == 0) { sh.n_far = sh.n_farnxt; sh.n_farnxt = 0; }
    int *tmp = fcur; fcur = fnxt; fnxt = tmp;
    __syncthreads();
  }
  return sh.done;
}

// ---- backtrace (wave 0) ---------------------------------------------------------------------
__device__ __forceinline__ void backtrace(const TraceArgs &a, Shared &sh, int t, int count, int *walk) {
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  const int ok_lane = lane < 26 ? lane : 0;
  int cur = (int)(unsigned)(sh.rail_best & 0xFFFFFFFFULL);
  int kept = 0;
  bool first = true;
  for (;;) {
    int kb = -1;
    const unsigned m = in_mask(a, cur);  // the predecessor's step to cur must be permitted
    if (cur != t) {
      float du = kInf;
      if (lane < 26 && ((m >> lane) & 1u)) {
        int u = cur + nb_offset(ok_lane, a.d);
        if (rec_pdrf(a.rec, u) != 0.0f) du = __uint_as_float(word_dist_bits(lp_u64(&a.rec[u].word)));
      }
      int db = __float_as_int(du);
      int mb = wave_min(db);
      if (mb < 0x7f800000) kb = wave_min(db == mb ? lane : 64);
    }
    if (first && kb >= 0) {  // the join at u = cur + N[kb]: u's first rail in dijkstra3d order
      const int u = cur + nb_offset(kb, a.d);
      const unsigned mu = (rec_meta(a.rec, u) & kNbrBits);
      int key = 64;
      if (lane < 26 && ((mu >> lane) & 1u) && rec_pdrf(a.rec, u + nb_offset(ok_lane, a.d)) == 0.0f) key = kD3Rank[lane];
      int r = wave_min(key);
      if (r < 26) {
        int ks = kD3Inv[r];
        cur = u + nb_offset(ks, a.d);
        kb = kOpp[ks];
      }
    }
    first = false;
    if (kept >= count) { kept = -1; break; }  // longer than the label: cannot happen on a valid field
    if (lane == 0) g_st(walk + kept, cur);
    ++kept;
    if (kb < 0) break;
    cur += nb_offset(kb, a.d);
  }
  if (lane == 0) sh.path_len = kept;
}

// ---- fix_branching=False: parent field and parent walk ---------------------------------------
// The field (batched.cu:246-251): the railroad from the root in `source` mode (the near-far
// bands of flood_lab_nf, batched_coop.inc:125-267) gives every voxel of the label its
// vertex-weighted distance over the PDRF; parents_lab (batched_regular.inc:113-134) then takes
// the same-label neighbour of smallest distance, first k on ties (-1 at the root and at voxels the
// run does not reach, as a label that filled holes split has), and the words go back to ~0.
// Returns the railroad's status: kNoRail is the complete field; kRailFound means the run met a
// voxel of PDRF 0 besides the root (it would stop there as at a rail; PDRF underflows to 0 only for
// extreme exponents), and the label is ejected; kStalled / kOverflowed end the label as failed.
__device__ __forceinline__ int parent_field(const TraceArgs &a, Shared &sh, int root, const int *voxels, int count,
                                            const Front &na, const Front &nb, int *l2, int *l3, int *touched) {
  for (int j = threadIdx.x; j < count; j += blockDim.x) g_st(a.parent + g_ld(voxels + j), -1);  // unreached: no parent
  __syncthreads();
  const int status = railroad(a, sh, root, count, na, nb, l2, l3, touched, true);
  const int nt = sh.n_touched;
  for (int j = threadIdx.x; j < nt; j += blockDim.x) {
    const int v = g_ld(touched + j);
    const unsigned dv = word_dist_bits(lp_u64(&a.rec[v].word));
    int p = -1;
    if (v != root && dv != 0u && dv < 0x7f800000u) {
      unsigned m = in_mask(a, v), best = 0x7f800000u;
      while (m) {  // ascending k: the first of equal distances stays
        const int k = __ffs(m) - 1;
        m &= m - 1;
        const int u = v + sh.off[k];
        const unsigned du = word_dist_bits(lp_u64(&a.rec[u].word));
        if (du < best) { best = du; p = u; }
      }
    }
    g_st(a.parent + v, p);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nt; j += blockDim.x) sp_u64(&a.rec[g_ld(touched + j)].word, kMaxKey);
  __syncthreads();
  return status;
}

// The walk (one thread):
// backtrace_parent_lab (lockstep_regular.inc:166-202): from the target along the parent field,
// keeping every vertex up to and including the first without a parent (-1 at the root, STOP -2
// at a vertex an earlier path emitted). The walk is a pointer chase; its length is the path's.
__device__ __forceinline__ void parent_walk(const TraceArgs &a, Shared &sh, int t, int count, int *walk) {
  if (threadIdx.x != 0) return;
  int cur = t, kept = 0;
  for (;;) {
    const int p = lp_i(a.parent + cur);
    if (kept >= count) { kept = -1; break; }  // longer than the label: cannot happen on a valid field
    g_st(walk + kept, cur);
    ++kept;
    if (p < 0) break;
    cur = p;
  }
  sh.path_len = kept;
}

// ---- invalidation ---------------------------------------------------------------------------
// A claim word is (bottleneck bits << 32) | (source << 1) | queued: claims order by (bottleneck,
// source), and the improving atomicMin returns whether the voxel is already queued for the next
// round (an equal claim never re-queues). A round's snapshot clears the queued bit with the atomic
// that reads the claim. ~0 is unclaimed (and not queued).
__device__ __forceinline__ unsigned long long claim_word(unsigned mb, unsigned su) {
  return ((unsigned long long)mb << 32) | ((unsigned long long)su << 1) | 1ULL;
}
__device__ __forceinline__ unsigned claim_source(unsigned long long w) { return (unsigned)w >> 1; }

// The invalidation ball around source s: scale * DBF(s) + constant. Every source of an
// invalidation is a voxel of its path, so the radii are staged once in an LDS table of (source,
// radius) slots, direct-mapped by a multiplicative hash (of colliding path voxels the lowest index
// keeps the slot); a source found in its slot costs no device-memory load, any other is read as
// before. Table and fallback give the same bits (one expression).
__device__ __forceinline__ float ball_radius(const Ball &ball, float dbf) { return ball.scale * dbf + ball.constant; }
__device__ __forceinline__ int radius_slot(unsigned s) { return int((s * 0x9E3779B1u) >> 20); }
static_assert(kRadiusSlots == 1 << 12, "radius_slot yields 12 bits");
__device__ __forceinline__ float source_radius(const TraceArgs &a, const Ball &ball, const unsigned long long *tab,
                                               unsigned s) {
  const unsigned long long e = tab[radius_slot(s)];
  return (unsigned)(e >> 32) == s ? __uint_as_float((unsigned)e) : ball_radius(ball, g_ld(a.dbf + s));
}
// An improving claim whose old word says the voxel still has to be queued.
__device__ __forceinline__ bool needs_queue(unsigned long long old) { return old == kMaxKey || !(old & 1ULL); }

__device__ __forceinline__ float offer_distance(const TraceArgs &a, int x, int y, int z, int qx, int qy, int qz) {
  float dx = a.wx * (float)(x - qx), dy = a.wy * (float)(y - qy), dz = a.wz * (float)(z - qz);
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

// A neighbour whose offer is beyond the band waits for it: if its record showed it unclaimed and
// not yet waiting, one atomic sets the waiting flag and its winner reports it in p.far, to be
// pulled at the band boundary. (A voxel claimed meanwhile may be listed too: the boundary skips
// claimed entries; a claimed voxel keeps the flag until the claimed set dies.)
__device__ __forceinline__ bool may_wait(const RecWords &r) { return r.w == 0xFFFFFFFFu && !(r.y & kWaiting); }

// Masked neighbours B..E-1 of u at once: alive and key loads, then the claim and waiting atomics,
// in flight together. Only the key's high word (the bottleneck) is pre-loaded: a claim whose
// bottleneck is above it cannot win and issues no atomic. The outcome is a set of bit masks.
template <int B, int E>
__device__ __forceinline__ void offer_part(const TraceArgs &a, const int *off, int u, unsigned m, int x, int y, int z,
                                           int qx, int qy, int qz, unsigned su, unsigned mu, float maxd, unsigned hi,
                                           Pushes &p) {
  bool al[E - B];
  unsigned kh[E - B], fresh = 0;
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    const bool on = (m >> k) & 1u;
    const RecWords r = on ? rec_load(a.rec, u + nb_offset<k>(a.d)) : RecWords{0u, 0u, 0u, 0u};
    al[i] = r.y & kAlive;
    kh[i] = r.w;
    if (may_wait(r)) fresh |= 1u << k;
  });
  unsigned mbs[E - B];
  unsigned wait = 0;
  unsigned long long old[E - B];
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    const float d = offer_distance(a, x + kDX[k], y + kDY[k], z + kDZ[k], qx, qy, qz);
    const unsigned mb = max(__float_as_uint(d), mu);  // M = max(M[u], d)
    mbs[i] = mb;
    old[i] = 0;  // "no improvement" unless the atomic says otherwise
    if (al[i] && d < maxd) {
      if (mb >= hi) {
        if ((fresh >> k) & 1u) {
          wait |= 1u << k;
          old[i] = g_or(&a.rec[u + nb_offset<k>(a.d)].meta, kWaiting);
        }
      } else if (kh[i] >= mb) {
        old[i] = g_min(&a.rec[u + nb_offset<k>(a.d)].word, claim_word(mb, su));
      }
    }
  });
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    if ((wait >> k) & 1u) {  // old: the meta before the waiting flag
      if (!(old[i] & kWaiting)) p.far |= 1u << k;
    } else if (old[i] > claim_word(mbs[i], su)) {  // won (old is 0 where no atomic was issued)
      if (old[i] == kMaxKey) p.touched |= 1u << k;
      if (needs_queue(old[i])) p.near |= 1u << k;
    }
  });
}

// Neighbours kb..kb+cnt-1 (those set in m; cnt <= kSpan) of u: offer_part with a run-time span.
__device__ __forceinline__ void offer_span(const TraceArgs &a, const int *off, int u, unsigned m, int kb, int cnt, int x,
                                           int y, int z, int qx, int qy, int qz, unsigned su, unsigned mu, float maxd,
                                           unsigned hi, Pushes &p) {
  constexpr int S = kSpan;
  bool al[S];
  unsigned kh[S], fresh = 0;
  int v[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int k = kb + j;
    const bool on = j < cnt && ((m >> k) & 1u);
    v[j] = u + off[on ? k : 0];
    const RecWords r = on ? rec_load(a.rec, v[j]) : RecWords{0u, 0u, 0u, 0u};
    al[j] = r.y & kAlive;
    kh[j] = r.w;
    if (may_wait(r)) fresh |= 1u << j;
  }
  unsigned mbs[S];
  unsigned wait = 0;  // bit j: slot j's atomic set the waiting flag
  unsigned long long old[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int k = j < cnt ? kb + j : 0;  // slots past the span are off (al[j] false)
    const float d = offer_distance(a, x + nb_dx(k), y + nb_dy(k), z + nb_dz(k), qx, qy, qz);
    const unsigned mb = max(__float_as_uint(d), mu);  // M = max(M[u], d)
    mbs[j] = mb;
    old[j] = 0;  // "no improvement" unless the atomic says otherwise
    if (al[j] && d < maxd) {
      if (mb >= hi) {
        if ((fresh >> j) & 1u) {
          wait |= 1u << j;
          old[j] = g_or(&a.rec[v[j]].meta, kWaiting);
        }
      } else if (kh[j] >= mb) {
        old[j] = g_min(&a.rec[v[j]].word, claim_word(mb, su));
      }
    }
  }
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int k = kb + j;
    if ((wait >> j) & 1u) {  // old: the meta before the waiting flag
      if (!(old[j] & kWaiting)) p.far |= 1u << k;
    } else if (old[j] > claim_word(mbs[j], su)) {  // won
      if (old[j] == kMaxKey) p.touched |= 1u << k;
      if (needs_queue(old[j])) p.near |= 1u << k;
    }
  }
}

// One invalidation round over a front of at most kGroupFront entries, a group of lanes per entry.
// The group's first lane takes the entry's snapshot (re-arming it) and its ball radius before the
// barrier that separates the round's snapshots from its offers; the values stay in registers.
__device__ __forceinline__ void offer_groups(const TraceArgs &a, const Ball &ball, Shared &sh, const Front &cur, int nc, unsigned hi,
                                             const Front &nxt, int *n_nxt, int *far, int *touched) {
  const Group q = group_of(nc);
  int u = 0;
  unsigned m = 0;
  unsigned long long ku = kMaxKey;
  float maxd = 0.0f;
  if (q.i < nc) {
    u = front_get(cur, q.i);
    m = rec_meta(a.rec, u) & kNbrBits;
    if (q.g == 0) {
      ku = g_and(&a.rec[u].word, ~1ULL);
      maxd = source_radius(a, ball, sh.radius, claim_source(ku));
    }
  }
  __syncthreads();
  ku = __shfl(ku, q.lead, 64);
  maxd = __shfl(maxd, q.lead, 64);
  Pushes p;
  if (q.i < nc) {
    const unsigned su = claim_source(ku), mu = (unsigned)(ku >> 32);
    int x, y, z, qx, qy, qz;
    unravel(u, a.d, x, y, z);
    unravel(int(su), a.d, qx, qy, qz);
    offer_span(a, sh.off, u, m, q.kb, q.cnt, x, y, z, qx, qy, qz, su, mu, maxd, hi, p);
  }
  emit_neighbours(touched, &sh.n_touched, p.touched, u, sh.off);
  emit_front(nxt, n_nxt, p.near, u, sh.off);
  emit_neighbours(far, &sh.n_far, p.far, u, sh.off);
}

__device__ __forceinline__ void offer_all(const TraceArgs &a, const int *off, int u, unsigned m, unsigned su, unsigned mu,
                                          float maxd, unsigned hi, Pushes &p) {
  int x, y, z, qx, qy, qz;
  unravel(u, a.d, x, y, z);
  unravel(int(su), a.d, qx, qy, qz);
  for_parts<kNbBatch>(m, [&](auto B, auto E) {
    offer_part<decltype(B)::value, decltype(E)::value>(a, off, u, m, x, y, z, qx, qy, qz, su, mu, maxd, hi, p);
  });
}

// The best claim waiting voxel c (at x, y, z) can pull from its claimed neighbours B..E-1: the
// smallest (max(bottleneck, distance to the neighbour's source), source) among those whose source's
// ball holds c. Claim loads, then the sources' DBF loads, each issued together.
template <int B, int E>
__device__ __forceinline__ void pull_part(const TraceArgs &a, const Ball &ball, const unsigned long long *tab, int c,
                                          unsigned mc, int x, int y, int z, unsigned long long &best) {
  unsigned long long ku[E - B];
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    ku[i] = (mc >> k) & 1u ? lp_u64(&a.rec[c + nb_offset<k>(a.d)].word) : kMaxKey;
  });
  float r[E - B];
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    r[i] = ku[i] != kMaxKey ? source_radius(a, ball, tab, claim_source(ku[i])) : 0.0f;
  });
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    if (ku[i] == kMaxKey) return;
    const unsigned su = claim_source(ku[i]);
    int qx, qy, qz;
    unravel(int(su), a.d, qx, qy, qz);
    const float d = offer_distance(a, x, y, z, qx, qy, qz);
    if (d >= r[i]) return;
    const unsigned mb = max(__float_as_uint(d), (unsigned)(ku[i] >> 32));
    const unsigned long long cand = ((unsigned long long)mb << 32) | su;
    if (cand < best) best = cand;
  });
}

// Neighbours kb..kb+cnt-1 (those set in mc; cnt <= kSpan) of waiting voxel c: pull_part with a
// run-time span.
__device__ __forceinline__ void pull_span(const TraceArgs &a, const Ball &ball, const unsigned long long *tab, const int *off, int c,
                                          unsigned mc, int kb, int cnt, int x, int y, int z, unsigned long long &best) {
  constexpr int S = kSpan;
  unsigned long long ku[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int k = kb + j;
    ku[j] = j < cnt && ((mc >> k) & 1u) ? lp_u64(&a.rec[c + off[k]].word) : kMaxKey;
  }
  float r[S];
#pragma unroll
  for (int j = 0; j < S; ++j) r[j] = ku[j] != kMaxKey ? source_radius(a, ball, tab, claim_source(ku[j])) : 0.0f;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    if (ku[j] == kMaxKey) continue;
    const unsigned su = claim_source(ku[j]);
    int qx, qy, qz;
    unravel(int(su), a.d, qx, qy, qz);
    const float d = offer_distance(a, x, y, z, qx, qy, qz);
    if (d >= r[j]) continue;
    const unsigned mb = max(__float_as_uint(d), (unsigned)(ku[j] >> 32));
    const unsigned long long cand = ((unsigned long long)mb << 32) | su;
    if (cand < best) best = cand;
  }
}

// Band-boundary pulls of at most kGroupFront waiting entries, a group of lanes per entry, the
// group's minimum folded by shuffles: pull[i] = kClaimed (claimed meanwhile), the best claim, or
// kMaxKey.
__device__ __forceinline__ void pull_groups(const TraceArgs &a, const Ball &ball, Shared &sh, const int *far, int nf,
                                            unsigned long long *pull) {
  const Group q = group_of(nf);
  unsigned long long best = kMaxKey;
  bool claimed = false;
  if (q.i < nf) {
    const int c = lp_i(far + q.i);
    const unsigned long long w = lp_u64(&a.rec[c].word);
    const unsigned mc = in_mask(a, c);
    claimed = w != kMaxKey;
    if (!claimed) {
      int x, y, z;
      unravel(c, a.d, x, y, z);
      pull_span(a, ball, sh.radius, sh.off, c, mc, q.kb, q.cnt, x, y, z, best);
    }
  }
  for (int o = 1; o < q.G; o <<= 1) {
    const unsigned long long other = __shfl_xor(best, o, 64);
    best = other < best ? other : best;
  }
  if (q.g == 0 && q.i < nf) {
    g_st(pull + q.i, claimed ? kClaimed : best);
    if (!claimed && best != kMaxKey) atomicMin(&sh.lowest, (unsigned)(best >> 32));
  }
}

__device__ __forceinline__ int band_of(unsigned mbits, float delta, int kmin) {
  float m = __uint_as_float(mbits);
  int k = (int)(m / delta);
  while ((float)(k + 1) * delta <= m) k++;
  while (k > 0 && (float)k * delta > m) k--;
  return k < kmin ? kmin : k;
}

// Returns false when the rounds did not converge within the budget (the label is then given up;
// the claimed set still dies).
__device__ __forceinline__ bool invalidate(const TraceArgs &a, const Ball &ball, Shared &sh, const int *path, int len,
                                           int count, const Front &la, const Front &lb, int *lfar, int *lkeep,
                                           int *touched, unsigned long long *snap, unsigned long long *pull) {
  if (threadIdx.x == 0) {
    sh.cnt[0] = sh.cnt[1] = sh.cnt[2] = sh.n_far = sh.n_keep = sh.n_touched = 0;
    sh.band = 0;
  }
  for (int j = threadIdx.x; j < kRadiusSlots; j += blockDim.x) sh.radius[j] = kMaxKey;
  __syncthreads();
  Front cur = la, nxt = lb;
  int *far = lfar, *keep = lkeep;
  int r = 0;  // the current list's length is sh.cnt[r]
  for (int j = threadIdx.x; j < len; j += blockDim.x) {  // sources: path voxels still alive
    int p = g_ld(path + j);
    atomicMin(&sh.radius[radius_slot(unsigned(p))],
              ((unsigned long long)unsigned(p) << 32) | __float_as_uint(ball_radius(ball, g_ld(a.dbf + p))));
    if (!(rec_meta(a.rec, p) & kAlive)) continue;
    if (g_cas(&a.rec[p].word, kMaxKey, claim_word(0u, unsigned(p))) == kMaxKey) {
      front_put(cur, wave_slot(&sh.cnt[0]), p);
      g_st(touched + wave_slot(&sh.n_touched), p);
    }
  }
  __syncthreads();
  long long rounds = 0;  // rounds and bands so far (the same in every thread)
  const long long budget = round_budget(count);
  bool converged = true;
  for (;;) {
    unsigned hi = __float_as_uint((float)(sh.band + 1) * a.delta);
    for (;;) {  // synchronous rounds inside the band
      const int nc = sh.cnt[r];
      if (nc == 0) break;
      if (++rounds > budget) { converged = false; break; }
      const int rn = rot_next(r);
      if (threadIdx.x == 0) { sh.cnt[rot_prev(r)] = 0; sh.prof[6] += 1; sh.prof[11] += nc; }
      if (nc <= kGroupFront) {
        offer_groups(a, ball, sh, cur, nc, hi, nxt, &sh.cnt[rn], far, touched);
      } else {
        for (int i = threadIdx.x; i < nc; i += blockDim.x)  // snapshot and re-arm in one atomic
          g_st(snap + i, g_and(&a.rec[front_get(cur, i)].word, ~1ULL));
        __syncthreads();
        for (int base = 0; base < nc; base += blockDim.x) {  // uniform: every lane reaches the appends
          const int i = base + threadIdx.x;
          Pushes p;
          int u = 0;
          if (i < nc) {
            u = front_get(cur, i);
            const unsigned m = (rec_meta(a.rec, u) & kNbrBits);
            const unsigned long long ku = g_ld(snap + i);
            const unsigned su = claim_source(ku), mu = (unsigned)(ku >> 32);
            offer_all(a, sh.off, u, m, su, mu, source_radius(a, ball, sh.radius, su), hi, p);
          }
          emit_neighbours(touched, &sh.n_touched, p.touched, u, sh.off);
          emit_front(nxt, &sh.cnt[rn], p.near, u, sh.off);
          emit_neighbours(far, &sh.n_far, p.far, u, sh.off);
        }
      }
      __syncthreads();
      r = rn;
      const Front tmp = cur; cur = nxt; nxt = tmp;
    }
    const int nf = sh.n_far;
    if (nf == 0 || !converged || ++rounds > budget) { converged = converged && nf == 0; break; }
    if (threadIdx.x == 0) { sh.lowest = 0x7f800000u; sh.prof[12] += nf; sh.prof[13] += 1; }
    __syncthreads();
    if (nf <= kGroupFront) {  // band boundary: exact pulls
      pull_groups(a, ball, sh, far, nf, pull);
    } else {
      for (int i = threadIdx.x; i < nf; i += blockDim.x) {
        int c = lp_i(far + i);
        if (lp_u64(&a.rec[c].word) != kMaxKey) { g_st(pull + i, kClaimed); continue; }
        const unsigned mc = in_mask(a, c);
        int x, y, z;
        unravel(c, a.d, x, y, z);
        unsigned long long best = kMaxKey;
        for_parts<kNbBatch>(mc, [&](auto B, auto E) { pull_part<decltype(B)::value, decltype(E)::value>(a, ball, sh.radius, c, mc, x, y, z, best); });
        g_st(pull + i, best);
        if (best != kMaxKey) atomicMin(&sh.lowest, (unsigned)(best >> 32));
      }
    }
    __syncthreads();
    if (threadIdx.x == 0 && sh.lowest != 0x7f800000u) sh.band = band_of(sh.lowest, a.delta, sh.band + 1);
    __syncthreads();
    hi = __float_as_uint((float)(sh.band + 1) * a.delta);
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {  // promote / keep waiting / drop
      int c = lp_i(far + i);
      unsigned long long p = g_ld(pull + i);
      if (p == kClaimed) continue;
      if (p == kMaxKey) { g_and(&a.rec[c].meta, ~kWaiting); continue; }
      if ((unsigned)(p >> 32) < hi) {
        sp_u64(&a.rec[c].word, claim_word((unsigned)(p >> 32), (unsigned)p));
        g_st(touched + wave_slot(&sh.n_touched), c);
        front_put(cur, wave_slot(&sh.cnt[r]), c);
        g_and(&a.rec[c].meta, ~kWaiting);
      } else {
        g_st(keep + wave_slot(&sh.n_keep), c);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) { sh.n_far = sh.n_keep; sh.n_keep = 0; }
    int *tmp = far; far = keep; keep = tmp;
    __syncthreads();
  }
  __syncthreads();
  const int nt = sh.n_touched;
  if (threadIdx.x == 0) {
    g_add(reinterpret_cast<unsigned long long *>(a.counters + 6), (unsigned long long)nt);
    if (!converged) report_failure(a, kStalled);
  }
  for (int j = threadIdx.x; j < nt; j += blockDim.x) {  // the claimed set dies
    int v = g_ld(touched + j);
    sp_u64(&a.rec[v].word, kMaxKey);
    g_and(&a.rec[v].meta, ~(kAlive | kWaiting));
  }
  __syncthreads();
  return converged;
}

// ---- target selection -----------------------------------------------------------------------
__device__ __forceinline__ int next_target(const TraceArgs &a, Shared &sh, int L) {
  if (threadIdx.x == 0) {
    int pc = sh.pcount;
    sh.ordered = 0;
    if (pc >= g_ld(a.maxp + L)) sh.target = -1;
    else if (pc < g_ld(a.mcount + L)) sh.target = g_ld(a.mflat + g_ld(a.mstart + L) + pc);
    else sh.target = -2;
  }
  __syncthreads();
  if (sh.target != -2) return sh.target;
  const int per = 4, window = per * blockDim.x;
  for (;;) {
    long long c = sh.cursor, e = sh.seg_end;
    if (c >= e) {  // the order is spent: the extra targets after it, last given first (trace.cu:197-199)
      __syncthreads();
      if (threadIdx.x == 0) {
        const int after = a.acount ? g_ld(a.acount + L) : 0;
        if (sh.after_used < after) {
          sh.target = g_ld(a.mflat + g_ld(a.mstart + L) + g_ld(a.mcount + L) + sh.after_used);
          sh.after_used += 1;
        } else {
          sh.target = -1;
        }
      }
      __syncthreads();
      return sh.target;
    }
    if (threadIdx.x == 0) sh.found = INT32_MAX;
    __syncthreads();
    for (int r = 0; r < per; ++r) {
      long long j = c + r * blockDim.x + threadIdx.x;
      if (j < e) {
        int v = g_ld(a.order + j);
        if ((rec_meta(a.rec, v) & kAlive)) atomicMin(&sh.found, int(j - c));
      }
    }
    __syncthreads();
    if (sh.found != INT32_MAX) {
      __syncthreads();
      if (threadIdx.x == 0) {
        sh.cursor = c + sh.found;
        sh.target = g_ld(a.order + c + sh.found);
        sh.ordered = 1;
      }
      __syncthreads();
      return sh.target;
    }
    __syncthreads();
    if (threadIdx.x == 0) sh.cursor = c + window;
    __syncthreads();
  }
}

// ---- somata ------------------------------------------------------------------------------------
// Physical distance from the soma root above the soma radius, in float64 as numpy's
// norm(anisotropy_f32 * (vertex - root)) (lockstep_regular.inc:35-45).
__device__ __forceinline__ bool soma_far(const TraceArgs &a, int v, int rx, int ry, int rz, double radius) {
  int x, y, z;
  unravel(v, a.d, x, y, z);
  const double dx = (double)a.wx * ((double)x - (double)rx), dy = (double)a.wy * ((double)y - (double)ry),
               dz = (double)a.wz * ((double)z - (double)rz);
  double ss = dx * dx + dy * dy;
  ss = ss + dz * dz;
  return sqrt(ss) > radius;
}

// One thread: a soma's walk compacted in place to the vertices Kimimaro keeps
// (lockstep_regular.inc:128-147, 184-195). A railroad walk runs rail end -> target: the target
// and the far vertices stay, and a rail end within the radius is replaced by the root. A parent
// walk runs target -> root: the far vertices and the root stay. Every write lands at or before
// the entry just read, except that a walk of one vertex gets two entries (a label has more voxels).
__device__ __forceinline__ void soma_trim(const TraceArgs &a, Shared &sh, int *walk, int root, double radius) {
  int rx, ry, rz;
  unravel(root, a.d, rx, ry, rz);
  const int n = sh.path_len;
  int kept = 0;
  for (int j = 0; j < n; ++j) {
    const int v = g_ld(walk + j);
    const bool far = soma_far(a, v, rx, ry, rz, radius);
    if (!a.parent && j == 0 && !far) g_st(walk + kept++, root);
    if (far || j == n - 1) g_st(walk + kept++, v);
  }
  sh.path_len = kept;
}

// The somata's own balls (lockstep.cu:197-206), one workgroup per soma label, before the trace:
// the ball of the soma parameters around the root dies, and the label's path budget becomes its
// voxel count minus what the ball took (or max_paths), none when the manual targets reach it.
__global__ void __launch_bounds__(kBlock) soma_balls(const TraceArgs *__restrict__ all, const int2 *__restrict__ queue,
                                                    int nqueue) {
  __shared__ Shared sh;
  const int2 q = queue[blockIdx.x];
  const TraceArgs a = all[__builtin_amdgcn_readfirstlane(q.x)];
  const int L = __builtin_amdgcn_readfirstlane(q.y);
  if (threadIdx.x < 26) sh.off[threadIdx.x] = nb_offset(int(threadIdx.x), a.d);
  const int seg = g_ld(a.seg_start + L), count = g_ld(a.seg_count + L);
  // lists 2 and 3 (the far lists) hold two entries per voxel: segment 2 * seg
  int *l2 = a.lists[2] + 2LL * seg, *l3 = a.lists[3] + 2LL * seg, *l4 = a.lists[4] + seg, *walk = a.lists[5] + seg;
  const Front na{a.lists[0] + seg, sh.front[0]}, nb{a.lists[1] + seg, sh.front[1]};
  unsigned long long *s0 = a.snaps[0] + seg, *s1 = a.snaps[1] + seg;
  const SomaLabel *soma = a.soma + L;
  if (threadIdx.x == 0) {
    for (int j = 0; j < kProf; ++j) sh.prof[j] = 0;
    g_st(walk, int(g_ld(a.root + L)));
  }
  __syncthreads();
  invalidate(a, soma->ball, sh, walk, 1, count, na, nb, l2, l3, l4, s0, s1);
  if (threadIdx.x == 0) {
    const long long limit = g_ld(&soma->max_paths);
    const long long budget = limit >= 0 ? limit : (long long)count - sh.n_touched;
    const int manual = g_ld(a.mcount + L) + (a.acount ? g_ld(a.acount + L) : 0);
    g_st(a.maxp + L, manual >= budget ? 0 : int(budget < INT32_MAX ? budget : INT32_MAX));
  }
}

// The queue is sorted largest first. Workgroup w < nfixed (the workgroups resident from the start,
// at most the queue's length) takes item w first; every other item comes from the shared cursor,
// from item nfixed on. Workgroups go to the XCDs in turn (w mod 8), so the largest labels start on
// different XCDs instead of wherever the first atomics land, and no XCD's L2 carries more than
// its share of them.
__global__ void __launch_bounds__(kBlock) trace_labels(const TraceArgs *__restrict__ all, const int2 *__restrict__ queue,
                                                      int nqueue, int nfixed, int *next) {
  __shared__ Shared sh;
  __shared__ int item;
  for (bool first_item = true;; first_item = false) {
    __syncthreads();
    if (threadIdx.x == 0) item = first_item && int(blockIdx.x) < nfixed ? int(blockIdx.x) : nfixed + g_add(next, 1);
    __syncthreads();
    if (item >= nqueue) return;
    const int2 q = queue[item];
    const int sample = __builtin_amdgcn_readfirstlane(q.x);
    const TraceArgs a = all[sample];  // uniform: scalar registers
    const int L = __builtin_amdgcn_readfirstlane(q.y);
    if (threadIdx.x < 26) sh.off[threadIdx.x] = nb_offset(int(threadIdx.x), a.d);
    const int seg = g_ld(a.seg_start + L), count = g_ld(a.seg_count + L);
    // lists 2 and 3 (the far lists) hold two entries per voxel: segment 2 * seg
    int *l2 = a.lists[2] + 2LL * seg, *l3 = a.lists[3] + 2LL * seg, *l4 = a.lists[4] + seg, *walk = a.lists[5] + seg;
    const Front na{a.lists[0] + seg, sh.front[0]}, nb{a.lists[1] + seg, sh.front[1]};
    unsigned long long *s0 = a.snaps[0] + seg, *s1 = a.snaps[1] + seg;
    if (threadIdx.x == 0) {
      sh.cursor = seg;
      sh.seg_end = (long long)seg + count;
      sh.pcount = 0;
      sh.after_used = 0;
      for (int j = 0; j < kProf; ++j) sh.prof[j] = 0;
      sh.tick = wall_clock64();
    }
    const long long label_begin = wall_clock64();
    __syncthreads();
    bool traceable = true;
    if (a.parent) {
      const int field = parent_field(a, sh, int(g_ld(a.root + L)), a.order + seg, count, na, nb, l2, l3, l4);
      lap(sh, 2);
      traceable = field == kNoRail;
      if (field == kRailFound && threadIdx.x == 0) g_st(a.ejected + g_add(a.counters + 2, 1) % a.ejected_cap, L);
      if ((field == kStalled || field == kOverflowed) && threadIdx.x == 0) report_failure(a, field);
    }
    // A soma's paths keep only the vertices beyond its radius (its ball died in soma_balls).
    const bool soma = a.soma && g_ld(&a.soma[L].radius) >= 0.0;
    for (; traceable;) {
      const int t = next_target(a, sh, L);
      lap(sh, 1);
      if (t < 0) break;
      int status = kRailFound;
      if (a.parent) {
        parent_walk(a, sh, t, count, walk);
        if (threadIdx.x == 0) sh.n_touched = 0;
      } else {
        status = railroad(a, sh, t, count, na, nb, l2, l3, l4);
        lap(sh, 2);
        if (status == kRailFound) backtrace(a, sh, t, count, walk);
      }
      __syncthreads();
      lap(sh, 3);
      const int nt = sh.n_touched;
      for (int j = threadIdx.x; j < nt; j += blockDim.x) {  // restore the railroad state
        sp_u64(&a.rec[g_ld(l4 + j)].word, kMaxKey);
      }
      if (status == kStalled || status == kOverflowed) {  // given up: the host raises
        if (threadIdx.x == 0) report_failure(a, status);
        __syncthreads();
        break;
      }
      if (status == kNoRail) {
        // No rail reachable from t (a component that filled holes split): as the reference's
        // per-component tracer does for such a label (trace.cu:198-215), an ordered target ends
        // the label with the paths it has, a manual one counts as an empty path.
        if (sh.ordered) break;
        if (threadIdx.x == 0) sh.path_len = 0;
        __syncthreads();
      } else if (sh.path_len < 0) {
        if (threadIdx.x == 0) g_st(a.ejected + g_add(a.counters + 2, 1) % a.ejected_cap, L);
        __syncthreads();
        break;
      }
      if (soma && threadIdx.x == 0) soma_trim(a, sh, walk, int(g_ld(a.root + L)), g_ld(&a.soma[L].radius));
      __syncthreads();
      const int len = sh.path_len;
      if (threadIdx.x == 0 && len > 0) {
        g_add(reinterpret_cast<unsigned long long *>(a.counters + 4), (unsigned long long)nt);
        int pos = g_add(a.counters + 0, len);
        int rec = g_add(a.counters + 1, 1);
        if ((long long)pos + len > a.store_cap || rec >= a.rec_cap) {
          g_or(a.counters + 3, 1);  // overflow: reported to the host
          sh.store_pos = -1;
        } else {
          sh.store_pos = pos;
          g_st(a.rec_label + rec, L);
          g_st(a.rec_len + rec, len);
          g_st(a.rec_start + rec, pos);
        }
      }
      __syncthreads();
      const int pos = sh.store_pos;
      for (int j = threadIdx.x; j < len; j += blockDim.x) {
        int v = g_ld(walk + j);
        if (pos >= 0) g_st(a.store + pos + (len - 1 - j), v);
        if (a.parent) {
          if (!soma) g_st(a.parent + v, -2);  // STOP: later walks end on this vertex (not for somata)
        }
        else sp_f(&a.rec[v].pdrf, 0.0f);       // the path becomes rail
      }
      __syncthreads();
      const bool converged = invalidate(a, a.ball, sh, walk, len, count, na, nb, l2, l3, l4, s0, s1);
      lap(sh, 4);
      if (!converged) break;  // reported by invalidate; the host raises
      if (threadIdx.x == 0) { sh.pcount += 1; sh.prof[7] += 1; }
      __syncthreads();
    }
    if (threadIdx.x == 0 && a.label_stats) {
      sh.prof[0] = (unsigned long long)(wall_clock64() - label_begin);
      for (int j = 0; j < kProf; ++j) g_st(a.label_stats + (long long)kProf * L + j, sh.prof[j]);
    }
  }
}

}  // namespace

void trace_all(Context &ctx, const std::vector<TraceArgs> &args, const std::vector<LabelItem> &items,
               const std::vector<LabelItem> &somata) {
  if (items.empty()) return;
  Buffer<TraceArgs> args_dev(ctx, args.size());
  args_dev.upload(ctx, args.data(), args.size());
  std::vector<int2> q(items.size());
  for (size_t i = 0; i < items.size(); ++i) q[i] = make_int2(items[i].sample, items[i].label);
  Buffer<int2> queue(ctx, q.size());
  queue.upload(ctx, q.data(), q.size());
  Buffer<int> next(ctx, 1);  // queue cursor
  next.clear(ctx);
  const int nq = int(q.size());
  // every workgroup takes labels from the queue until it is spent, so any grid drains it
  const int grid = std::min(nq, BROOK_TRACE_GRID > 0 ? BROOK_TRACE_GRID : ctx.compute_units() * 2);
  if (!somata.empty()) {
    std::vector<int2> sq(somata.size());
    for (size_t i = 0; i < somata.size(); ++i) sq[i] = make_int2(somata[i].sample, somata[i].label);
    Buffer<int2> soma_queue(ctx, sq.size());
    soma_queue.upload(ctx, sq.data(), sq.size());
    KernelTimer kt(ctx, "soma_balls");
    soma_balls<<<unsigned(sq.size()), kBlock, 0, ctx.stream()>>>(args_dev.data(), soma_queue.data(), int(sq.size()));
    BROOK_LAUNCHED();
  }
  {
    KernelTimer kt(ctx, "trace_labels");
    trace_labels<<<grid, kBlock, 0, ctx.stream()>>>(args_dev.data(), queue.data(), nq,
                                                    std::min(grid, ctx.compute_units()), next.data());
    BROOK_LAUNCHED();
  }
}

}  // namespace brook
