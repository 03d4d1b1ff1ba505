// Field preparation per component label (batched.cu:177-255 restated for gfx950):
//   * label voxel lists (segments laid out by id, from one wave-aggregated atomic per run),
//   * one workgroup per label from a device work queue: root-search flood + argmax, DAF flood +
//     argmax (label-gated geometric SSSP, the unique float32 min-plus fixpoint), no grid barriers,
//   * PDRF and the alive bit in one streaming pass, the root rails, the target-order keys.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fields.hpp"
#include "tuned.hpp"

namespace brook {
namespace {

// Voxel lists: vox[seg_start[l] + slot] = every voxel of id l (slot order is free; consumers
// that need an order sort the list or reduce over it). A workgroup walks kListChunk consecutive
// voxels twice: the first pass totals each label's voxels in LDS and reserves them with one global
// add per label, the second hands out places within the reservation per x-run.
constexpr int kListChunk = 16384;

template <class CC>
__global__ void __launch_bounds__(256) fill_segments(const CC *cc, int n, Dims d, const int *seg_start, int *fill,
                                                     int *vox) {
  __shared__ LabelSlots slots;
  __shared__ int cnt[kLabelSlots], base[kLabelSlots];
  slots.init();
  for (int s = threadIdx.x; s < kLabelSlots; s += 256) cnt[s] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long long begin = (long long)blockIdx.x * kListChunk;
  for (int t = 0; t < kListChunk; t += 256) {
    const long long g = begin + t + threadIdx.x;
    const bool valid = g < n;
    const int i = valid ? int(g) : 0;
    const CC l = valid ? cc[i] : CC(0);
    const WaveRun r = wave_run(l, valid, valid ? i % d.sx : 0, lane);
    if (!valid || l == CC(0) || !r.head) continue;
    const int s = slots.insert(int(l));
    if (s >= 0) atomicAdd(&cnt[s], r.len);
  }
  __syncthreads();
  for (int s = threadIdx.x; s < kLabelSlots; s += 256) {
    const int l = slots.tag[s];
    if (l >= 0) base[s] = atomicAdd(fill + l, cnt[s]);
    cnt[s] = 0;
  }
  __syncthreads();
  for (int t = 0; t < kListChunk; t += 256) {
    const long long g = begin + t + threadIdx.x;
    const bool valid = g < n;
    const int i = valid ? int(g) : 0;
    const CC l = valid ? cc[i] : CC(0);
    const WaveRun r = wave_run(l, valid, valid ? i % d.sx : 0, lane);
    int at = 0;
    if (valid && l != CC(0) && r.head) {
      const int s = slots.find(int(l));
      at = s >= 0 ? base[s] + atomicAdd(&cnt[s], r.len) : atomicAdd(fill + l, r.len);
    }
    at = __shfl(at, r.start_lane, 64);
    if (valid && l != CC(0)) vox[seg_start[l] + at + (lane - r.start_lane)] = i;
  }
}

// Voxel records: the neighbour mask (bit k = neighbour k of v is inside the volume and has v's
// nonzero id; every later neighbour walk reads this instead of 26 ids and bounds checks),
// PDRF +inf, word ~0. A workgroup covers a kRecX x kRecY x kRecZ box: its ids and a one-voxel
// halo (0 outside the volume, which never equals an id) are staged in LDS by coalesced row loads,
// the 26 comparisons of a voxel read LDS, and each wave stores one row of 64 records contiguously.
constexpr int kRecX = 64, kRecY = 4, kRecZ = 4;
template <class CC>
__global__ void __launch_bounds__(256) init_records(const CC *cc, Dims d, VoxelRec *rec) {
  constexpr int TX = kRecX + 2, TY = kRecY + 2, TZ = kRecZ + 2;
  __shared__ CC tile[TZ][TY][TX];
  const int bx = int(blockIdx.x) * kRecX, by = int(blockIdx.y) * kRecY, bz = int(blockIdx.z) * kRecZ;
  for (int t = threadIdx.x; t < TX * TY * TZ; t += 256) {
    const int tx = t % TX, ty = (t / TX) % TY, tz = t / (TX * TY);
    const int x = bx + tx - 1, y = by + ty - 1, z = bz + tz - 1;
    CC v = CC(0);
    if (x >= 0 && x < d.sx && y >= 0 && y < d.sy && z >= 0 && z < d.sz)
      v = cc[x + (long long)d.sx * (y + (long long)d.sy * z)];
    tile[tz][ty][tx] = v;
  }
  __syncthreads();
  const int tx = int(threadIdx.x) % kRecX, ty = int(threadIdx.x) / kRecX;
  const int x = bx + tx, y = by + ty;
  if (x >= d.sx || y >= d.sy) return;
  for (int tz = 0; tz < kRecZ && bz + tz < d.sz; ++tz) {
    const CC l = tile[tz + 1][ty + 1][tx + 1];
    unsigned m = 0;
    if (l != CC(0)) {
      for_each_nb([&](auto K) {
        constexpr int k = decltype(K)::value;
        if (tile[tz + 1 + kDZ[k]][ty + 1 + kDY[k]][tx + 1 + kDX[k]] == l) m |= 1u << k;
      });
    }
    VoxelRec r;
    r.pdrf = kInf;
    r.meta = m;
    r.word = kMaxKey;
    rec[x + (long long)d.sx * (y + (long long)d.sy * (bz + tz))] = r;
  }
}

__global__ void fill_float(float *p, long long n, float v) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

constexpr int kFloodBlock = BROOK_FLOOD_BLOCK;  // threads per label workgroup
// Near-far bands in the floods (see flood_bands); the band width is BROOK_FLOOD_BAND times the
// longest step.
#ifndef BROOK_FLOOD_NEARFAR
#define BROOK_FLOOD_NEARFAR 1
#endif
#ifndef BROOK_FLOOD_BAND
#define BROOK_FLOOD_BAND 2.0f
#endif
constexpr int kFloodBatch = BROOK_FLOOD_NB_BATCH;  // neighbours in flight per thread

// Relaxes masked neighbours B..E-1 of u (distance du): all loads, then all atomics, in flight
// together. Returns the neighbours this thread must queue (improved and found not queued).
template <int B, int E>
__device__ __forceinline__ unsigned flood_part(VoxelRec *rec, Dims d, int u, unsigned m, float du, const float *W) {
  unsigned seen[E - B];
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    seen[i] = (m >> k) & 1u ? rec_word_hi(rec, u + nb_offset<k>(d)) : 0u;
  });
  unsigned long long old[E - B];
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    const unsigned cb = __float_as_uint(du + W[k]);
    old[i] = 0;
    if (cb < seen[i]) old[i] = g_min(&rec[u + nb_offset<k>(d)].word, (unsigned long long)cb << 32);
  });
  unsigned push = 0;
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    const unsigned cb = __float_as_uint(du + W[k]);
    if ((old[i] & 1ULL) && old[i] > ((unsigned long long)cb << 32)) push |= 1u << k;
  });
  return push;
}

// Neighbours kb..kb+cnt-1 (those set in m; cnt <= kSpan) of u: flood_part with a run-time span.
__device__ __forceinline__ unsigned flood_span(VoxelRec *rec, int u, unsigned m, int kb, int cnt, float du, const float *W,
                                               const int *off) {
  constexpr int S = kSpan;
  unsigned seen[S];
  int v[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const bool on = j < cnt && ((m >> (kb + j)) & 1u);
    v[j] = u + off[on ? kb + j : 0];
    seen[j] = on ? rec_word_hi(rec, v[j]) : 0u;
  }
  unsigned long long old[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const unsigned cb = __float_as_uint(du + W[j < cnt ? kb + j : 0]);
    old[j] = 0;
    if (cb < seen[j]) old[j] = g_min(&rec[v[j]].word, (unsigned long long)cb << 32);
  }
  unsigned push = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const unsigned cb = __float_as_uint(du + W[j < cnt ? kb + j : 0]);
    if ((old[j] & 1ULL) && old[j] > ((unsigned long long)cb << 32)) push |= 1u << (kb + j);
  }
  return push;
}

// Geometric flood of one label from one source inside one workgroup, on the records' packed words
// (distance bits high, "not queued" bit 0 low; ~0 = +inf, not queued). Words of the label must be
// ~0 on entry; on exit they hold the distances, every voxel dequeued (bit 0 set).
// Label-correcting rounds: a voxel is dequeued by one atomic that also returns its distance, and
// an improvement that finds the voxel not queued queues it, so no update is ever lost.
// The front's words hold their start distances (queued) on entry; counts[0] is its length.
// Returns false when the rounds exceed `budget` (far above what the label-correcting rounds need:
// a logic error then ends as a host-visible error instead of a hang).
__device__ __forceinline__ bool flood_rounds(VoxelRec *rec, Dims d, const float *W, const int *off, int *cur, int *nxt,
                                             int *counts /*LDS [2]*/, long long budget) {
  for (long long round = 0;; ++round) {
    const int nc = counts[0];
    if (nc == 0) break;
    if (round >= budget) return false;  // uniform: counts[0] and round are the same in every thread
    if (nc <= kFloodBlock / 4) {  // a group of lanes per front voxel; its first lane dequeues
      const Group q = group_of(nc, kFloodBlock);
      int u = 0;
      unsigned m = 0;
      float du = 0.0f;
      if (q.g == 0 && q.i < nc) {
        u = lp_i(cur + q.i);
        m = rec_meta(rec, u) & kNbrBits;
        du = __uint_as_float((unsigned)(g_or(&rec[u].word, 1ULL) >> 32));
      }
      u = __shfl(u, q.lead, 64);
      m = __shfl(m, q.lead, 64);
      du = __shfl(du, q.lead, 64);
      const unsigned push = q.i < nc ? flood_span(rec, u, m, q.kb, q.cnt, du, W, off) : 0u;
      emit_neighbours(nxt, &counts[1], push, u, off);
    } else {
      for (int base = 0; base < nc; base += blockDim.x) {  // uniform: every lane reaches the append
        const int i = base + threadIdx.x;
        unsigned push = 0;
        int u = 0;
        if (i < nc) {
          u = lp_i(cur + i);
          const unsigned m = rec_meta(rec, u) & kNbrBits;
          const float du = __uint_as_float((unsigned)(g_or(&rec[u].word, 1ULL) >> 32));
          for_parts<kFloodBatch>(m, [&](auto B, auto E) {
            push |= flood_part<decltype(B)::value, decltype(E)::value>(rec, d, u, m, du, W);
          });
        }
        emit_neighbours(nxt, &counts[1], push, u, off);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) { counts[0] = counts[1]; counts[1] = 0; }
    int *t = cur; cur = nxt; nxt = t;
    __syncthreads();
  }
  return true;
}

// Rounds one flood of a label of `count` voxels may take (see flood_rounds).
__device__ __forceinline__ long long flood_budget(int count) { return 4LL * count + 4096; }

// ---- near-far bands -------------------------------------------------------------------------
// The same fixpoint, with the label-correcting rounds kept inside a distance band [.., T]: an
// improvement beyond T goes to a far list instead of the next round, and when the band has no
// front left, T moves to the nearest far distance plus the band width and the far voxels inside
// it start the next band. A voxel is then relaxed about once instead of re-queued each time a
// cheaper path with more steps arrives (30 % of pops on the 512^3 workload with anisotropy
// (8, 8, 40)). Word bits as the railroad's: bit 0 "not queued", bit 1 "far listed"; near
// candidates store 00 (queued), far candidates 11; a near improvement of a far-listed voxel
// clears its far bit (its far entry goes stale). Far lists hold 2 * count entries (one append per
// voxel per round; a list longer than count after a round is compacted to one entry per voxel).
struct FloodShared {
  int near[2];          // front lengths: current, next
  int n_far, n_farnxt;  // far list lengths
  unsigned minfar;      // smallest far distance (float bits) since the band moved
  float T;              // the band's upper end
  int overflow;         // a far append found the list full (never expected)
};

__device__ __forceinline__ bool flood_far_listed(unsigned long long w) {
  return (w & 2ULL) && (unsigned)(w >> 32) != 0xFFFFFFFFu;
}

// Masked neighbours B..E-1 of u: the near and far masks to append, the smallest far distance.
template <int B, int E>
__device__ __forceinline__ void band_part(VoxelRec *rec, Dims d, int u, unsigned m, float du, const float *W, float T,
                                          unsigned &near, unsigned &far, unsigned &minfar) {
  unsigned seen[E - B];
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    seen[i] = (m >> k) & 1u ? rec_word_hi(rec, u + nb_offset<k>(d)) : 0u;
  });
  unsigned long long old[E - B];
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    const float c = du + W[k];
    const unsigned cb = __float_as_uint(c);
    old[i] = 0;
    if (cb < seen[i]) old[i] = g_min(&rec[u + nb_offset<k>(d)].word, ((unsigned long long)cb << 32) | (c <= T ? 0ULL : 3ULL));
  });
  for_nb<B, E>([&](auto K) {
    constexpr int k = decltype(K)::value, i = k - B;
    const float c = du + W[k];
    const unsigned cb = __float_as_uint(c);
    const bool in = c <= T;
    if (!(old[i] > (((unsigned long long)cb << 32) | (in ? 0ULL : 3ULL)))) return;
    if (in) {
      if (old[i] & 1ULL) near |= 1u << k;
    } else {
      minfar = min(minfar, cb);
      if (!flood_far_listed(old[i])) far |= 1u << k;
    }
  });
}

// Neighbours kb..kb+cnt-1 (those set in m; cnt <= kSpan) of u: band_part with a run-time span.
__device__ __forceinline__ void band_span(VoxelRec *rec, int u, unsigned m, int kb, int cnt, float du, const float *W,
                                          const int *off, float T, unsigned &near, unsigned &far, unsigned &minfar) {
  constexpr int S = kSpan;
  unsigned seen[S];
  int v[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const bool on = j < cnt && ((m >> (kb + j)) & 1u);
    v[j] = u + off[on ? kb + j : 0];
    seen[j] = on ? rec_word_hi(rec, v[j]) : 0u;
  }
  unsigned long long old[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const float c = du + W[j < cnt ? kb + j : 0];
    const unsigned cb = __float_as_uint(c);
    old[j] = 0;
    if (cb < seen[j]) old[j] = g_min(&rec[v[j]].word, ((unsigned long long)cb << 32) | (c <= T ? 0ULL : 3ULL));
  }
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const float c = du + W[j < cnt ? kb + j : 0];
    const unsigned cb = __float_as_uint(c);
    const bool in = c <= T;
    if (!(old[j] > (((unsigned long long)cb << 32) | (in ? 0ULL : 3ULL)))) continue;
    if (in) {
      if (old[j] & 1ULL) near |= 1u << (kb + j);
    } else {
      minfar = min(minfar, cb);
      if (!flood_far_listed(old[j])) far |= 1u << (kb + j);
    }
  }
}

__device__ __forceinline__ void fold_minfar(FloodShared &fs, unsigned m) {
  for (int o = 32; o > 0; o >>= 1) m = min(m, (unsigned)__shfl_xor((int)m, o, 64));
  if (__lane_id() == 0 && m != 0x7f800000u) atomicMin(&fs.minfar, m);
}

// Appends u + off[k] for every set bit k of mask to a list of `cap` entries (all lanes active);
// a wave whose appends would pass the end writes nothing and raises `overflow`.
__device__ __forceinline__ void emit_capped_list(int *list, int *counter, unsigned mask, int u, const int *off, int cap,
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

// The near list `cur` (fs.near[0] queued entries) starts the flood; far lists hold 2 * count.
__device__ __forceinline__ bool flood_bands(VoxelRec *rec, Dims d, const float *W, const int *off, int *cur, int *nxt,
                                            int *fcur, int *fnxt, FloodShared &fs, int count, long long budget) {
  float wmax = 0.0f;
  for (int k = 0; k < 26; ++k) wmax = fmaxf(wmax, W[k]);
  const float width = BROOK_FLOOD_BAND * wmax;
  if (threadIdx.x == 0) {
    fs.near[1] = fs.n_far = fs.n_farnxt = fs.overflow = 0;
    fs.minfar = 0x7f800000u;
    fs.T = width;
  }
  __syncthreads();
  long long rounds = 0;
  for (;;) {
    for (;;) {  // the band's rounds
      const int nc = fs.near[0];
      if (nc == 0) break;
      if (++rounds > budget || fs.overflow) return false;  // uniform: read after the round's barrier
      const float T = fs.T;
      unsigned minfar = 0x7f800000u;
      if (nc <= kFloodBlock / 4) {  // a group of lanes per front voxel; its first lane dequeues
        const Group q = group_of(nc, kFloodBlock);
        int u = 0;
        unsigned m = 0;
        float du = 0.0f;
        if (q.g == 0 && q.i < nc) {
          u = lp_i(cur + q.i);
          m = rec_meta(rec, u) & kNbrBits;
          du = __uint_as_float((unsigned)(g_or(&rec[u].word, 1ULL) >> 32));
        }
        u = __shfl(u, q.lead, 64);
        m = __shfl(m, q.lead, 64);
        du = __shfl(du, q.lead, 64);
        unsigned near = 0, far = 0;
        if (q.i < nc) band_span(rec, u, m, q.kb, q.cnt, du, W, off, T, near, far, minfar);
        emit_neighbours(nxt, &fs.near[1], near, u, off);
        emit_capped_list(fcur, &fs.n_far, far, u, off, 2 * count, &fs.overflow);
      } else {
        for (int base = 0; base < nc; base += blockDim.x) {  // uniform: every lane reaches the appends
          const int i = base + threadIdx.x;
          unsigned near = 0, far = 0;
          int u = 0;
          if (i < nc) {
            u = lp_i(cur + i);
            const unsigned m = rec_meta(rec, u) & kNbrBits;
            const float du = __uint_as_float((unsigned)(g_or(&rec[u].word, 1ULL) >> 32));
            for_parts<kFloodBatch>(m, [&](auto B, auto E) {
              band_part<decltype(B)::value, decltype(E)::value>(rec, d, u, m, du, W, T, near, far, minfar);
            });
          }
          emit_neighbours(nxt, &fs.near[1], near, u, off);
          emit_capped_list(fcur, &fs.n_far, far, u, off, 2 * count, &fs.overflow);
        }
      }
      fold_minfar(fs, minfar);
      __syncthreads();
      if (threadIdx.x == 0) { fs.near[0] = fs.near[1]; fs.near[1] = 0; }
      int *t = cur; cur = nxt; nxt = t;
      __syncthreads();
      if (fs.n_far > count) {  // one entry per far-listed voxel (the one whose atomic clears its far bit)
        const int nf = fs.n_far;
        for (int i = threadIdx.x; i < nf; i += blockDim.x) {
          const int w = lp_i(fcur + i);
          if (flood_far_listed(g_and(&rec[w].word, ~2ULL))) g_st(fnxt + wave_slot(&fs.n_farnxt), w);
        }
        __syncthreads();
        const int kept = fs.n_farnxt;
        for (int i = threadIdx.x; i < kept; i += blockDim.x) g_or(&rec[lp_i(fnxt + i)].word, 2ULL);
        __syncthreads();
        if (threadIdx.x == 0) { fs.n_far = kept; fs.n_farnxt = 0; }
        int *tf = fcur; fcur = fnxt; fnxt = tf;
        __syncthreads();
      }
    }
    const int nf = fs.n_far;
    if (nf == 0) return true;
    if (++rounds > budget || fs.overflow) return false;
    __syncthreads();
    if (threadIdx.x == 0) {  // the band moves to the nearest far distance
      fs.T = __uint_as_float(fs.minfar) + width;
      fs.minfar = 0x7f800000u;
    }
    __syncthreads();
    const float T = fs.T;
    unsigned minfar = 0x7f800000u;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {  // far voxels inside the band start it
      const int w = lp_i(fcur + i);
      if (!flood_far_listed(lp_u64(&rec[w].word))) continue;
      const unsigned long long was = g_and(&rec[w].word, ~3ULL);
      if (!flood_far_listed(was)) continue;
      const float dw = __uint_as_float((unsigned)(was >> 32));
      if (dw <= T) {
        g_st(cur + wave_slot(&fs.near[0]), w);
      } else {
        g_or(&rec[w].word, 3ULL);
        minfar = min(minfar, (unsigned)(was >> 32));
        g_st(fnxt + wave_slot(&fs.n_farnxt), w);
      }
    }
    fold_minfar(fs, minfar);
    __syncthreads();
    if (threadIdx.x == 0) { fs.n_far = fs.n_farnxt; fs.n_farnxt = 0; }
    int *tf = fcur; fcur = fnxt; fnxt = tf;
    __syncthreads();
  }
}

// The flood from the queued front `cur` (fs.near[0] entries).
__device__ __forceinline__ bool flood_front(VoxelRec *rec, Dims d, const float *W, const int *off, int *cur, int *nxt,
                                            int *fcur, int *fnxt, FloodShared &fs, int count) {
#if BROOK_FLOOD_NEARFAR
  return flood_bands(rec, d, W, off, cur, nxt, fcur, fnxt, fs, count, flood_budget(count));
#else
  return flood_rounds(rec, d, W, off, cur, nxt, fs.near, flood_budget(count));
#endif
}

__device__ __forceinline__ bool flood_label(VoxelRec *rec, Dims d, int src, const float *W, const int *off, int *cur,
                                            int *nxt, int *fcur, int *fnxt, FloodShared &fs, int count) {
  if (threadIdx.x == 0) {
    sp_u64(&rec[src].word, 0ULL);  // distance 0, queued
    g_st(cur, src);
    fs.near[0] = 1;
    fs.near[1] = 0;
  }
  __syncthreads();
  return flood_front(rec, d, W, off, cur, nxt, fcur, fnxt, fs, count);
}


// The free-space length of a soma voxel at offset (dxi, dyi, dzi) from the root, when its physical
// distance is within `radius`: the shortest 26-connected path through empty space, in float32 in
// the order of freespace_seed_lab (batched_regular.inc:51-73).
__device__ __forceinline__ bool free_space(int dxi, int dyi, int dzi, float wx, float wy, float wz, float radius, float &val) {
  const float rdx = wx * (float)dxi, rdy = wy * (float)dyi, rdz = wz * (float)dzi;
  if (rdx * rdx + rdy * rdy + rdz * rdz > radius * radius) return false;
  const float dx = fabsf((float)dxi), dy = fabsf((float)dyi), dz = fabsf((float)dzi);
  const float dxyz = fminf(fminf(dx, dy), dz);
  const float dxy = fminf(dx, dy), dyz = fminf(dy, dz), dxz = fminf(dx, dz);
  val = dxyz * sqrtf(wx * wx + wy * wy + wz * wz) + wx * (dx - dxyz) + wy * (dy - dxyz) + wz * (dz - dxyz) +
        (dxy - dxyz) * (sqrtf(wx * wx + wy * wy) - wx - wy) + (dxz - dxyz) * (sqrtf(wx * wx + wz * wz) - wx - wz) +
        (dyz - dxyz) * (sqrtf(wy * wy + wz * wz) - wy - wz);
  return true;
}

// A soma's DAF flood (batched.cu:222-231): every voxel of the label inside the free-space ball
// starts at its free-space length (the root at 0), then the label-correcting rounds.
__device__ __forceinline__ bool flood_free_space(VoxelRec *rec, Dims d, const int *list, int cnt, int root, float radius,
                                                 const float *W, const int *off, int *cur, int *nxt, int *fcur, int *fnxt,
                                                 FloodShared &fs) {
  if (threadIdx.x == 0) { fs.near[0] = 0; fs.near[1] = 0; }
  __syncthreads();
  int rx, ry, rz;
  unravel(root, d, rx, ry, rz);
  for (int j = threadIdx.x; j < cnt; j += blockDim.x) {
    const int v = g_ld(list + j);
    int x, y, z;
    unravel(v, d, x, y, z);
    float val;
    if (!free_space(x - rx, y - ry, z - rz, W[0], W[2], W[4], radius, val)) continue;
    sp_u64(&rec[v].word, (unsigned long long)__float_as_uint(val) << 32);  // queued
    g_st(cur + wave_slot(&fs.near[0]), v);
  }
  __syncthreads();
  return flood_front(rec, d, W, off, cur, nxt, fcur, fnxt, fs, cnt);
}

// Argmax of the flooded distance over the label's voxel list: largest finite value, ties -> lowest
// flat index (the packed key of batched_regular.inc:9-22). Returns the key (0: none).
__device__ __forceinline__ unsigned long long label_argmax(const VoxelRec *rec, const int *list, int count, unsigned long long *lds) {
  if (threadIdx.x == 0) *lds = 0;
  __syncthreads();
  unsigned long long best = 0;
  for (int j = threadIdx.x; j < count; j += blockDim.x) {
    int v = g_ld(list + j);
    unsigned fb = word_dist_bits(lp_u64(&rec[v].word));
    if (!(fb < 0x7f800000u)) continue;
    unsigned long long key = ((unsigned long long)fb << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)v);
    best = key > best ? key : best;
  }
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_down(best, off, 64);
    best = o > best ? o : best;
  }
  if ((threadIdx.x & 63) == 0 && best) atomicMax(lds, best);
  __syncthreads();
  unsigned long long r = *lds;
  __syncthreads();
  return r;
}

// stalled: set when a flood exceeded its round budget (the label's fields are then left as they
// are and the host raises). The queue is sorted largest first: workgroup w < nfixed (resident from
// the start, at most the queue's length) takes item w first (workgroups go to the XCDs in turn, so
// the largest labels start on different XCDs); every other item comes from the shared cursor,
// from item nfixed on.
__global__ void __launch_bounds__(kFloodBlock) prepare_labels(const FloodJob *__restrict__ jobs,
                                                              const int2 *__restrict__ queue, int nqueue, int nfixed,
                                                              int *next, int *stalled) {
  __shared__ float W[26];
  __shared__ int off[26];
  __shared__ FloodShared fs;
  __shared__ unsigned long long best;
  __shared__ int item;
  for (bool first_item = true;; first_item = false) {
    __syncthreads();
    if (threadIdx.x == 0) item = first_item && int(blockIdx.x) < nfixed ? int(blockIdx.x) : nfixed + atomicAdd(next, 1);
    __syncthreads();
    if (item >= nqueue) return;
    const int2 q = queue[item];
    const FloodJob job = jobs[__builtin_amdgcn_readfirstlane(q.x)];  // uniform: scalar registers
    const int L = __builtin_amdgcn_readfirstlane(q.y);
    const Dims d = job.d;
    const LabelTables &t = job.t;
    if (threadIdx.x < 26) {
      W[threadIdx.x] = g_ld(t.weights + threadIdx.x);
      off[threadIdx.x] = nb_offset(int(threadIdx.x), d);
    }
    __syncthreads();
    const int s0 = g_ld(t.seg_start + L), cnt = g_ld(t.seg_count + L);
    const int *list = t.vox + s0;
    int *a = job.lists_a + s0, *b = job.lists_b + s0, *fa = job.far_a + 2LL * s0, *fb = job.far_b + 2LL * s0;
    long long root = g_ld(t.root + L);
    if (root < 0) {  // no border target: flood from the component's first voxel, root = argmax
      if (!flood_label(t.rec, d, int(g_ld(t.first + L)), W, off, a, b, fa, fb, fs, cnt)) {
        if (threadIdx.x == 0) atomicOr(stalled, 1);
        continue;
      }
      unsigned long long key = label_argmax(t.rec, list, cnt, &best);
      root = key ? (long long)(0xFFFFFFFFu - (unsigned)key) : -1;
      for (int j = threadIdx.x; j < cnt; j += blockDim.x) sp_u64(&t.rec[g_ld(list + j)].word, kMaxKey);
      __syncthreads();
      if (threadIdx.x == 0) g_st(t.root + L, root);
      if (job.root_only) continue;
    }
    const float radius = t.free_radius ? g_ld(t.free_radius + L) : 0.0f;
    const bool converged = radius > 0.0f ? flood_free_space(t.rec, d, list, cnt, int(root), radius, W, off, a, b, fa, fb, fs)
                                         : flood_label(t.rec, d, int(root), W, off, a, b, fa, fb, fs, cnt);
    if (!converged) {
      if (threadIdx.x == 0) atomicOr(stalled, 1);
      continue;
    }
    unsigned long long key = label_argmax(t.rec, list, cnt, &best);
    if (threadIdx.x == 0) {
      g_st(t.target + L, key ? (long long)(0xFFFFFFFFu - (unsigned)key) : -1LL);
      g_st(t.max_daf + L, key ? __uint_as_float((unsigned)(key >> 32)) : 0.0f);
    }
    for (int j = threadIdx.x; j < cnt; j += blockDim.x) {  // DAF out, words back to ~0
      const int v = g_ld(list + j);
      g_st(job.daf + v, __uint_as_float(word_dist_bits(lp_u64(&t.rec[v].word))));
      sp_u64(&t.rec[v].word, kMaxKey);
    }
  }
}

// PDRF (batched_regular.inc:93-108) and the alive mask in one pass. use[l] 0: not traced.
template <class CC>
__global__ void __launch_bounds__(256) pdrf_alive(const CC *cc, const float *dbf, const float *daf, const float *M,
                                                  const float *inv, const unsigned char *use, int squarings, float scale,
                                                  long long n, float *pdrf, VoxelRec *rec) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  CC l = cc[i];
  unsigned char u = use[l];
  if (l == CC(0) || u == 0) {
    if (pdrf) pdrf[i] = kInf;
    return;
  }
  float dv = dbf[i];
  if (dv == 0.0f) dv = kInf;
  float p = 1.0f - dv * M[l];
  for (int k = 0; k < squarings; ++k) p = p * p;
  p = p * scale;
  if (u == 2) {
    float a = daf[i];
    if (isinf(a)) a = 0.0f;
    p = p + a * inv[l];
  }
  if (pdrf) pdrf[i] = p;
  // the record's first half (PDRF, meta) in one 8-byte load and store
  unsigned long long *head = reinterpret_cast<unsigned long long *>(rec + i);
  const unsigned meta = unsigned(*head >> 32) | kAlive;
  *head = ((unsigned long long)meta << 32) | __float_as_uint(p);
}

__global__ void set_rails(VoxelRec *rec, const long long *root, const int *labels, int count) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < count) rec[root[labels[i]]].pdrf = 0.0f;
}

unsigned grid_for(long long n) { return unsigned((n + 255) / 256); }
unsigned blocks_of(long long n, int per) { return unsigned((n + per - 1) / per); }

}  // namespace

void build_segments(Context &ctx, const void *cc, int cc_bytes, Dims d, const std::vector<int> &seg_start,
                    Buffer<int> &seg_start_dev, Buffer<int> &vox) {
  const int n = d.sx * d.sy * d.sz;
  seg_start_dev = Buffer<int>(ctx, seg_start.size());
  seg_start_dev.upload(ctx, seg_start.data(), seg_start.size());
  Buffer<int> fill(ctx, seg_start.size());
  fill.clear(ctx);
  vox = Buffer<int>(ctx, std::max(1, seg_start.back()));
  if (cc_bytes == 1)
    fill_segments<uint8_t><<<blocks_of(n, kListChunk), 256, 0, ctx.stream()>>>(static_cast<const uint8_t *>(cc), n, d,
                                                                  seg_start_dev.data(), fill.data(), vox.data());
  else if (cc_bytes == 2)
    fill_segments<uint16_t><<<blocks_of(n, kListChunk), 256, 0, ctx.stream()>>>(static_cast<const uint16_t *>(cc), n, d,
                                                                   seg_start_dev.data(), fill.data(), vox.data());
  else
    fill_segments<uint32_t><<<blocks_of(n, kListChunk), 256, 0, ctx.stream()>>>(static_cast<const uint32_t *>(cc), n, d,
                                                                   seg_start_dev.data(), fill.data(), vox.data());
  BROOK_LAUNCHED();
}

void voxel_records(Context &ctx, const void *cc, int cc_bytes, Dims d, VoxelRec *rec) {
  const long long n = (long long)d.sx * d.sy * d.sz;
  if (n == 0) return;
  KernelTimer kt(ctx, "init_records");
  if ((d.sy + kRecY - 1) / kRecY > 65535 || (d.sz + kRecZ - 1) / kRecZ > 65535)
    throw std::invalid_argument("volume too tall for the record tiles");
  const dim3 g(unsigned((d.sx + kRecX - 1) / kRecX), unsigned((d.sy + kRecY - 1) / kRecY), unsigned((d.sz + kRecZ - 1) / kRecZ));
  if (cc_bytes == 1)
    init_records<uint8_t><<<g, 256, 0, ctx.stream()>>>(static_cast<const uint8_t *>(cc), d, rec);
  else if (cc_bytes == 2)
    init_records<uint16_t><<<g, 256, 0, ctx.stream()>>>(static_cast<const uint16_t *>(cc), d, rec);
  else
    init_records<uint32_t><<<g, 256, 0, ctx.stream()>>>(static_cast<const uint32_t *>(cc), d, rec);
  BROOK_LAUNCHED();
}

void fill_value(Context &ctx, float *p, long long n, float v) {
  if (n <= 0) return;
  fill_float<<<grid_for(n), 256, 0, ctx.stream()>>>(p, n, v);
  BROOK_LAUNCHED();
}

void prepare_fields(Context &ctx, const std::vector<FloodJob> &jobs, const std::vector<LabelItem> &items, int max_grid) {
  if (items.empty()) return;
  Buffer<FloodJob> jobs_dev(ctx, jobs.size());
  jobs_dev.upload(ctx, jobs.data(), jobs.size());
  std::vector<int2> q(items.size());
  for (size_t i = 0; i < items.size(); ++i) q[i] = make_int2(items[i].sample, items[i].label);
  Buffer<int2> queue(ctx, q.size());
  queue.upload(ctx, q.data(), q.size());
  Buffer<int> next(ctx, 2);  // [0] queue cursor, [1] stalled
  next.clear(ctx);
  const int nq = int(q.size());
  // every workgroup takes labels from the queue until it is spent, so any grid drains it
  int grid = std::min(nq, BROOK_FLOOD_GRID > 0 ? BROOK_FLOOD_GRID : ctx.compute_units() * 2);
  if (max_grid > 0) grid = std::min(grid, max_grid);
  {
    KernelTimer kt(ctx, "prepare_labels");
    prepare_labels<<<grid, kFloodBlock, 0, ctx.stream()>>>(jobs_dev.data(), queue.data(), nq,
                                                           std::min(grid, ctx.compute_units()), next.data(),
                                                           next.data() + 1);
    BROOK_LAUNCHED();
  }
  if (next.download(ctx)[1]) throw std::runtime_error("flood did not converge: a label exceeded its round budget");
}

void pdrf_and_alive(Context &ctx, const void *cc, int cc_bytes, long long n, const float *dbf, const float *daf,
                    const float *M, const float *inv, const unsigned char *use, int squarings, float scale, float *pdrf,
                    VoxelRec *rec) {
  KernelTimer kt(ctx, "pdrf_alive");
  if (cc_bytes == 1)
    pdrf_alive<uint8_t><<<grid_for(n), 256, 0, ctx.stream()>>>(static_cast<const uint8_t *>(cc), dbf, daf, M, inv, use,
                                                               squarings, scale, n, pdrf, rec);
  else if (cc_bytes == 2)
    pdrf_alive<uint16_t><<<grid_for(n), 256, 0, ctx.stream()>>>(static_cast<const uint16_t *>(cc), dbf, daf, M, inv, use,
                                                                squarings, scale, n, pdrf, rec);
  else
    pdrf_alive<uint32_t><<<grid_for(n), 256, 0, ctx.stream()>>>(static_cast<const uint32_t *>(cc), dbf, daf, M, inv, use,
                                                                squarings, scale, n, pdrf, rec);
  BROOK_LAUNCHED();
}

void root_rails(Context &ctx, VoxelRec *rec, const long long *root, const int *labels, int count) {
  if (!count) return;
  set_rails<<<grid_for(count), 256, 0, ctx.stream()>>>(rec, root, labels, count);
  BROOK_LAUNCHED();
}

}  // namespace brook
