// Multi-label anisotropic Euclidean distance transform on gfx950, with the reference's rounding
// rule (edt_kernels.cuh:15-144, edt.cu:12-23): squared x-distances from a float32 accumulation
// table, y/z lower envelopes with float64 intersections, sqrtf at the end.
//
// Layout for HBM: the x pass gives each row to one wave (64 consecutive voxels per load, run
// boundaries from two ballots, no atomics); the y and z passes give each line to one thread with
// neighbouring threads on neighbouring x, so every step of every thread is one coalesced wave
// load. Background runs are never evaluated (their values are discarded by the finish rule) and
// the last pass fuses sqrt, the background zero and the per-label maximum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "stages.hpp"

namespace brook {
namespace {

__device__ __forceinline__ unsigned long long mask_upto(int lane) { return lane == 63 ? ~0ULL : ((2ULL << lane) - 1ULL); }

template <class CC, int MAXCH>
__global__ void __launch_bounds__(256) edt_x(const CC *cc, int sx, int rows, const float *acc, int bb, float *D) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= rows) return;  // whole wave
  const long long base = (long long)row * sx;
  const int nch = (sx + 63) >> 6;
  CC lab[MAXCH];
  int cst[MAXCH];
  int carry = 0;
  CC tail = CC(0);
#pragma unroll
  for (int j = 0; j < MAXCH; ++j) {
    if (j < nch) {
      int p = j * 64 + lane;
      bool valid = p < sx;
      CC l = valid ? cc[base + p] : CC(0);
      CC prev = shfl_up_any(l, 1);
      if (lane == 0) prev = tail;
      bool start = valid && (p == 0 || prev != l);
      unsigned long long S = __ballot(start);
      unsigned long long below = S & mask_upto(lane);
      cst[j] = below ? j * 64 + 63 - __clzll((long long)below) : carry;
      if (S) carry = j * 64 + 63 - __clzll((long long)S);
      lab[j] = l;
      tail = CC(__shfl((unsigned long long)l, 63, 64));
    }
  }
  int carry_d = sx - 1;
  CC head = CC(0);
#pragma unroll
  for (int jj = MAXCH - 1; jj >= 0; --jj) {
    if (jj < nch) {
      int p = jj * 64 + lane;
      bool valid = p < sx;
      CC l = lab[jj];
      CC next = CC(__shfl_down((unsigned long long)l, 1, 64));
      if (lane == 63) next = head;
      bool end = valid && (p == sx - 1 || next != l);
      unsigned long long E = __ballot(end);
      unsigned long long above = lane == 0 ? E : (E & ~((1ULL << lane) - 1ULL));
      int dd = above ? jj * 64 + __ffsll((long long)above) - 1 : carry_d;
      if (E) carry_d = jj * 64 + __ffsll((long long)E) - 1;
      head = CC(__shfl((unsigned long long)l, 0, 64));
      if (valid && l != CC(0)) {
        int c = cst[jj];
        float best = kInf;
        if (c > 0 || bb) { float t = acc[p - c + 1]; best = t * t; }
        if (dd < sx - 1 || bb) { float t = acc[dd + 1 - p]; float v = t * t; if (v < best) best = v; }
        D[base + p] = best;
      }
    }
  }
}

// Serial x pass for rows longer than the wave path handles (one thread per row).
template <class CC>
__global__ void __launch_bounds__(256) edt_x_serial(const CC *cc, int sx, int rows, const float *acc, int bb, float *D) {
  int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= rows) return;
  long long base = (long long)row * sx;
  int i = 0;
  while (i < sx) {
    CC l = cc[base + i];
    int c = i, d = i;
    while (d + 1 < sx && cc[base + d + 1] == l) d++;
    if (l != CC(0))
      for (int p = c; p <= d; ++p) {
        float best = kInf;
        if (c > 0 || bb) { float t = acc[p - c + 1]; best = t * t; }
        if (d < sx - 1 || bb) { float t = acc[d + 1 - p]; float v = t * t; if (v < best) best = v; }
        D[base + p] = best;
      }
    i = d + 1;
  }
}

// Lower envelope along y (axis 1) or z (axis 2), one thread per line, neighbouring threads on
// neighbouring x. The line is read in chunks of kChunk voxels (all loads of a chunk in flight at
// once, staged in the thread's own LDS slots). The envelope stack (position, height) lives in LDS
// for its first kStack entries and spills to a global column for deeper stacks; the float64
// intersection Z[j] of entries j-1 and j is recomputed from them when needed -- the same
// expression on the same operands as when the entry was pushed, so the same bits.
constexpr int kStack = 20, kChunk = 8, kFhBlock = 256;

template <class CC>
__global__ void __launch_bounds__(kFhBlock) edt_fh(const CC *cc, Dims dm, int axis, float w, int bb, int nlines,
                                                   long long spill_pitch, int *spill_v, float *spill_h, float *D,
                                                   int finish, unsigned *label_max) {
  __shared__ int sV[kStack * kFhBlock];
  __shared__ float sH[kStack * kFhBlock];
  __shared__ CC sL[kChunk * kFhBlock];
  __shared__ float sD[kChunk * kFhBlock];
  const int tid = threadIdx.x;
  const int t = blockIdx.x * kFhBlock + tid;
  if (t >= nlines) return;
  long long base, stride;
  int n;
  if (axis == 1) {
    int x = t % dm.sx, z = t / dm.sx;
    base = x + (long long)dm.sx * dm.sy * z;
    n = dm.sy;
    stride = dm.sx;
  } else {
    base = t;
    n = dm.sz;
    stride = (long long)dm.sx * dm.sy;
  }
  const double w2 = (double)(w * w);
  auto getV = [&](int k) -> int { return k < kStack ? sV[k * kFhBlock + tid] : spill_v[(k - kStack) * spill_pitch + t]; };
  auto getH = [&](int k) -> float { return k < kStack ? sH[k * kFhBlock + tid] : spill_h[(k - kStack) * spill_pitch + t]; };
  auto set = [&](int k, int v, float h) {
    if (k < kStack) { sV[k * kFhBlock + tid] = v; sH[k * kFhBlock + tid] = h; }
    else { spill_v[(k - kStack) * spill_pitch + t] = v; spill_h[(k - kStack) * spill_pitch + t] = h; }
  };
  // intersection of entries j-1 and j (j >= 1), as computed when entry j was placed
  auto inter = [&](int j) -> double {
    double v1 = (double)getV(j - 1), h1 = (double)getH(j - 1), v2 = (double)getV(j), h2 = (double)getH(j);
    return ((h2 + w2 * v2 * v2) - (h1 + w2 * v1 * v1)) / (2.0 * w2 * (v2 - v1));
  };
  int k = -1, c = 0;
  bool active = false;
  CC cur = CC(0);
  auto place = [&](int pp, double ph) {
    bool placed = false;
    while (k >= 0 && !placed) {
      double vk = (double)getV(k), hk = (double)getH(k);
      double s = ((ph + w2 * (double)pp * (double)pp) - (hk + w2 * vk * vk)) / (2.0 * w2 * ((double)pp - vk));
      double zk = k == 0 ? -1e300 : inter(k);
      if (s <= zk) {
        k--;
      } else {
        k++;
        set(k, pp, (float)ph);
        placed = true;
      }
    }
    if (!placed) { k = 0; set(0, pp, (float)ph); }
  };
  auto close_run = [&](int d, bool right_wall) {
    if (right_wall) place(d + 1, 0.0);
    float run_max = 0.0f;
    if (k < 0) {
      for (int q = c; q <= d; ++q) D[base + q * stride] = kInf;
      run_max = kInf;
    } else {
      int kk = 0;
      double znext = k >= 1 ? inter(1) : 0.0;
      int vkk = getV(0);
      float hkk = getH(0);
      for (int q = c; q <= d; ++q) {
        while (kk < k && znext < (double)q) {
          kk++;
          vkk = getV(kk);
          hkk = getH(kk);
          if (kk < k) znext = inter(kk + 1);
        }
        double dq = (double)q - (double)vkk;
        float v = (float)(w2 * dq * dq + (double)hkk);
        if (finish) { v = sqrtf(v); run_max = fmaxf(run_max, v); }
        D[base + q * stride] = v;
      }
    }
    if (finish && label_max && run_max < kInf) {
      unsigned bits = __float_as_uint(run_max);
      unsigned *slot = label_max + (unsigned long long)cur;
      if (ld_u(slot) < bits) atomicMax(slot, bits);
    }
  };
  for (int y0 = 0; y0 < n; y0 += kChunk) {
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      int y = y0 + j;
      CC l = y < n ? cc[base + y * stride] : CC(0);
      sL[j * kFhBlock + tid] = l;
      sD[j * kFhBlock + tid] = (y < n && l != CC(0)) ? D[base + y * stride] : kInf;
    }
    const int m = min(kChunk, n - y0);
    for (int j = 0; j < m; ++j) {
      const int y = y0 + j;
      const CC l = sL[j * kFhBlock + tid];
      if (active && l != cur) { close_run(y - 1, true); active = false; }
      if (l == CC(0)) {
        if (finish) D[base + y * stride] = 0.0f;
        continue;
      }
      if (!active) {
        active = true;
        cur = l;
        c = y;
        k = -1;
        if (y > 0 || bb) { k = 0; set(0, y - 1, 0.0f); }
      }
      const float hf = sD[j * kFhBlock + tid];
      if (hf < kInf) place(y, (double)hf);
    }
  }
  if (active) close_run(n - 1, bb != 0);
}

// finish for a single pass (x only): background 0, else sqrt.
template <class CC>
__global__ void __launch_bounds__(256) edt_finish(const CC *cc, long long n, float *D, unsigned *label_max) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  CC l = cc[i];
  if (l == CC(0)) { D[i] = 0.0f; return; }
  float v = sqrtf(D[i]);
  D[i] = v;
  if (label_max && v < kInf) atomicMax(label_max + (unsigned long long)l, __float_as_uint(v));
}

template <class CC>
void run_edt(Context &ctx, const CC *cc, Dims dm, const float aniso[3], bool bb, int dims_used, float *out,
             unsigned *label_max) {
  auto s = ctx.stream();
  const int rows = dm.sy * dm.sz;
  std::vector<float> table(dm.sx + 2, 0.0f);
  for (size_t k = 1; k < table.size(); ++k) table[k] = table[k - 1] + aniso[0];
  Buffer<float> acc(ctx, table.size());
  acc.upload(ctx, table.data(), table.size());
  {
    KernelTimer x_timer(ctx, "edt_x");
    if (dm.sx <= 128)
      edt_x<CC, 2><<<(rows + 3) / 4, 256, 0, s>>>(cc, dm.sx, rows, acc.data(), bb, out);
    else if (dm.sx <= 512)
      edt_x<CC, 8><<<(rows + 3) / 4, 256, 0, s>>>(cc, dm.sx, rows, acc.data(), bb, out);
    else if (dm.sx <= 2048)
      edt_x<CC, 32><<<(rows + 3) / 4, 256, 0, s>>>(cc, dm.sx, rows, acc.data(), bb, out);
    else
      edt_x_serial<CC><<<(rows + 255) / 256, 256, 0, s>>>(cc, dm.sx, rows, acc.data(), bb, out);
    BROOK_LAUNCHED();
  }
  const int last = dims_used - 1;
  if (last == 0) {
    long long n = (long long)dm.sx * dm.sy * dm.sz;
    edt_finish<CC><<<unsigned((n + 255) / 256), 256, 0, s>>>(cc, n, out, label_max);
    BROOK_LAUNCHED();
    return;
  }
  // envelope stacks deeper than the LDS part spill to one global column per line
  for (int axis = 1; axis <= last; ++axis) {
    int n = axis == 1 ? dm.sy : dm.sz;
    int nlines = axis == 1 ? dm.sx * dm.sz : dm.sx * dm.sy;
    size_t spill_rows = size_t(std::max(1, n + 3 - kStack));
    Buffer<int> V(ctx, spill_rows * nlines);
    Buffer<float> H(ctx, spill_rows * nlines);
    KernelTimer fh_timer(ctx, axis == 1 ? "edt_y" : "edt_z");
    edt_fh<CC><<<(nlines + kFhBlock - 1) / kFhBlock, kFhBlock, 0, s>>>(cc, dm, axis, aniso[axis], bb, nlines, nlines,
                                                                      V.data(), H.data(), out, axis == last, label_max);
    BROOK_LAUNCHED();
  }
}

}  // namespace

void edt(Context &ctx, const void *cc, int cc_bytes, Dims dims, const float aniso[3], bool black_border, int dims_used,
         float *out, unsigned *label_max) {
  for (int a = 0; a < 3; ++a)
    if (!(aniso[a] > 0) || !std::isfinite(aniso[a])) throw std::invalid_argument("anisotropy must be finite and positive");
  if (int64_t(dims.sx) * dims.sy * dims.sz == 0) return;
  if (cc_bytes == 1) run_edt<uint8_t>(ctx, static_cast<const uint8_t *>(cc), dims, aniso, black_border, dims_used, out, label_max);
  else if (cc_bytes == 2) run_edt<uint16_t>(ctx, static_cast<const uint16_t *>(cc), dims, aniso, black_border, dims_used, out, label_max);
  else run_edt<uint32_t>(ctx, static_cast<const uint32_t *>(cc), dims, aniso, black_border, dims_used, out, label_max);
}

}  // namespace brook
