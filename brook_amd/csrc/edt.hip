// Multi-label anisotropic Euclidean distance transform on gfx950, with the reference's rounding
// rule (edt_kernels.cuh:15-144, edt.cu:12-23): squared x-distances from a float32 accumulation
// table, y/z lower envelopes with float64 intersections, sqrtf at the end.
//
// Layout for HBM: the x pass gives each row to one wave (64 consecutive voxels per load, run
// boundaries from two ballots, no atomics); the y and z passes first list the label runs of every
// line (one thread per line, coalesced across neighbouring x) and then give each run its own
// thread. Background runs are never evaluated (their values are discarded by the finish rule) and
// the last pass fuses sqrt, the background zero and the per-label maximum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "stages.hpp"

namespace brook {
namespace {

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
      unsigned long long below = S & lanes_upto(lane);
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

// Lower envelopes along y (axis 1) or z (axis 2), run-parallel. A run is a maximal stretch of one
// nonzero id along a line; runs are independent (the walls at their ends stand for the
// neighbouring id or the border), so each run gets its own thread and no wave waits on another
// line's runs.
//   edt_runs:   one thread per line: lists the line's runs (count-only when runs == nullptr) and
//               writes the background zeros of the last pass;
//   edt_run_fh: one thread per run: the envelope of the run's parabolas (heights from the previous
//               pass) with float64 intersections, then the run's values.
// The envelope stack keeps its first kRunStack entries in LDS; deeper stacks spill to a per-run
// global column. The intersection of entries j-1 and j is recomputed from them when needed -- the
// same expression on the same operands as when entry j was placed, so the same bits.
constexpr int kRunStack = 24, kRunBlock = 256, kLineChunk = 8;

struct LineGeom {
  long long base, stride;
  int n;
};
__device__ __forceinline__ LineGeom line_geom(Dims dm, int axis, int t) {
  LineGeom g;
  if (axis == 1) {
    const int x = t % dm.sx, z = t / dm.sx;
    g.base = x + (long long)dm.sx * dm.sy * z;
    g.n = dm.sy;
    g.stride = dm.sx;
  } else {
    g.base = t;
    g.n = dm.sz;
    g.stride = (long long)dm.sx * dm.sy;
  }
  return g;
}

// Run record: x = line, y = first position, z = last position | left wall << 29 | right wall << 30,
// w = first entry of the run's spill column. The count pass stores each line's number of runs and
// of spill entries; the list pass writes the line's runs from the scanned offsets (no shared
// counters).
template <class CC>
__global__ void __launch_bounds__(256) edt_runs(const CC *cc, Dims dm, int axis, int bb, int nlines, int finish, float *D,
                                                int *line_runs, long long *line_spill, const int *run_at,
                                                const long long *spill_at, int4 *runs) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nlines) return;
  const LineGeom g = line_geom(dm, axis, t);
  const bool list = runs != nullptr;
  int nr = 0;
  long long ns = 0, at = list ? spill_at[t] : 0;
  const int first = list ? run_at[t] : 0;
  auto emit = [&](int c, int d, bool right) {
    const int spill = max(0, d - c + 3 - kRunStack);
    if (list) runs[first + nr] = make_int4(t, c, d | ((c > 0 || bb) ? 1 << 29 : 0) | (right ? 1 << 30 : 0), int(at + ns));
    ++nr;
    ns += spill;
  };
  CC cur = CC(0);
  int c = 0;
  for (int y0 = 0; y0 < g.n; y0 += kLineChunk) {
    CC lab[kLineChunk];
#pragma unroll
    for (int j = 0; j < kLineChunk; ++j) lab[j] = y0 + j < g.n ? cc[g.base + (y0 + j) * g.stride] : CC(0);
    const int m = min(kLineChunk, g.n - y0);
    for (int j = 0; j < m; ++j) {
      const int y = y0 + j;
      const CC l = lab[j];
      if (l != cur) {
        if (cur != CC(0)) emit(c, y - 1, true);
        c = y;
        cur = l;
      }
      if (l == CC(0) && finish && list) D[g.base + y * g.stride] = 0.0f;
    }
  }
  if (cur != CC(0)) emit(c, g.n - 1, bb != 0);
  if (!list) {
    line_runs[t] = nr;
    line_spill[t] = ns;
  }
}

template <class CC>
__global__ void __launch_bounds__(kRunBlock) edt_run_fh(const CC *cc, Dims dm, int axis, float w, int nruns,
                                                        const int4 *runs, int *spill_v, float *spill_h, float *D,
                                                        int finish, unsigned *label_max) {
  __shared__ int sV[kRunStack * kRunBlock];
  __shared__ float sH[kRunStack * kRunBlock];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * kRunBlock + tid;
  if (i >= nruns) return;
  const int4 r = runs[i];
  const LineGeom g = line_geom(dm, axis, r.x);
  const int c = r.y, d = r.z & ((1 << 29) - 1);
  const bool left = (r.z >> 29) & 1, right = (r.z >> 30) & 1;
  int *sv = spill_v + r.w;
  float *sh = spill_h + r.w;
  const double w2 = (double)(w * w);
  auto getV = [&](int k) -> int { return k < kRunStack ? sV[k * kRunBlock + tid] : sv[k - kRunStack]; };
  auto getH = [&](int k) -> float { return k < kRunStack ? sH[k * kRunBlock + tid] : sh[k - kRunStack]; };
  auto set = [&](int k, int v, float h) {
    if (k < kRunStack) { sV[k * kRunBlock + tid] = v; sH[k * kRunBlock + tid] = h; }
    else { sv[k - kRunStack] = v; sh[k - kRunStack] = h; }
  };
  // intersection of entries j-1 and j (j >= 1), as computed when entry j was placed
  auto inter = [&](int j) -> double {
    double v1 = (double)getV(j - 1), h1 = (double)getH(j - 1), v2 = (double)getV(j), h2 = (double)getH(j);
    return ((h2 + w2 * v2 * v2) - (h1 + w2 * v1 * v1)) / (2.0 * w2 * (v2 - v1));
  };
  int k = -1;
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
  if (left) { k = 0; set(0, c - 1, 0.0f); }
  for (int q0 = c; q0 <= d; q0 += kLineChunk) {  // heights of the previous pass, a chunk in flight
    float hv[kLineChunk];
#pragma unroll
    for (int j = 0; j < kLineChunk; ++j) hv[j] = q0 + j <= d ? D[g.base + (q0 + j) * g.stride] : kInf;
    const int m = min(kLineChunk, d - q0 + 1);
    for (int j = 0; j < m; ++j)
      if (hv[j] < kInf) place(q0 + j, (double)hv[j]);
  }
  if (right) place(d + 1, 0.0);
  float run_max = 0.0f;
  if (k < 0) {
    for (int q = c; q <= d; ++q) D[g.base + q * g.stride] = kInf;
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
      D[g.base + q * g.stride] = v;
    }
  }
  if (finish && label_max && run_max < kInf) {
    unsigned bits = __float_as_uint(run_max);
    unsigned *slot = label_max + (unsigned long long)cc[g.base + c * g.stride];
    if (ld_u(slot) < bits) atomicMax(slot, bits);
  }
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
  for (int axis = 1; axis <= last; ++axis) {
    const int nlines = axis == 1 ? dm.sx * dm.sz : dm.sx * dm.sy;
    const unsigned lgrid = unsigned((nlines + 255) / 256);
    const char *timer = axis == 1 ? "edt_y" : "edt_z";  // two scopes: the host read between them is not timed
    Buffer<int> line_runs(ctx, size_t(nlines)), run_at(ctx, size_t(nlines));
    Buffer<long long> line_spill(ctx, size_t(nlines)), spill_at(ctx, size_t(nlines));
    {
      KernelTimer count_timer(ctx, timer);
      edt_runs<CC><<<lgrid, 256, 0, s>>>(cc, dm, axis, bb, nlines, 0, out, line_runs.data(), line_spill.data(), nullptr,
                                         nullptr, nullptr);
      BROOK_LAUNCHED();
      exclusive_sum(ctx, line_runs.data(), run_at.data(), nlines);
      exclusive_sum(ctx, line_spill.data(), spill_at.data(), nlines);
    }
    int last_runs = 0, last_at = 0;
    long long last_spill = 0, last_spill_at = 0;
    BROOK_HIP(hipMemcpyAsync(&last_runs, line_runs.data() + nlines - 1, 4, hipMemcpyDeviceToHost, s));
    BROOK_HIP(hipMemcpyAsync(&last_at, run_at.data() + nlines - 1, 4, hipMemcpyDeviceToHost, s));
    BROOK_HIP(hipMemcpyAsync(&last_spill, line_spill.data() + nlines - 1, 8, hipMemcpyDeviceToHost, s));
    BROOK_HIP(hipMemcpyAsync(&last_spill_at, spill_at.data() + nlines - 1, 8, hipMemcpyDeviceToHost, s));
    ctx.synchronize();
    const long long total_runs = (long long)last_at + last_runs, total_spill = last_spill_at + last_spill;
    if (total_runs > INT32_MAX || total_spill > INT32_MAX)
      throw Unsupported("EDT envelope spill exceeds 2^31 entries (runs too long for this build)");
    const int nruns = int(total_runs);
    Buffer<int4> runs(ctx, size_t(std::max(nruns, 1)));
    Buffer<int> V(ctx, size_t(std::max<long long>(total_spill, 1)));
    Buffer<float> H(ctx, size_t(std::max<long long>(total_spill, 1)));
    KernelTimer fh_timer(ctx, timer);
    edt_runs<CC><<<lgrid, 256, 0, s>>>(cc, dm, axis, bb, nlines, axis == last, out, nullptr, nullptr, run_at.data(),
                                       spill_at.data(), runs.data());
    BROOK_LAUNCHED();
    if (nruns) {
      edt_run_fh<CC><<<unsigned((nruns + kRunBlock - 1) / kRunBlock), kRunBlock, 0, s>>>(
          cc, dm, axis, aniso[axis], nruns, runs.data(), V.data(), H.data(), out, axis == last, label_max);
      BROOK_LAUNCHED();
    }
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
