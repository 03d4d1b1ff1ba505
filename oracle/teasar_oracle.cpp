// TEST INFRASTRUCTURE ONLY -- the CPU oracle of the TEASAR hot path.
//
// A plain C++17 restatement of giorgioangel/brook's skeletonize path (the lockstep tracer's
// semantics), used as the parity checker for the HIP engine and as the timed CPU baseline.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library;
// the product (brook_amd/) never links, imports or calls it.
//
// Every function cites the reference file:line it restates. Arithmetic follows the reference's
// floating-point regime (float32 with no contraction, IEEE sqrt/div; the EDT envelope in float64).
//
// Pinning (what checks each function independently of this file's own formulas):
//   ccl            brute-force BFS labelling on random volumes; c_smoke.c:27-34 known answer
//   analyze        numpy bincount / bbox on random volumes
//   edt            float64 exact EDT (edt_exact_f64, itself pinned by brute force) to 1e-5 rel;
//                  known answer c_smoke.c:53-56 ({2,1,0,...,1,2}); golden skeleton digests
//   edt_exact_f64  brute-force nearest-different-label search on small volumes
//   PointSet       CPython's own set iteration order (python3 in the test process)
//   border targets golden skeleton digests (fix_borders=True) from the reference's test suite
//   geometric flood  brute-force Bellman-Ford to fixpoint on small volumes
//   target order   numpy lexsort of the reference rule (component_order_smoke.cpp:9-21) on its grid
//                  of lengths, densities and tie rates, with +inf and -0 keys
//   railroad/join  test_package.py:751-811 known answers (26 singles, 325 pairs, all 26, 9^3 road)
//   invalidation   one source: the claimed set is the ball's component and every key the exact
//                  minimax bottleneck (a Dijkstra over max); many sources: every claim valid and the
//                  claimed set closed under the final keys; golden skeleton digests
//   whole path     golden digests test_package.py:2079-2106 (Kimimaro 5.8.1 results)
//   vertex-weighted flood and parents (fix_branching=False)  brute-force float32 Bellman-Ford of
//                  the vw_step fixpoint; every parent chain satisfies d[v] = vw_step(d[parent], pdrf[v])
//                  and reaches the root; the STOP rule's union of vertices and edges equals the
//                  union of full walks to the root (Kimimaro's paths) on random volumes
#include "avocado_oracle.hpp"
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <queue>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace oracle {
using i64 = int64_t;
using u64 = uint64_t;

// Neighbour table, common.cuh:33-47 (faces, xy/yz/xz edges, corners).
static const int NX[26] = {-1, 1, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, -1, -1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1};
static const int NY[26] = {0, 0, -1, 1, 0, 0, -1, 1, -1, 1, -1, -1, 1, 1, 0, 0, 0, 0, -1, -1, 1, 1, -1, -1, 1, 1};
static const int NZ[26] = {0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1};
// dijkstra3d order ranks, common.cuh:74-81; opposite offsets, common.cuh:62-66.
static const int D3RANK[26] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 21, 20, 24, 19, 23, 22, 25};
static const int D3INV[26] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 22, 20, 19, 24, 23, 21, 25};
static const int OPP[26] = {1, 0, 3, 2, 5, 4, 9, 8, 7, 6, 13, 12, 11, 10, 17, 16, 15, 14, 25, 24, 23, 22, 21, 20, 19, 18};

// cc3d's voxel-graph bit of neighbour k (common.cuh:49-56): the step v -> v + N[k] is permitted iff
// bit BIT[k] of graph[v] is set.
static const int BIT[26] = {1, 0, 3, 2, 5, 4, 9, 7, 8, 6, 17, 13, 16, 12, 15, 11, 14, 10, 25, 21, 23, 19, 24, 20, 22, 18};

// The gate a voxel graph puts on each step (common.cuh:84-100): a push from u to u + N[k] needs
// bit BIT[k] of graph[u]; a pull by c from its neighbour u = c + N[k] needs the step u -> c, bit
// BIT[OPP[k]] of graph[u]. Without a graph every step is open.
struct Gate {
  const uint32_t *vg = nullptr;
  bool push(int64_t u, int k) const { return !vg || ((vg[u] >> BIT[k]) & 1u); }
  bool pull(int64_t u, int k) const { return !vg || ((vg[u] >> BIT[OPP[k]]) & 1u); }
};

static const float FINF = std::numeric_limits<float>::infinity();
static const u64 MAXKEY = ~u64(0);

struct Shape {
  i64 sx = 0, sy = 0, sz = 0;
  i64 n() const { return sx * sy * sz; }
  i64 at(i64 x, i64 y, i64 z) const { return x + sx * (y + sy * z); }
  void unravel(i64 v, i64 &x, i64 &y, i64 &z) const {
    z = v / (sx * sy);
    i64 r = v - z * sx * sy;
    y = r / sx;
    x = r - y * sx;
  }
  bool inside(i64 x, i64 y, i64 z) const { return x >= 0 && y >= 0 && z >= 0 && x < sx && y < sy && z < sz; }
};

static inline uint32_t fbits(float f) { uint32_t b; std::memcpy(&b, &f, 4); return b; }
static inline float bitsf(uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; }

// brook_vw_step, common.cuh:123-127.
static inline float vw_step(float d, float f) {
  float c = d + f;
  if (c == d && f > 0.0f && d < FINF) c = bitsf(fbits(d) + 1);
  return c;
}

// ---------------------------------------------------------------------------------------------
// Connected components, ccl.cu:24-82 / ccl_kernels.cuh:15-133: 26-connectivity between equal
// nonzero values; ids 1..N in ascending order of each component's minimum flat index.
struct Components {
  std::vector<uint32_t> cc;
  std::vector<i64> mapping;  // [0] = 0, [id] = input value
  std::vector<i64> roots;    // [id-1] = minimum flat index
};

static i64 uf_find(std::vector<i64> &p, i64 v) {
  while (p[v] != v) { p[v] = p[p[v]]; v = p[v]; }
  return v;
}

Components ccl(const i64 *labels, const Shape &s) {
  const i64 n = s.n();
  std::vector<i64> parent(n);
  for (i64 i = 0; i < n; ++i) parent[i] = i;
  for (i64 v = 0; v < n; ++v) {
    i64 lv = labels[v];
    if (!lv) continue;
    i64 x, y, z;
    s.unravel(v, x, y, z);
    for (int k = 0; k < 26; ++k) {
      i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
      if (!s.inside(nx, ny, nz)) continue;
      i64 u = s.at(nx, ny, nz);
      if (u >= v || labels[u] != lv) continue;
      i64 a = uf_find(parent, u), b = uf_find(parent, v);
      if (a == b) continue;
      if (a < b) parent[b] = a; else parent[a] = b;  // the smaller index stays the root
    }
  }
  Components out;
  out.cc.assign(n, 0);
  out.mapping.push_back(0);
  for (i64 v = 0; v < n; ++v) {
    if (!labels[v]) continue;
    i64 r = uf_find(parent, v);
    if (r == v) {
      out.roots.push_back(v);
      out.mapping.push_back(labels[v]);
      out.cc[v] = uint32_t(out.roots.size());
    } else {
      out.cc[v] = out.cc[r];  // r < v is numbered already
    }
  }
  return out;
}

// Components of a voxel graph (graph_components.cu:18-166, cc3d.color_connectivity_graph): the
// graph's permitted steps alone join voxels, background included, whatever their labels. Ids go to
// the components that hold a foreground voxel, 1..N in ascending order of each component's lowest
// voxel (background included); roots[id - 1] is the lowest foreground voxel; mapping[id] the label
// where the id's last run of voxels in flat order starts; keys[id] the component's rank among all
// of them (background-only ones included), 1-based. A uint8 graph (one XY plane) holds the bits
// of the steps within the plane, the corner bits (6..9) moved down by 2.
struct GraphComponents : Components {
  std::vector<i64> keys;  // [0] = 0
  bool input_uniform = true;  // every voxel has the label of the first
};

GraphComponents graph_ccl(const i64 *labels, const uint32_t *vg, bool compact_2d, const Shape &s) {
  const i64 n = s.n();
  std::vector<i64> parent(n);
  for (i64 i = 0; i < n; ++i) parent[i] = i;
  for (i64 v = 0; v < n; ++v) {
    i64 x, y, z;
    s.unravel(v, x, y, z);
    for (int k = 0; k < 26; ++k) {
      int bit = BIT[k];
      if (compact_2d && bit >= 6) bit -= 2;
      if (!((vg[v] >> bit) & 1u)) continue;
      i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
      if (!s.inside(nx, ny, nz)) continue;
      i64 u = s.at(nx, ny, nz);
      if (u >= v) continue;
      i64 a = uf_find(parent, u), b = uf_find(parent, v);
      if (a == b) continue;
      if (a < b) parent[b] = a; else parent[a] = b;
    }
  }
  std::vector<i64> rank(n, 0), compact(n, 0);  // by root: 1-based rank among all roots; compact id
  i64 roots_seen = 0;
  for (i64 v = 0; v < n; ++v)
    if (uf_find(parent, v) == v) rank[v] = ++roots_seen;
  std::vector<uint8_t> has_fg(n, 0);
  for (i64 v = 0; v < n; ++v)
    if (labels[v] > 0) has_fg[uf_find(parent, v)] = 1;
  GraphComponents out;
  out.keys.push_back(0);
  for (i64 r = 0; r < n; ++r)
    if (has_fg[r]) { compact[r] = i64(out.keys.size()); out.keys.push_back(rank[r]); }
  const size_t count = out.keys.size() - 1;
  out.cc.assign(n, 0);
  out.roots.assign(count, -1);
  out.mapping.assign(count + 1, 0);
  std::vector<i64> last_run(count + 1, -1);
  for (i64 v = 0; v < n; ++v) {
    const uint32_t id = labels[v] > 0 ? uint32_t(compact[uf_find(parent, v)]) : 0;
    out.cc[v] = id;
    if (labels[v] != labels[0]) out.input_uniform = false;
    if (!id) continue;
    if (out.roots[id - 1] < 0) out.roots[id - 1] = v;
    if (v == 0 || out.cc[v - 1] != id) last_run[id] = v;
  }
  for (size_t id = 1; id <= count; ++id) out.mapping[id] = labels[last_run[id]];
  return out;
}

// The graph-aware EDT (graph_edt.cu:14-132, after the edt package): a binary volume on the doubled
// grid, the voxel centres at even positions; a voxel's +x, +y, +z half-steps (odd in one axis only)
// are foreground when the graph permits that step, every other position of a foreground voxel is
// foreground, and with a black border the last doubled position on each axis is background. Its
// EDT with half the anisotropy (the reference's rounding rule, edt_rule), sampled at the voxel
// centres; 0 on background. `vg` is the tracing graph (a uint8 plane graph's value widened as is,
// graph_components.cu:167-176).
template <class L>
void edt_rule(const L *lab, const Shape &s, const float aniso[3], bool black_border, int dims, float *D);

template <class L>
void graph_edt(const L *lab, const uint32_t *vg, const Shape &s, const float aniso[3], bool black_border, float *D) {
  const Shape d2{2 * s.sx, 2 * s.sy, 2 * s.sz};
  std::vector<uint8_t> mask(size_t(d2.n()), 0);
  for (i64 z = 0; z < d2.sz; ++z)
    for (i64 y = 0; y < d2.sy; ++y)
      for (i64 x = 0; x < d2.sx; ++x) {
        if (black_border && (x == d2.sx - 1 || y == d2.sy - 1 || z == d2.sz - 1)) continue;
        const i64 src = s.at(x >> 1, y >> 1, z >> 1);
        if (!(lab[src] > 0)) continue;
        const int sub = int((x & 1) | ((y & 1) << 1) | ((z & 1) << 2));
        uint8_t m = 1;
        if (sub == 1) m = (vg[src] & 1u) != 0;
        else if (sub == 2) m = (vg[src] & 4u) != 0;
        else if (sub == 4) m = (vg[src] & 16u) != 0;
        mask[size_t(d2.at(x, y, z))] = m;
      }
  const float half[3] = {aniso[0] * 0.5f, aniso[1] * 0.5f, aniso[2] * 0.5f};
  std::vector<float> D2(size_t(d2.n()));
  edt_rule<uint8_t>(mask.data(), d2, half, black_border, 3, D2.data());
  for (i64 z = 0; z < s.sz; ++z)
    for (i64 y = 0; y < s.sy; ++y)
      for (i64 x = 0; x < s.sx; ++x) {
        const i64 v = s.at(x, y, z);
        D[v] = lab[v] > 0 ? D2[size_t(d2.at(2 * x, 2 * y, 2 * z))] : 0.0f;
      }
}

// ---------------------------------------------------------------------------------------------
// analyze, preamble.cu:10-29 / preamble_kernels.cuh:31-56: per-id bbox and voxel count.
struct Box {
  i64 lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {-1, -1, -1};
  i64 count = 0;
};

std::vector<Box> analyze(const std::vector<uint32_t> &cc, const Shape &s, size_t count) {
  std::vector<Box> boxes(count + 1);
  for (i64 z = 0, v = 0; z < s.sz; ++z)
    for (i64 y = 0; y < s.sy; ++y)
      for (i64 x = 0; x < s.sx; ++x, ++v) {
        uint32_t id = cc[v];
        if (!id) continue;
        Box &b = boxes[id];
        i64 p[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) { b.lo[a] = std::min(b.lo[a], p[a]); b.hi[a] = std::max(b.hi[a], p[a]); }
        b.count++;
      }
  i64 fg = 0;
  for (size_t i = 1; i <= count; ++i) fg += boxes[i].count;
  boxes[0].count = s.n() - fg;
  return boxes;
}

// ---------------------------------------------------------------------------------------------
// Multi-label anisotropic squared EDT with the reference's rounding rule:
// x pass edt_kernels.cuh:15-44 with the float accumulation table of edt.cu:21-22,
// y/z passes edt_kernels.cuh:48-144 (FH lower envelope, float64 intersections),
// finish edt.cu:12-15 (background 0, else sqrtf). `dims` 2 skips the z pass (a plane).
template <class L>
void edt_rule(const L *lab, const Shape &s, const float aniso[3], bool black_border, int dims, float *D) {
  std::vector<float> acc(s.sx + 2, 0.0f);
  for (size_t k = 1; k < acc.size(); ++k) acc[k] = acc[k - 1] + aniso[0];
  // pass x
  for (i64 line = 0; line < s.sy * s.sz; ++line) {
    i64 base = s.sx * line;
    i64 i = 0;
    while (i < s.sx) {
      L lv = lab[base + i];
      i64 c = i, d = i;
      while (d + 1 < s.sx && lab[base + d + 1] == lv) d++;
      bool lw = c > 0 || black_border, rw = d < s.sx - 1 || black_border;
      for (i64 p = c; p <= d; ++p) {
        float best = FINF;
        if (lw) { float dd = acc[p - (c - 1)]; float v = dd * dd; if (v < best) best = v; }
        if (rw) { float dd = acc[(d + 1) - p]; float v = dd * dd; if (v < best) best = v; }
        D[base + p] = best;
      }
      i = d + 1;
    }
  }
  // passes y, z
  std::vector<int> V;
  std::vector<float> H;
  std::vector<double> Z;
  for (int axis = 1; axis < dims; ++axis) {
    i64 n = axis == 1 ? s.sy : s.sz, stride = axis == 1 ? s.sx : s.sx * s.sy;
    i64 nlines = axis == 1 ? s.sx * s.sz : s.sx * s.sy;
    V.assign(n + 3, 0); H.assign(n + 3, 0.0f); Z.assign(n + 3, 0.0);
    float w = aniso[axis];
    double w2 = (double)(w * w);
    for (i64 line = 0; line < nlines; ++line) {
      i64 base;
      if (axis == 1) { i64 x = line % s.sx, z = line / s.sx; base = x + s.sx * s.sy * z; }
      else base = line;
      i64 i = 0;
      while (i < n) {
        L lv = lab[base + i * stride];
        i64 c = i, d = i;
        while (d + 1 < n && lab[base + (d + 1) * stride] == lv) d++;
        bool lw = c > 0 || black_border, rw = d < n - 1 || black_border;
        int k = -1;
        if (lw) { k = 0; V[0] = (int)(c - 1); H[0] = 0.0f; Z[0] = -1e300; }
        auto place = [&](i64 pp, double ph) {
          bool placed = false;
          while (k >= 0 && !placed) {
            double s_ = ((ph + w2 * (double)pp * (double)pp) - ((double)H[k] + w2 * (double)V[k] * (double)V[k])) /
                        (2.0 * w2 * ((double)pp - (double)V[k]));
            if (s_ <= Z[k]) k--;
            else { k++; V[k] = (int)pp; H[k] = (float)ph; Z[k] = s_; placed = true; }
          }
          if (!placed) { k = 0; V[0] = (int)pp; H[0] = (float)ph; Z[0] = -1e300; }
        };
        for (i64 p = c; p <= d; ++p) {
          float hf = D[base + p * stride];
          if (!(hf < FINF)) continue;
          place(p, (double)hf);
        }
        if (rw) place(d + 1, 0.0);
        if (k < 0) {
          for (i64 q = c; q <= d; ++q) D[base + q * stride] = FINF;
        } else {
          int kk = 0;
          for (i64 q = c; q <= d; ++q) {
            while (kk < k && Z[kk + 1] < (double)q) kk++;
            double dq = (double)q - (double)V[kk];
            D[base + q * stride] = (float)(w2 * dq * dq + (double)H[kk]);
          }
        }
        i = d + 1;
      }
    }
  }
  for (i64 v = 0; v < s.n(); ++v) D[v] = lab[v] == 0 ? 0.0f : std::sqrt(D[v]);
}

// Exact float64 multi-label EDT (the FP64 reference for the 1e-5 tolerance): distance to the
// nearest voxel of a different label (or the outside, with a black border), computed by the
// textbook separable lower envelope in double with exact integer spacings. Not a restatement of
// the reference; its own pin is a brute-force search on small volumes.
template <class L>
void edt_exact_f64(const L *lab, const Shape &s, const double aniso[3], bool black_border, int dims, double *D) {
  const double BIG = std::numeric_limits<double>::infinity();
  i64 dimsz[3] = {s.sx, s.sy, s.sz};
  i64 strides[3] = {1, s.sx, s.sx * s.sy};
  for (i64 v = 0; v < s.n(); ++v) D[v] = BIG;
  std::vector<double> f, g, zz;
  std::vector<i64> vv;
  for (int axis = 0; axis < dims; ++axis) {
    i64 n = dimsz[axis], stride = strides[axis];
    double w = aniso[axis];
    f.resize(n + 2); g.resize(n + 2); vv.resize(n + 3); zz.resize(n + 4);
    for (i64 base = 0; base < s.n(); ++base) {
      i64 x, y, z; s.unravel(base, x, y, z);
      i64 p[3] = {x, y, z};
      if (p[axis] != 0) continue;
      i64 i = 0;
      while (i < n) {
        L lv = lab[base + i * stride];
        i64 c = i, d = i;
        while (d + 1 < n && lab[base + (d + 1) * stride] == lv) d++;
        bool lw = c > 0 || black_border, rw = d < n - 1 || black_border;
        // sites: walls (height 0) and voxels of the run with their previous-pass value
        std::vector<std::pair<double, double>> sites;  // (position, height)
        if (lw) sites.push_back({double(c - 1), 0.0});
        for (i64 q = c; q <= d; ++q) {
          double h = axis == 0 ? BIG : D[base + q * stride];
          if (h < BIG) sites.push_back({double(q), h});
        }
        if (rw) sites.push_back({double(d + 1), 0.0});
        for (i64 q = c; q <= d; ++q) {
          double best = BIG;
          for (auto &st : sites) {
            double dq = w * (double(q) - st.first);
            best = std::min(best, dq * dq + st.second);
          }
          g[q - c] = best;
        }
        for (i64 q = c; q <= d; ++q) D[base + q * stride] = g[q - c];
        i = d + 1;
      }
    }
  }
  for (i64 v = 0; v < s.n(); ++v) D[v] = lab[v] == 0 ? 0.0 : std::sqrt(D[v]);
}

// ---------------------------------------------------------------------------------------------
// CPython insertion-only set of 3-tuples of ints (point_set.hpp:11-55): tuple hash (xxHash-based,
// CPython >= 3.8), open addressing with 9 linear probes and perturbed jumps, resize at 3/5 full.
class PointSet {
  struct Slot { std::array<i64, 3> p{}; u64 h = 0; bool live = false; };
  std::vector<Slot> table_ = std::vector<Slot>(8);
  size_t used_ = 0;
  static u64 hash(const std::array<i64, 3> &pt) {
    const u64 P1 = 11400714785074694791ull, P2 = 14029467366897019727ull, P5 = 2870177450012600261ull;
    u64 acc = P5;
    for (i64 v : pt) {
      // hash(int) for |v| < 2^61 - 1 is v itself, with -1 mapped to -2.
      const u64 M = (u64(1) << 61) - 1;
      u64 mag = v < 0 ? u64(-(v + 1)) + 1 : u64(v);
      i64 h = i64(mag % M);
      if (v < 0) h = -h;
      if (h == -1) h = -2;
      acc += u64(h) * P2;
      acc = (acc << 31) | (acc >> 33);
      acc *= P1;
    }
    acc += u64(3) ^ (P5 ^ 3527539ull);
    return acc == ~u64(0) ? 1546275796ull : acc;
  }
  bool insert(const std::array<i64, 3> &pt, u64 h) {
    size_t mask = table_.size() - 1, i = h & mask;
    u64 perturb = h;
    for (;;) {
      size_t probes = i + 9 <= mask ? 9 : 0;
      for (size_t j = 0; j <= probes; ++j) {
        Slot &sl = table_[i + j];
        if (!sl.live) { sl.p = pt; sl.h = h; sl.live = true; return true; }
        if (sl.h == h && sl.p == pt) return false;
      }
      perturb >>= 5;
      i = (i * 5 + 1 + perturb) & mask;
    }
  }

 public:
  void add(const std::array<i64, 3> &pt) {
    if (!insert(pt, hash(pt))) return;
    ++used_;
    if (used_ * 5 < (table_.size() - 1) * 3) return;
    size_t target = used_ * (used_ > 50000 ? 2 : 4), size = 8;
    while (size <= target) size *= 2;
    std::vector<Slot> old = std::move(table_);
    table_ = std::vector<Slot>(size);
    for (auto &sl : old) if (sl.live) insert(sl.p, sl.h);
  }
  std::vector<std::array<i64, 3>> values() const {
    std::vector<std::array<i64, 3>> r;
    for (auto &sl : table_) if (sl.live) r.push_back(sl.p);
    return r;
  }
};

// ---------------------------------------------------------------------------------------------
// Border targets, borders.cu:53-116 (the serial plan; the batched plan equals it,
// test_package.py:1026-1041). Faces (a,b,fixed c) = (0,1,2),(0,2,1),(1,2,0), low then high.
static float sqdist(float x, float y, float a, float b, float wx, float wy) {
  float dx = wx * (x - a), dy = wy * (y - b);
  return dx * dx + dy * dy;
}
static std::array<float, 4> face_keys(i64 ix, i64 iy, float cx, float cy, i64 nx, i64 ny, float wx, float wy) {
  float x = float(ix), y = float(iy), sx = float(nx), sy = float(ny);
  float center_x = float(double(wx * sx) / 2.0), center_y = float(double(wy * sy) / 2.0);
  float far_x = float(double(sx) - 0.5), far_y = float(double(sy) - 0.5);
  float corner = std::min({sqdist(x, y, -0.5f, -0.5f, wx, wy), sqdist(x, y, far_x, -0.5f, wx, wy),
                           sqdist(x, y, far_x, far_y, wx, wy), sqdist(x, y, -0.5f, far_x, wx, wy)});
  float edge = float(std::min({double(wx) * (double(x) - 0.5), double(wx) * (double(sx) - 0.5 - double(x)),
                               double(wy) * (double(y) - 0.5), double(wy) * (double(sy) - 0.5 - double(y))}));
  return {sqdist(x, y, cx, cy, wx, wy), sqdist(x, y, center_x, center_y, wx, wy), corner, edge};
}

std::vector<std::vector<std::array<i64, 3>>> border_targets(const std::vector<uint32_t> &cc, const Shape &s, size_t count,
                                                            const float aniso[3]) {
  std::vector<PointSet> sets(count + 1);
  const int axes[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
  i64 shp[3] = {s.sx, s.sy, s.sz};
  for (int face = 0; face < 6; ++face) {
    int a = axes[face / 2][0], b = axes[face / 2][1], c = axes[face / 2][2];
    i64 fixed = face % 2 ? shp[c] - 1 : 0, fx = shp[a], fy = shp[b];
    // plane of shape (fy, fx, 1): flat i = y + fy * x with y along b, x along a
    std::vector<i64> plane(fx * fy);
    for (i64 x = 0; x < fx; ++x)
      for (i64 y = 0; y < fy; ++y) {
        i64 p[3];
        p[a] = x; p[b] = y; p[c] = fixed;
        plane[y + fy * x] = cc[s.at(p[0], p[1], p[2])];
      }
    Shape ps{fy, fx, 1};
    Components c2 = ccl(plane.data(), ps);
    std::vector<float> d(fx * fy);
    float pa[3] = {aniso[b], aniso[a], 1.0f};
    edt_rule<uint32_t>(c2.cc.data(), ps, pa, true, 2, d.data());
    size_t k = c2.mapping.size();
    std::vector<float> maxdist(k, 0), sumx(k, 0), sumy(k, 0);
    std::vector<uint32_t> sizes(k, 0);
    std::vector<i64> first(k, INT64_MAX), pick(k, -1);
    for (i64 x = 0; x < fx; ++x)
      for (i64 y = 0; y < fy; ++y) {
        size_t i = y + fy * x;
        uint32_t l = c2.cc[i];
        if (!l) continue;
        sumx[l] += float(x); sumy[l] += float(y); ++sizes[l];
        if (d[i] > 0) { maxdist[l] = std::max(maxdist[l], d[i]); first[l] = std::min(first[l], x + fx * y); }
      }
    float wx = aniso[a], wy = aniso[b];
    std::vector<float> cxs(k, 0), cys(k, 0);
    for (size_t l = 1; l < k; ++l)
      if (sizes[l]) {
        float px = (wx * sumx[l]) / float(sizes[l]), py = (wy * sumy[l]) / float(sizes[l]);
        if (px - (wx * float(fx)) / 2.0f < 0) px = px + wx;
        if (py - (wy * float(fy)) / 2.0f < 0) py = py + wy;
        cxs[l] = float(int(px / wx)); cys[l] = float(int(py / wy));
      }
    std::vector<std::array<float, 4>> best(k);
    for (i64 y = 0; y < fy; ++y)
      for (i64 x = 0; x < fx; ++x) {
        size_t i = y + fy * x;
        uint32_t l = c2.cc[i];
        if (!l || d[i] == 0 || d[i] != maxdist[l]) continue;
        auto key = face_keys(x, y, cxs[l], cys[l], fx, fy, wx, wy);
        if (pick[l] < 0 || key < best[l]) { best[l] = key; pick[l] = x + fx * y; }
      }
    std::vector<size_t> order;
    for (size_t l = 1; l < k; ++l) if (pick[l] >= 0) order.push_back(l);
    std::sort(order.begin(), order.end(), [&](size_t u, size_t v) { return first[u] < first[v]; });
    for (size_t l : order) {
      std::array<i64, 3> p;
      p[a] = pick[l] % fx; p[b] = pick[l] / fx; p[c] = fixed;
      sets.at(size_t(c2.mapping[l])).add(p);
    }
  }
  std::vector<std::vector<std::array<i64, 3>>> out;
  for (auto &st : sets) out.push_back(st.values());
  return out;
}

// ---------------------------------------------------------------------------------------------
// Label-gated multi-source geometric flood (batched_coop.inc:10-113, mode 0): the unique least
// fixpoint of d[v] = min_u fl(d[u] + W[k]) over same-label 26-neighbours. float addition is monotone
// and non-decreasing, so Dijkstra reproduces the fixpoint bit for bit.
struct HeapItem {
  float d;
  i64 v;
  bool operator>(const HeapItem &o) const { return d > o.d || (d == o.d && v > o.v); }
};
using MinHeap = std::priority_queue<HeapItem, std::vector<HeapItem>, std::greater<HeapItem>>;

void geometric_flood(const std::vector<uint32_t> &cc, const Shape &s, const float W[26], const std::vector<i64> &sources,
                     std::vector<float> &dist, const Gate &gate = {}) {
  MinHeap heap;
  for (i64 src : sources) { dist[src] = 0.0f; heap.push({0.0f, src}); }
  while (!heap.empty()) {
    HeapItem it = heap.top(); heap.pop();
    if (it.d > dist[it.v]) continue;
    i64 x, y, z; s.unravel(it.v, x, y, z);
    uint32_t l = cc[it.v];
    for (int k = 0; k < 26; ++k) {
      i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
      if (!s.inside(nx, ny, nz)) continue;
      i64 u = s.at(nx, ny, nz);
      if (cc[u] != l || !gate.push(it.v, k)) continue;
      float c = it.d + W[k];
      if (c < dist[u]) { dist[u] = c; heap.push({c, u}); }
    }
  }
}

// The same fixpoint from sources with initial values (BatchedFlooder::flood seeded,
// batched.cu:122-148): every finite seed starts the flood with its value.
void geometric_flood_seeded(const std::vector<uint32_t> &cc, const Shape &s, const float W[26],
                            const std::vector<std::pair<i64, float>> &seeds, std::vector<float> &dist,
                            const Gate &gate = {}) {
  MinHeap heap;
  for (auto &sd : seeds) { dist[sd.first] = sd.second; heap.push({sd.second, sd.first}); }
  while (!heap.empty()) {
    HeapItem it = heap.top(); heap.pop();
    if (it.d > dist[it.v]) continue;
    i64 x, y, z; s.unravel(it.v, x, y, z);
    uint32_t l = cc[it.v];
    for (int k = 0; k < 26; ++k) {
      i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
      if (!s.inside(nx, ny, nz)) continue;
      i64 u = s.at(nx, ny, nz);
      if (cc[u] != l || !gate.push(it.v, k)) continue;
      float c = it.d + W[k];
      if (c < dist[u]) { dist[u] = c; heap.push({c, u}); }
    }
  }
}

// Free-space seed of a soma's DAF (freespace_seed_lab, batched_regular.inc:51-73): a voxel of the
// label within `radius` (physical, float32) of the root starts at the closed-form length of the
// shortest 26-connected free-space path to the root, evaluated in float32 in this order.
static bool freespace_seed(i64 dx_i, i64 dy_i, i64 dz_i, float wx, float wy, float wz, float radius, float &val) {
  float rdx = wx * (float)dx_i, rdy = wy * (float)dy_i, rdz = wz * (float)dz_i;
  if (rdx * rdx + rdy * rdy + rdz * rdz > radius * radius) return false;
  float dx = std::fabs((float)dx_i), dy = std::fabs((float)dy_i), dz = std::fabs((float)dz_i);
  float dxyz = std::min(std::min(dx, dy), dz);
  float dxy = std::min(dx, dy), dyz = std::min(dy, dz), dxz = std::min(dx, dz);
  val = dxyz * std::sqrt(wx * wx + wy * wy + wz * wz)
        + wx * (dx - dxyz) + wy * (dy - dxyz) + wz * (dz - dxyz)
        + (dxy - dxyz) * (std::sqrt(wx * wx + wy * wy) - wx - wy)
        + (dxz - dxyz) * (std::sqrt(wx * wx + wz * wz) - wx - wz)
        + (dyz - dxyz) * (std::sqrt(wy * wy + wz * wz) - wy - wz);
  return true;
}

// fill_voids' count on a label's bounding box (fillvoids_kernels.cuh, fillvoids_flood.cuh): the
// voxels of the box that are not the label and are not 6-connected through such voxels to a face
// of the box.
i64 void_count(const std::vector<uint32_t> &cc, const Shape &s, uint32_t id, const Box &b) {
  Shape c{b.hi[0] - b.lo[0] + 1, b.hi[1] - b.lo[1] + 1, b.hi[2] - b.lo[2] + 1};
  auto fg = [&](i64 x, i64 y, i64 z) { return cc[s.at(x + b.lo[0], y + b.lo[1], z + b.lo[2])] == id; };
  std::vector<uint8_t> seen(size_t(c.n()), 0);
  std::vector<i64> stack;
  for (i64 z = 0; z < c.sz; ++z)
    for (i64 y = 0; y < c.sy; ++y)
      for (i64 x = 0; x < c.sx; ++x)
        if ((x == 0 || y == 0 || z == 0 || x == c.sx - 1 || y == c.sy - 1 || z == c.sz - 1) && !fg(x, y, z)) {
          seen[c.at(x, y, z)] = 1;
          stack.push_back(c.at(x, y, z));
        }
  while (!stack.empty()) {
    i64 v = stack.back(); stack.pop_back();
    i64 x, y, z; c.unravel(v, x, y, z);
    for (int k = 0; k < 6; ++k) {
      i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
      if (!c.inside(nx, ny, nz) || fg(nx, ny, nz)) continue;
      i64 u = c.at(nx, ny, nz);
      if (!seen[u]) { seen[u] = 1; stack.push_back(u); }
    }
  }
  i64 voids = 0;
  for (i64 z = 0; z < c.sz; ++z)
    for (i64 y = 0; y < c.sy; ++y)
      for (i64 x = 0; x < c.sx; ++x)
        if (!fg(x, y, z) && !seen[c.at(x, y, z)]) ++voids;
  return voids;
}

// The voids themselves (flat indices, ascending).
std::vector<i64> void_voxels(const std::vector<uint32_t> &cc, const Shape &s, uint32_t id, const Box &b) {
  std::vector<i64> out;
  if (!void_count(cc, s, id, b)) return out;
  Shape c{b.hi[0] - b.lo[0] + 1, b.hi[1] - b.lo[1] + 1, b.hi[2] - b.lo[2] + 1};
  auto fg = [&](i64 x, i64 y, i64 z) { return cc[s.at(x + b.lo[0], y + b.lo[1], z + b.lo[2])] == id; };
  std::vector<uint8_t> seen(size_t(c.n()), 0);
  std::vector<i64> stack;
  for (i64 z = 0; z < c.sz; ++z)
    for (i64 y = 0; y < c.sy; ++y)
      for (i64 x = 0; x < c.sx; ++x)
        if ((x == 0 || y == 0 || z == 0 || x == c.sx - 1 || y == c.sy - 1 || z == c.sz - 1) && !fg(x, y, z)) {
          seen[c.at(x, y, z)] = 1;
          stack.push_back(c.at(x, y, z));
        }
  while (!stack.empty()) {
    i64 v = stack.back(); stack.pop_back();
    i64 x, y, z; c.unravel(v, x, y, z);
    for (int k = 0; k < 6; ++k) {
      i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
      if (!c.inside(nx, ny, nz) || fg(nx, ny, nz)) continue;
      i64 u = c.at(nx, ny, nz);
      if (!seen[u]) { seen[u] = 1; stack.push_back(u); }
    }
  }
  for (i64 z = 0; z < c.sz; ++z)
    for (i64 y = 0; y < c.sy; ++y)
      for (i64 x = 0; x < c.sx; ++x)
        if (!fg(x, y, z) && !seen[c.at(x, y, z)]) out.push_back(s.at(x + b.lo[0], y + b.lo[1], z + b.lo[2]));
  std::sort(out.begin(), out.end());
  return out;
}

// The soma root (find_soma_roots, batched.cu:58-108): among the label's voxels whose DBF equals
// its maximum, in box-relative coordinates sorted by (x, y, z), the first of least squared
// distance (float64) to their centroid (float32 of the float64 mean).
i64 soma_root(const std::vector<uint32_t> &cc, const std::vector<float> &dbf, const Shape &s, uint32_t id, float maximum,
              const Box &b) {
  std::vector<std::pair<std::array<i64, 3>, i64>> pts;
  for (i64 v = 0; v < s.n(); ++v)
    if (cc[v] == id && dbf[v] == maximum) {
      i64 x, y, z; s.unravel(v, x, y, z);
      pts.push_back({{x - b.lo[0], y - b.lo[1], z - b.lo[2]}, v});
    }
  std::sort(pts.begin(), pts.end());
  i64 sum[3] = {0, 0, 0};
  for (auto &p : pts) for (int a = 0; a < 3; ++a) sum[a] += p.first[a];
  float center[3];
  for (int a = 0; a < 3; ++a) center[a] = float(double(sum[a]) / double(pts.size()));
  double best = INFINITY;
  i64 root = -1;
  for (auto &p : pts) {
    double cost = 0;
    for (int a = 0; a < 3; ++a) { double d = double(p.first[a]) - center[a]; cost += d * d; }
    if (cost < best) { best = cost; root = p.second; }
  }
  return root;
}

// Vertex-weighted flood (batched.cu:246-251 with flood mode 1): the unique least fixpoint of
// d[v] = min_u vw_step(d[u], vw[v]) over same-label 26-neighbours. vw_step is monotone
// non-decreasing in d and adds at least one ulp for a positive weight, so Dijkstra reproduces it.
void vw_flood(const std::vector<uint32_t> &cc, const Shape &s, const std::vector<float> &vw,
              const std::vector<i64> &sources, std::vector<float> &dist, const Gate &gate = {}) {
  MinHeap heap;
  for (i64 src : sources) { dist[src] = 0.0f; heap.push({0.0f, src}); }
  while (!heap.empty()) {
    HeapItem it = heap.top(); heap.pop();
    if (it.d > dist[it.v]) continue;
    i64 x, y, z; s.unravel(it.v, x, y, z);
    uint32_t l = cc[it.v];
    for (int k = 0; k < 26; ++k) {
      i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
      if (!s.inside(nx, ny, nz)) continue;
      i64 u = s.at(nx, ny, nz);
      if (cc[u] != l || !gate.push(it.v, k)) continue;
      float c = vw_step(it.d, vw[u]);
      if (c < dist[u]) { dist[u] = c; heap.push({c, u}); }
    }
  }
}

// parents_lab, batched_regular.inc:113-134: the same-label neighbour with the smallest distance,
// first k on ties; -1 for background, inactive labels, unreached voxels and the sources (d 0).
void parent_field(const std::vector<uint32_t> &cc, const Shape &s, const std::vector<uint8_t> &active,
                  const std::vector<float> &dist, std::vector<int64_t> &parent, const Gate &gate = {}) {
  parent.assign(size_t(s.n()), -1);
  for (i64 v = 0; v < s.n(); ++v) {
    uint32_t l = cc[v];
    if (!l || !active[l] || !(dist[v] < FINF) || dist[v] == 0.0f) continue;
    i64 x, y, z; s.unravel(v, x, y, z);
    float best = FINF;
    i64 arg = -1;
    for (int k = 0; k < 26; ++k) {
      i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
      if (!s.inside(nx, ny, nz)) continue;
      i64 u = s.at(nx, ny, nz);
      if (cc[u] != l || !gate.pull(u, k)) continue;
      if (dist[u] < best) { best = dist[u]; arg = u; }
    }
    parent[v] = arg;
  }
}

// Geometric weights, batched.cu:211-214.
void geometric_weights(const float aniso[3], float W[26]) {
  double x = aniso[0], y = aniso[1], z = aniso[2];
  std::vector<float> w = {float(x), float(x), float(y), float(y), float(z), float(z)};
  for (double d : {std::sqrt(x * x + y * y), std::sqrt(y * y + z * z), std::sqrt(x * x + z * z)})
    for (int i = 0; i < 4; ++i) w.push_back(float(d));
  for (int i = 0; i < 8; ++i) w.push_back(float(std::sqrt(x * x + y * y + z * z)));
  for (int k = 0; k < 26; ++k) W[k] = w[k];
}

// ---------------------------------------------------------------------------------------------
struct Params {
  double scale = 10, constant = 10, pdrf_scale = 5000, pdrf_exponent = 16;
  double soma_detection = 1100, soma_acceptance = 4000, soma_scale = 0.5, soma_constant = 0;
  i64 max_paths = -1;  // < 0: none
  float aniso[3] = {1, 1, 1};
  double dust = 1000;
  bool fix_branching = true, fix_borders = true, fill_holes = false, fix_avocados = false;
  bool stop_emitted = true;  // fix_branching=False: walks end on the first emitted vertex (lockstep_regular.inc:160-165)
  // voxel_graph (skeletonize.cu:76-91): the tracing graph (a uint8 plane graph's values widened
  // as they are) and whether the input was such a plane graph (its CCL reads the corner bits moved)
  const uint32_t *graph = nullptr;
  bool graph_compact_2d = false;
  // A soma's private arena (soma.cu:54-72, 141-237): the run on its filled box takes the label's
  // border targets as given (the last one the root) and places vertices at origin + voxel.
  const std::vector<std::array<i64, 3>> *given_borders = nullptr;
  i64 origin[3] = {0, 0, 0};
  bool arena = false;  // the input is one filled component: no graph components, plain black-border rule
};

// A soma whose filled box takes part of another traced label is traced in a private arena: its
// filled box alone, while the main volume keeps every other label's voxels.
struct Arena {
  uint32_t id = 0;
  i64 lo[3] = {0, 0, 0};
  Shape c{0, 0, 0};
  std::vector<i64> labels;  // 1 on the filled soma, Fortran order over c
};

struct Prepared {
  Shape s;
  Components comp;
  std::vector<Box> boxes;
  std::vector<float> dbf, daf, pdrf;
  std::vector<uint8_t> active, soma;
  std::vector<float> dbf_max, max_daf;
  std::vector<i64> root, target;
  std::vector<std::vector<std::array<i64, 3>>> borders;
  std::vector<Arena> arenas;
};

// Per-label argmax, batched_regular.inc:9-22: largest finite value, ties -> lowest flat index.
static void seg_argmax(const std::vector<float> &val, const std::vector<uint32_t> &cc, size_t k,
                       std::vector<float> &value, std::vector<i64> &index) {
  std::vector<u64> best(k, 0);
  for (size_t i = 0; i < val.size(); ++i) {
    uint32_t l = cc[i];
    if (!l) continue;
    float v = val[i];
    if (!(v < FINF)) continue;
    u64 key = (u64(fbits(v)) << 32) | u64(0xFFFFFFFFu - uint32_t(i));
    if (key > best[l]) best[l] = key;
  }
  value.assign(k, 0.0f); index.assign(k, -1);
  for (size_t i = 0; i < k; ++i) {
    value[i] = bitsf(uint32_t(best[i] >> 32));
    index[i] = best[i] ? i64(0xFFFFFFFFu - uint32_t(best[i])) : -1;
  }
}

static void parallel_for(size_t count, int threads, const std::function<void(size_t)> &fn) {
  if (threads <= 1 || count <= 1) { for (size_t i = 0; i < count; ++i) fn(i); return; }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  std::exception_ptr error;
  std::mutex m;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      for (;;) {
        size_t i = next++;
        if (i >= count) return;
        try { fn(i); } catch (...) { std::lock_guard<std::mutex> g(m); if (!error) error = std::current_exception(); }
      }
    });
  for (auto &th : pool) th.join();
  if (error) std::rethrow_exception(error);
}

// prepare_sample (skeletonize.cu:70-116) + prepare_batched (batched.cu:177-255) + the somata of
// soma.cu (voids filled, enclosed labels evicted, refreshed DBF; roots, free-space seeded DAF).
Prepared prepare(const i64 *labels, const Shape &s, const Params &p, int threads) {
  Prepared out;
  out.s = s;
  const Gate gate{p.graph};
  std::vector<i64> keys;  // Kimimaro's ids of the graph components (avocado order)
  bool uniform = false;
  if (p.graph && !p.arena) {
    GraphComponents g = graph_ccl(labels, p.graph, p.graph_compact_2d, s);
    keys = std::move(g.keys);
    uniform = g.input_uniform;
    out.comp = std::move(g);
  } else {
    out.comp = ccl(labels, s);
  }
  size_t count = out.comp.roots.size();
  size_t k = count + 1;
  out.boxes = analyze(out.comp.cc, s, count);
  // of the volume before any fill; with a graph, every voxel of the input has one label
  bool black_border = p.graph && !p.arena ? uniform : count == 1 && out.boxes[1].count == s.n();
  auto field = [&](const std::vector<uint32_t> &lab, std::vector<float> &dist) {
    if (p.graph) graph_edt<uint32_t>(lab.data(), p.graph, s, p.aniso, black_border, dist.data());
    else edt_rule<uint32_t>(lab.data(), s, p.aniso, black_border, 3, dist.data());
  };
  if (p.fill_holes) {
    // fill_all_holes (holes.cu:37-62, skeletonize.cu:83-90): in ascending id and with the original
    // boxes, a component's box voids become the component; a component painted over, even in
    // part, fills nothing later; then the boxes and first voxels are recomputed.
    std::vector<uint8_t> skip(k, 0);
    bool painted = false;
    for (size_t id = 1; id < k; ++id) {
      if (!out.boxes[id].count || skip[id]) continue;
      std::vector<i64> voids = void_voxels(out.comp.cc, s, uint32_t(id), out.boxes[id]);
      for (i64 v : voids) {
        uint32_t o = out.comp.cc[v];
        if (o && o != id) skip[o] = 1;
        out.comp.cc[v] = uint32_t(id);
      }
      painted |= !voids.empty();
    }
    if (painted) {
      out.boxes = analyze(out.comp.cc, s, count);
      std::vector<i64> first(k, -1);
      for (i64 v = s.n() - 1; v >= 0; --v) first[out.comp.cc[v]] = v;
      for (size_t id = 1; id < k; ++id) out.comp.roots[id - 1] = first[id];
    }
  }
  out.dbf.resize(s.n());
  field(out.comp.cc, out.dbf);
  if (p.fix_avocados) {
    // skeletonize.cu:91-95: the correction (avocado_oracle.hpp), the field refreshed after each
    // changing pass, then fastremap.renumber with get_mapping's rule (the input label under each
    // run start of a new id, the last one kept)
    const std::vector<uint32_t> before = out.comp.cc;
    oracle_avocado::correct(out.comp.cc, out.dbf, {s.sx, s.sy, s.sz}, k, p.soma_detection, field, keys);
    std::vector<uint32_t> renum(k, 0);
    std::vector<i64> mapping{0}, roots;
    uint32_t prev = UINT32_MAX;
    for (i64 v = 0; v < s.n(); ++v) {
      const uint32_t l = out.comp.cc[v];
      if (l && !renum[l]) { renum[l] = uint32_t(roots.size() + 1); roots.push_back(v); mapping.push_back(0); }
      const uint32_t nl = l ? renum[l] : 0;
      if (nl != prev) {
        if (nl) mapping[nl] = before[v] ? out.comp.mapping[before[v]] : 0;
        prev = nl;
      }
      out.comp.cc[v] = nl;
    }
    out.comp.mapping = mapping;
    out.comp.roots = roots;
    count = roots.size();
    k = count + 1;
    out.boxes = analyze(out.comp.cc, s, count);
  }
  out.borders.assign(k, {});
  if (p.fix_borders) out.borders = border_targets(out.comp.cc, s, count, p.aniso);
  if (p.given_borders && count == 1) out.borders[1] = *p.given_borders;
  std::vector<i64> unused;
  seg_argmax(out.dbf, out.comp.cc, k, out.dbf_max, unused);
  out.active.assign(k, 0);
  out.soma.assign(k, 0);
  out.root.assign(k, -1);
  for (size_t id = 1; id < k; ++id) {
    const Box &b = out.boxes[id];
    i64 ext = (b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
    if (double(b.count) <= p.dust || ext <= 1) continue;
    out.active[id] = 1;
  }
  // Somata (batched.cu:194-205, soma.cu:141-237), in ascending id: a label above the detection
  // threshold whose box encloses voids takes them (fill_voids), evicting every traced label wholly
  // inside them (one partly inside is the reference's private-arena case: not restated); its DBF
  // becomes the EDT of its filled box (black border when the filled mask fills the box). A label
  // is a soma when its (refreshed) DBF maximum exceeds the acceptance threshold.
  std::vector<uint8_t> evicted(k, 0);
  for (size_t id = 1; id < k; ++id) {
    if (!out.active[id] || evicted[id] || !(double(out.dbf_max[id]) > p.soma_detection)) continue;
    Box &b = out.boxes[id];
    std::vector<i64> voids = void_voxels(out.comp.cc, s, uint32_t(id), b);
    if (!voids.empty()) {
      std::vector<i64> taken(k, 0);
      for (i64 v : voids) if (out.comp.cc[v]) taken[out.comp.cc[v]]++;
      auto owner = [&](size_t o) { return taken[o] && double(out.boxes[o].count) > p.dust && out.active[o] && !evicted[o]; };
      bool partial = false;
      for (size_t o = 1; o < k; ++o) partial |= owner(o) && taken[o] != out.boxes[o].count;
      if (partial) {  // soma.cu:207-215: the label goes to a private arena; nothing is evicted or painted
        Arena a;
        a.id = uint32_t(id);
        for (int ax = 0; ax < 3; ++ax) a.lo[ax] = b.lo[ax];
        a.c = Shape{b.hi[0] - b.lo[0] + 1, b.hi[1] - b.lo[1] + 1, b.hi[2] - b.lo[2] + 1};
        a.labels.assign(size_t(a.c.n()), 0);
        for (i64 z = 0; z < a.c.sz; ++z)
          for (i64 y = 0; y < a.c.sy; ++y)
            for (i64 x = 0; x < a.c.sx; ++x)
              a.labels[a.c.at(x, y, z)] = out.comp.cc[s.at(x + b.lo[0], y + b.lo[1], z + b.lo[2])] == id;
        for (i64 v : voids) {
          i64 x, y, z; s.unravel(v, x, y, z);
          a.labels[a.c.at(x - b.lo[0], y - b.lo[1], z - b.lo[2])] = 1;
        }
        out.arenas.push_back(std::move(a));
        out.active[id] = 0;
        continue;
      }
      for (size_t o = 1; o < k; ++o) {
        if (!owner(o)) continue;
        evicted[o] = 1;
        out.active[o] = 0;
      }
      for (i64 v : voids) out.comp.cc[v] = uint32_t(id);
      for (size_t o = 1; o < k; ++o) out.boxes[o].count -= taken[o];
      b.count += i64(voids.size());
      Shape c{b.hi[0] - b.lo[0] + 1, b.hi[1] - b.lo[1] + 1, b.hi[2] - b.lo[2] + 1};
      std::vector<uint8_t> mask(size_t(c.n()));
      for (i64 z = 0; z < c.sz; ++z)
        for (i64 y = 0; y < c.sy; ++y)
          for (i64 x = 0; x < c.sx; ++x)
            mask[c.at(x, y, z)] = out.comp.cc[s.at(x + b.lo[0], y + b.lo[1], z + b.lo[2])] == id;
      std::vector<float> D(size_t(c.n()));
      if (p.graph) {  // the graph EDT of the crop with the crop of the graph (trace.cu:107)
        std::vector<uint32_t> crop(size_t(c.n()));
        for (i64 z = 0; z < c.sz; ++z)
          for (i64 y = 0; y < c.sy; ++y)
            for (i64 x = 0; x < c.sx; ++x) crop[c.at(x, y, z)] = p.graph[s.at(x + b.lo[0], y + b.lo[1], z + b.lo[2])];
        graph_edt<uint8_t>(mask.data(), crop.data(), c, p.aniso, b.count == c.n(), D.data());
      } else {
        edt_rule<uint8_t>(mask.data(), c, p.aniso, b.count == c.n(), 3, D.data());
      }
      float maximum = 0.0f;
      for (i64 z = 0; z < c.sz; ++z)
        for (i64 y = 0; y < c.sy; ++y)
          for (i64 x = 0; x < c.sx; ++x) {
            i64 i = c.at(x, y, z);
            if (!mask[i]) continue;
            out.dbf[s.at(x + b.lo[0], y + b.lo[1], z + b.lo[2])] = D[i];
            maximum = std::max(maximum, D[i]);
          }
      out.dbf_max[id] = maximum;
    }
    out.soma[id] = double(out.dbf_max[id]) > p.soma_acceptance;
  }
  float W[26];
  geometric_weights(p.aniso, W);
  std::vector<i64> search;
  for (size_t id = 1; id < k; ++id)
    if (out.soma[id]) out.root[id] = soma_root(out.comp.cc, out.dbf, s, uint32_t(id), out.dbf_max[id], out.boxes[id]);
  for (size_t id = 1; id < k; ++id)
    if (out.active[id] && !out.soma[id]) {
      if (!out.borders[id].empty()) {
        auto &pt = out.borders[id].back();
        out.root[id] = s.at(pt[0], pt[1], pt[2]);
      } else search.push_back(id);
    }
  out.daf.assign(s.n(), FINF);
  if (!search.empty()) {
    // root search flood from each component's first voxel; labels are independent
    parallel_for(search.size(), threads, [&](size_t i) {
      size_t id = search[i];
      geometric_flood(out.comp.cc, s, W, {out.comp.roots[id - 1]}, out.daf, gate);
    });
    std::vector<float> v; std::vector<i64> idx;
    seg_argmax(out.daf, out.comp.cc, k, v, idx);
    for (size_t id : search) out.root[id] = idx[id];
    std::fill(out.daf.begin(), out.daf.end(), FINF);
  }
  std::vector<size_t> act;
  for (size_t id = 1; id < k; ++id) if (out.active[id]) act.push_back(id);
  // somata: the free-space ball around the root seeds the flood (batched.cu:222-231)
  std::vector<std::vector<i64>> members(k);
  for (i64 v = 0; v < s.n(); ++v)
    if (out.soma[out.comp.cc[v]]) members[out.comp.cc[v]].push_back(v);
  parallel_for(act.size(), threads, [&](size_t i) {
    size_t id = act[i];
    if (!out.soma[id]) { geometric_flood(out.comp.cc, s, W, {out.root[id]}, out.daf, gate); return; }
    i64 rx, ry, rz; s.unravel(out.root[id], rx, ry, rz);
    std::vector<std::pair<i64, float>> seeds = {{out.root[id], 0.0f}};
    for (i64 v : members[id]) {
      i64 x, y, z; s.unravel(v, x, y, z);
      float val;
      if (freespace_seed(x - rx, y - ry, z - rz, p.aniso[0], p.aniso[1], p.aniso[2], out.dbf_max[id], val)) seeds.push_back({v, val});
    }
    geometric_flood_seeded(out.comp.cc, s, W, seeds, out.daf, gate);
  });
  seg_argmax(out.daf, out.comp.cc, k, out.max_daf, out.target);
  double exponent = p.pdrf_exponent;
  if (!(exponent > 0 && exponent < 65536 && exponent == std::floor(exponent) && !(int(exponent) & (int(exponent) - 1))))
    throw std::runtime_error("pdrf_exponent must be a power of two for the lockstep path");
  std::vector<float> M(k, 0), inv(k, 0);
  std::vector<uint8_t> use(k, 0);
  for (size_t id = 1; id < k; ++id)
    if (out.active[id]) {
      M[id] = float(1.0 / std::pow(double(out.dbf_max[id]), 1.01));
      if (out.max_daf[id] != 0) { inv[id] = float(1.0 / double(out.max_daf[id])); use[id] = 2; }
      else use[id] = 1;
    }
  int squarings = 0;
  for (int e = int(exponent); e > 1; e >>= 1) ++squarings;
  float scale = float(p.pdrf_scale);
  out.pdrf.resize(s.n());
  // pdrf_lab, batched_regular.inc:93-108
  for (i64 i = 0; i < s.n(); ++i) {
    uint32_t l = out.comp.cc[i];
    if (l == 0 || use[l] == 0) { out.pdrf[i] = FINF; continue; }
    float d = out.dbf[i];
    if (d == 0.0f) d = FINF;
    float q = 1.0f - d * M[l];
    for (int j = 0; j < squarings; ++j) q = q * q;
    q = q * scale;
    if (use[l] == 2) { float a = out.daf[i]; if (std::isinf(a)) a = 0.0f; q = q + a * inv[l]; }
    out.pdrf[i] = q;
  }
  return out;
}

// ---------------------------------------------------------------------------------------------
// The per-label trace (lockstep.cu:91-622 restated label by label, which A.13 of the rule book
// shows to be exact: labels never interact).
struct Tracer {
  const Shape &s;
  const std::vector<uint32_t> &cc;
  const std::vector<float> &dbf;
  std::vector<float> &pdrf;
  std::vector<float> &dist;      // +inf between railroads
  std::vector<u64> &key;         // MAXKEY between invalidations
  std::vector<uint8_t> &alive;
  std::vector<uint8_t> &inq;     // 0 idle / 1 frontier / 2 waiting (invalidation)
  float wx, wy, wz, scale, constant, delta;
  // When set, each invalidation leaves here the final claim key of every voxel it claimed
  // ((bottleneck bits << 32) | source; MAXKEY elsewhere), for the stage-level checks.
  std::vector<u64> *claims = nullptr;
  Gate gate;  // a voxel graph's steps (pushes: railroad, join, offers; pulls: backtrace, band pulls)

  // Railroad (lockstep_coop.inc:71-305 with brook_vw_step) + backtrace and join
  // (lockstep_regular.inc:61-154). Returns the path target-first; empty: no rail reachable.
  std::vector<i64> railroad(uint32_t L, i64 t) {
    if (pdrf[t] == 0.0f) return {t};
    std::vector<i64> touched;
    MinHeap heap;
    dist[t] = 0.0f;
    touched.push_back(t);
    heap.push({0.0f, t});
    float rail_d = FINF;
    i64 rail_v = -1;
    while (!heap.empty()) {
      HeapItem it = heap.top(); heap.pop();
      if (it.d > dist[it.v]) continue;
      if (it.d > rail_d) break;
      i64 x, y, z; s.unravel(it.v, x, y, z);
      for (int k = 0; k < 26; ++k) {
        i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
        if (!s.inside(nx, ny, nz)) continue;
        i64 v = s.at(nx, ny, nz);
        if (cc[v] != L || !gate.push(it.v, k)) continue;
        float fv = pdrf[v];
        float c = vw_step(it.d, fv);
        if (c < dist[v]) {
          if (dist[v] == FINF) touched.push_back(v);
          dist[v] = c;
          if (fv == 0.0f) {
            if (c < rail_d || (c == rail_d && v < rail_v)) { rail_d = c; rail_v = v; }
          } else heap.push({c, v});
        }
      }
    }
    std::vector<i64> walk;
    if (rail_v >= 0) {
      i64 cur = rail_v;
      i64 x, y, z; s.unravel(cur, x, y, z);
      int cnt = 0;
      for (;;) {
        int kb = -1;
        if (cur != t) {
          float best = FINF;
          for (int k = 0; k < 26; ++k) {
            i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
            if (!s.inside(nx, ny, nz)) continue;
            i64 u = s.at(nx, ny, nz);
            if (cc[u] != L || pdrf[u] == 0.0f || !gate.pull(u, k)) continue;
            if (dist[u] < best) { best = dist[u]; kb = k; }
          }
        }
        if (cnt == 0 && kb >= 0) {  // the join: u's first rail in dijkstra3d order
          i64 ux = x + NX[kb], uy = y + NY[kb], uz = z + NZ[kb];
          const i64 uidx = s.at(ux, uy, uz);
          int r = 64;
          for (int k = 0; k < 26; ++k) {
            i64 nx = ux + NX[k], ny = uy + NY[k], nz = uz + NZ[k];
            if (!s.inside(nx, ny, nz)) continue;
            i64 v = s.at(nx, ny, nz);
            if (cc[v] == L && pdrf[v] == 0.0f && gate.push(uidx, k)) r = std::min(r, D3RANK[k]);
          }
          if (r < 26) {
            int ks = D3INV[r];
            x = ux + NX[ks]; y = uy + NY[ks]; z = uz + NZ[ks];
            cur = s.at(x, y, z);
            kb = OPP[ks];
          }
        }
        walk.push_back(cur);
        cnt++;
        if (kb < 0) break;
        x += NX[kb]; y += NY[kb]; z += NZ[kb];
        cur = s.at(x, y, z);
      }
      std::reverse(walk.begin(), walk.end());
    }
    for (i64 v : touched) dist[v] = FINF;
    return walk;
  }

  // Rolling-ball invalidation (invalidation_kernels.cuh:105-299; per label as lockstep_coop.inc:550-845):
  // bands of the bottleneck M on the fixed grid [k delta, (k+1) delta), synchronous rounds from a
  // snapshot inside a band, exact pulls of the waiting voxels at each band boundary.
  i64 invalidate(uint32_t L, const std::vector<i64> &path) { return invalidate(L, path, scale, constant); }
  i64 invalidate(uint32_t L, const std::vector<i64> &path, float ball_scale, float ball_constant) {
    std::vector<i64> touched, frontier, next, far, keep;
    for (i64 p : path) {
      if (!alive[p]) continue;
      if (key[p] == MAXKEY) {
        key[p] = u64(uint32_t(p));
        inq[p] = 1;
        frontier.push_back(p);
        touched.push_back(p);
      }
    }
    auto ball = [&](i64 src) { return ball_scale * dbf[src] + ball_constant; };
    auto offer_d = [&](i64 x, i64 y, i64 z, i64 src) {
      i64 sxx, syy, szz; s.unravel(src, sxx, syy, szz);
      float dvx = wx * (float)(x - sxx), dvy = wy * (float)(y - syy), dvz = wz * (float)(z - szz);
      return std::sqrt(dvx * dvx + dvy * dvy + dvz * dvz);
    };
    int band = 0;
    std::vector<u64> snap;
    for (;;) {
      uint32_t hi = fbits((float)(band + 1) * delta);
      while (!frontier.empty()) {
        snap.resize(frontier.size());
        for (size_t i = 0; i < frontier.size(); ++i) { snap[i] = key[frontier[i]]; inq[frontier[i]] = 0; }
        next.clear();
        for (size_t i = 0; i < frontier.size(); ++i) {
          i64 u = frontier[i];
          i64 su = i64(snap[i] & 0xffffffffULL);
          uint32_t mu = uint32_t(snap[i] >> 32);
          float maxd = ball(su);
          i64 x, y, z; s.unravel(u, x, y, z);
          for (int k = 0; k < 26; ++k) {
            i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
            if (!s.inside(nx, ny, nz)) continue;
            i64 v = s.at(nx, ny, nz);
            if (cc[v] != L || !alive[v] || !gate.push(u, k)) continue;
            float d = offer_d(nx, ny, nz, su);
            if (d >= maxd) continue;
            uint32_t mb = std::max(fbits(d), mu);
            if (mb < hi) {
              u64 cand = (u64(mb) << 32) | u64(uint32_t(su));
              if (cand < key[v]) {
                if (key[v] == MAXKEY) touched.push_back(v);
                key[v] = cand;
                if (inq[v] != 1) { inq[v] = 1; next.push_back(v); }
              }
            } else if (key[v] == MAXKEY && inq[v] == 0) {
              inq[v] = 2;
              far.push_back(v);
            }
          }
        }
        frontier.swap(next);
      }
      if (far.empty()) break;
      // band boundary: exact pulls from the claimed neighbours' keys
      std::vector<std::pair<i64, uint32_t>> pulled(far.size());  // (source or -1 claimed / -2 none, M bits)
      uint32_t lowest = 0x7f800000u;
      for (size_t i = 0; i < far.size(); ++i) {
        i64 c = far[i];
        if (key[c] != MAXKEY) { pulled[i] = {-1, 0}; continue; }
        i64 x, y, z; s.unravel(c, x, y, z);
        u64 best = MAXKEY;
        for (int k = 0; k < 26; ++k) {
          i64 nx = x + NX[k], ny = y + NY[k], nz = z + NZ[k];
          if (!s.inside(nx, ny, nz)) continue;
          i64 u = s.at(nx, ny, nz);
          if (cc[u] != L || !gate.pull(u, k)) continue;
          u64 ku = key[u];
          if (ku == MAXKEY) continue;
          i64 su = i64(ku & 0xffffffffULL);
          float d = offer_d(x, y, z, su);
          if (d >= ball(su)) continue;
          uint32_t mb = std::max(fbits(d), uint32_t(ku >> 32));
          u64 cand = (u64(mb) << 32) | u64(uint32_t(su));
          best = std::min(best, cand);
        }
        if (best == MAXKEY) { pulled[i] = {-2, 0}; continue; }
        pulled[i] = {i64(best & 0xffffffffULL), uint32_t(best >> 32)};
        lowest = std::min(lowest, uint32_t(best >> 32));
      }
      if (lowest != 0x7f800000u) {
        // inval_band (invalidation_kernels.cuh:98-104)
        float m = bitsf(lowest);
        int kb = (int)(m / delta);
        while ((float)(kb + 1) * delta <= m) kb++;
        while (kb > 0 && (float)kb * delta > m) kb--;
        band = std::max(kb, band + 1);
      }
      hi = fbits((float)(band + 1) * delta);
      keep.clear();
      frontier.clear();
      for (size_t i = 0; i < far.size(); ++i) {
        i64 c = far[i];
        if (pulled[i].first == -1) continue;
        if (pulled[i].first == -2) { inq[c] = 0; continue; }
        uint32_t mb = pulled[i].second;
        if (mb < hi) {
          key[c] = (u64(mb) << 32) | u64(uint32_t(pulled[i].first));
          touched.push_back(c);
          frontier.push_back(c);
          inq[c] = 1;
        } else keep.push_back(c);
      }
      far.swap(keep);
    }
    if (claims)
      for (i64 v : touched) (*claims)[v] = key[v];
    for (i64 v : touched) { alive[v] = 0; key[v] = MAXKEY; inq[v] = 0; }
    return i64(touched.size());
  }
};

struct Record { uint32_t cc; std::vector<i64> path; };

// Target order of one label's voxels (lockstep_order.cu:21-43, the rule of
// component_order_smoke.cpp:12-16): DAF descending with +-inf and +-0 as 0, ties by flat index
// descending.
void sort_target_order(std::vector<i64> &voxels, const float *daf) {
  std::sort(voxels.begin(), voxels.end(), [&](i64 a, i64 b) {
    float da = daf[a], db = daf[b];
    if (std::isinf(da) || da == 0) da = 0;
    if (std::isinf(db) || db == 0) db = 0;
    if (da != db) return da > db;
    return a > b;
  });
}

struct Skeleton {
  i64 label = 0;
  std::vector<std::array<float, 3>> vertices;
  std::vector<std::array<uint32_t, 2>> edges;
  std::vector<float> radii;
};

// consolidate, skeleton.cpp:13-40: vertices sorted and deduplicated, edges (min,max) sorted unique
// without self-loops, vertices on no edge dropped.
void consolidate(Skeleton &sk) {
  if (sk.vertices.empty() || sk.edges.empty()) { sk.vertices.clear(); sk.edges.clear(); sk.radii.clear(); return; }
  std::vector<uint32_t> order(sk.vertices.size()), inverse(sk.vertices.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = uint32_t(i);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return sk.vertices[a] < sk.vertices[b]; });
  std::vector<std::array<float, 3>> verts;
  std::vector<float> radii;
  for (uint32_t o : order) {
    if (verts.empty() || verts.back() != sk.vertices[o]) { verts.push_back(sk.vertices[o]); radii.push_back(sk.radii[o]); }
    inverse[o] = uint32_t(verts.size() - 1);
  }
  std::vector<std::array<uint32_t, 2>> edges;
  for (auto e : sk.edges) {
    e = {inverse[e[0]], inverse[e[1]]};
    if (e[0] > e[1]) std::swap(e[0], e[1]);
    if (e[0] != e[1]) edges.push_back(e);
  }
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  std::vector<uint8_t> used(verts.size(), 0);
  for (auto &e : edges) used[e[0]] = used[e[1]] = 1;
  sk.vertices.clear(); sk.radii.clear();
  for (size_t i = 0; i < verts.size(); ++i)
    if (used[i]) { inverse[i] = uint32_t(sk.vertices.size()); sk.vertices.push_back(verts[i]); sk.radii.push_back(radii[i]); }
  for (auto &e : edges) e = {inverse[e[0]], inverse[e[1]]};
  sk.edges = std::move(edges);
}

struct Result {
  std::vector<Skeleton> skeletons;
  i64 paths = 0, components = 0, invalidated = 0;
};

// skeletonize (skeletonize.cu:184-198): prepare, trace every eligible label, assemble and merge.
Result skeletonize(const i64 *labels, const Shape &s, const Params &p, const std::vector<std::array<i64, 3>> &before,
                   const std::vector<std::array<i64, 3>> &after_points, int threads) {
  Result res;
  if (!s.n() || double(s.n()) <= p.dust) return res;
  Prepared pr = prepare(labels, s, p, threads);
  const size_t count = pr.comp.roots.size(), k = count + 1;
  res.components = i64(count);
  if (!count) return res;
  // manual targets per label (lockstep.cu:116-137)
  std::vector<std::vector<i64>> manual(k), after(k);
  std::vector<std::vector<std::array<i64, 3>>> extra(k), extra_after(k);
  auto bucket = [&](const std::vector<std::array<i64, 3>> &points, std::vector<std::vector<std::array<i64, 3>>> &out) {
    for (auto pt : points) {
      std::array<i64, 3> q = pt;
      i64 shp[3] = {s.sx, s.sy, s.sz};
      for (int a = 0; a < 3; ++a) {
        if (q[a] < 0) q[a] += shp[a];
        if (q[a] < 0 || q[a] >= shp[a]) throw std::invalid_argument("extra target is outside the input volume");
      }
      const uint32_t id = pr.comp.cc[s.at(q[0], q[1], q[2])];
      const i64 flat = s.at(pt[0], pt[1], pt[2]);  // traced at the coordinates as written
      if (id && (flat < 0 || flat >= s.n() || pr.comp.cc[flat] != id))
        throw std::invalid_argument("extra target's coordinates do not flatten onto its component");
      out[id].push_back(pt);
    }
  };
  bucket(before, extra);
  bucket(after_points, extra_after);
  std::vector<i64> maxp(k, 0);
  std::vector<size_t> labels_list;
  for (size_t id = 1; id < k; ++id) {
    if (!pr.active[id]) continue;
    labels_list.push_back(id);
    std::vector<std::array<i64, 3>> m = pr.borders[id];
    if (!m.empty()) m.pop_back();
    if (pr.soma[id] && !pr.borders[id].empty()) m.insert(m.begin(), pr.borders[id].back());
    m.insert(m.end(), extra[id].begin(), extra[id].end());
    if (m.empty() && !pr.soma[id]) manual[id].push_back(pr.target[id]);
    else for (auto it = m.rbegin(); it != m.rend(); ++it) manual[id].push_back(s.at((*it)[0], (*it)[1], (*it)[2]));
    // the targets after the DAF order, taken from the back (trace.cu:206-212)
    for (auto it = extra_after[id].rbegin(); it != extra_after[id].rend(); ++it)
      after[id].push_back(s.at((*it)[0], (*it)[1], (*it)[2]));
    i64 budget = p.max_paths >= 0 ? p.max_paths : pr.boxes[id].count;
    maxp[id] = i64(manual[id].size() + after[id].size()) >= budget ? 0 : std::min<i64>(INT32_MAX, budget);
  }
  // target order (lockstep_order.cu:21-43): DAF descending (inf and 0 as 0), flat index descending
  std::vector<std::vector<i64>> order(k);
  for (i64 v = 0; v < s.n(); ++v) {
    uint32_t l = pr.comp.cc[v];
    if (l && pr.active[l]) order[l].push_back(v);
  }
  std::vector<float> dist(s.n(), FINF);
  std::vector<u64> key(s.n(), MAXKEY);
  std::vector<uint8_t> alive(s.n(), 0), inq(s.n(), 0);
  for (i64 v = 0; v < s.n(); ++v) alive[v] = pr.active[pr.comp.cc[v]] ? 1 : 0;
  // fix_branching=False: one vertex-weighted flood from the roots over the PDRF and the parent
  // field (batched.cu:246-251); paths are parent walks (lockstep_regular.inc:166-202).
  std::vector<int64_t> parent;
  if (!p.fix_branching) {
    std::vector<float> vdist(s.n(), FINF);
    std::vector<i64> roots;
    for (size_t id : labels_list) roots.push_back(pr.root[id]);
    vw_flood(pr.comp.cc, s, pr.pdrf, roots, vdist, Gate{p.graph});
    parent_field(pr.comp.cc, s, pr.active, vdist, parent, Gate{p.graph});
  }
  // the root becomes a rail (lockstep.cu:183)
  for (size_t id : labels_list) pr.pdrf[pr.root[id]] = 0.0f;
  const float aniso_min = std::min({p.aniso[0], p.aniso[1], p.aniso[2]});
  std::vector<std::vector<Record>> records(k);
  std::vector<i64> inval_counts(k, 0);
  parallel_for(labels_list.size(), threads, [&](size_t li) {
    uint32_t L = uint32_t(labels_list[li]);
    auto &ord = order[L];
    sort_target_order(ord, pr.daf.data());
    Tracer tr{s, pr.comp.cc, pr.dbf, pr.pdrf, dist, key, alive, inq, p.aniso[0], p.aniso[1], p.aniso[2],
              float(p.scale), float(p.constant), aniso_min};
    tr.gate = Gate{p.graph};
    size_t cursor = 0;
    i64 pcount = 0;
    // a soma: its ball dies before the first path and the budget shrinks by what it took
    // (lockstep.cu:197-206); paths are trimmed to the vertices beyond the soma radius
    // (lockstep_regular.inc:35-45, 128-147, 184-195)
    const bool soma = pr.soma[L] != 0;
    const double radius = soma ? double(pr.dbf_max[L]) * p.soma_scale + p.soma_constant : -1.0;
    i64 rx = 0, ry = 0, rz = 0;
    if (soma) {
      s.unravel(pr.root[L], rx, ry, rz);
      i64 removed = tr.invalidate(L, {pr.root[L]}, float(p.soma_scale), float(p.soma_constant));
      inval_counts[L] += removed;
      i64 budget = p.max_paths >= 0 ? p.max_paths : pr.boxes[L].count - removed;
      maxp[L] = i64(manual[L].size() + after[L].size()) >= budget ? 0 : std::min<i64>(INT32_MAX, budget);
    }
    auto far = [&](i64 v) {
      i64 x, y, z; s.unravel(v, x, y, z);
      double dx = double(p.aniso[0]) * (double(x) - double(rx)), dy = double(p.aniso[1]) * (double(y) - double(ry)),
             dz = double(p.aniso[2]) * (double(z) - double(rz));
      double ss = dx * dx + dy * dy;
      ss = ss + dz * dz;
      return std::sqrt(ss) > radius;
    };
    size_t after_used = 0;
    for (;;) {
      if (pcount >= maxp[L]) break;
      i64 t;
      bool ordered = false;
      if (pcount < i64(manual[L].size())) t = manual[L][pcount];
      else {
        while (cursor < ord.size() && !alive[ord[cursor]]) cursor++;
        if (cursor < ord.size()) t = ord[cursor], ordered = true;
        else if (after_used < after[L].size()) t = after[L][after_used++];  // nothing alive is left
        else break;
      }
      std::vector<i64> path;
      if (p.fix_branching) {
        path = tr.railroad(L, t);
        if (path.empty()) {  // no rail reachable (trace.cu:218-224): an ordered target ends the label,
          if (ordered) break;  // a manual one is an empty path
          pcount++;
          continue;
        }
      } else {
        // backtrace_parent_lab (lockstep_regular.inc:166-202): the walk keeps every vertex up to
        // and including the first with no parent (-1 at the root, STOP -2 at an emitted vertex).
        for (i64 cur = t;;) {
          path.push_back(cur);
          i64 par = cur == pr.root[L] ? -1 : parent[cur];
          if (par < 0) break;
          cur = par;
        }
      }
      if (soma) {
        std::vector<i64> kept;
        if (p.fix_branching) {  // target-first: the target, the far vertices, the root for a near rail end
          kept.push_back(path.front());
          for (size_t j = 1; j < path.size(); ++j) if (far(path[j])) kept.push_back(path[j]);
          if (!far(path.back())) kept.push_back(pr.root[L]);
        } else {  // target .. root: the far vertices and the root
          for (size_t j = 0; j + 1 < path.size(); ++j) if (far(path[j])) kept.push_back(path[j]);
          kept.push_back(path.back());
        }
        path.swap(kept);
      }
      records[L].push_back({L, path});
      if (p.fix_branching) for (i64 v : path) pr.pdrf[v] = 0.0f;
      else if (p.stop_emitted && !soma) for (i64 v : path) parent[v] = -2;
      inval_counts[L] += tr.invalidate(L, path);
      pcount++;
    }
  });
  // assembly (assembly.cu:20-160) then same-label merge in CC order (skeletonize.cu:126-170)
  std::vector<Skeleton> per_cc(k);
  for (size_t id : labels_list) {
    Skeleton &sk = per_cc[id];
    for (auto &rec : records[id]) {
      res.paths++;
      if (rec.path.size() < 2) continue;
      size_t base = sk.vertices.size();
      for (size_t j = 0; j < rec.path.size(); ++j) {
        i64 x, y, z; s.unravel(rec.path[j], x, y, z);
        sk.vertices.push_back({float(x + p.origin[0]), float(y + p.origin[1]), float(z + p.origin[2])});
        sk.radii.push_back(pr.dbf[rec.path[j]]);
        if (j) sk.edges.push_back({uint32_t(base + j - 1), uint32_t(base + j)});
      }
    }
    consolidate(sk);
    for (auto &v : sk.vertices) for (int a = 0; a < 3; ++a) v[a] *= p.aniso[a];
    res.invalidated += inval_counts[id];
  }
  // private arenas (soma.cu:54-139): each soma runs alone on its filled box, with its border and
  // manual targets moved into the box, its vertices placed back at the box's origin
  for (const Arena &a : pr.arenas) {
    Params q = p;
    q.fix_borders = false;
    q.fill_holes = false;
    q.fix_avocados = false;
    auto shift = [&](std::vector<std::array<i64, 3>> pts) {
      for (auto &pt : pts) for (int ax = 0; ax < 3; ++ax) pt[ax] -= a.lo[ax];
      return pts;
    };
    const std::vector<std::array<i64, 3>> borders = shift(pr.borders[a.id]);
    q.given_borders = &borders;
    q.arena = true;
    for (int ax = 0; ax < 3; ++ax) q.origin[ax] = p.origin[ax] + a.lo[ax];
    std::vector<uint32_t> graph_crop;
    if (p.graph) {
      graph_crop.resize(size_t(a.c.n()));
      for (i64 z = 0; z < a.c.sz; ++z)
        for (i64 y = 0; y < a.c.sy; ++y)
          for (i64 x = 0; x < a.c.sx; ++x) graph_crop[a.c.at(x, y, z)] = p.graph[s.at(x + a.lo[0], y + a.lo[1], z + a.lo[2])];
      q.graph = graph_crop.data();
      q.graph_compact_2d = false;
    }
    Result r = skeletonize(a.labels.data(), a.c, q, shift(extra[a.id]), shift(extra_after[a.id]), threads);
    res.paths += r.paths;
    res.invalidated += r.invalidated;
    if (!r.skeletons.empty()) per_cc[a.id] = std::move(r.skeletons[0]);
  }
  std::unordered_map<i64, size_t> index;
  for (size_t id = 1; id < k; ++id) {
    Skeleton &sk = per_cc[id];
    if (sk.edges.empty()) continue;
    sk.label = pr.comp.mapping[id];
    auto it = index.find(sk.label);
    if (it == index.end()) { index[sk.label] = res.skeletons.size(); res.skeletons.push_back(std::move(sk)); }
    else {
      Skeleton &o = res.skeletons[it->second];
      size_t base = o.vertices.size();
      o.vertices.insert(o.vertices.end(), sk.vertices.begin(), sk.vertices.end());
      o.radii.insert(o.radii.end(), sk.radii.begin(), sk.radii.end());
      for (auto e : sk.edges) o.edges.push_back({uint32_t(base + e[0]), uint32_t(base + e[1])});
    }
  }
  for (auto &sk : res.skeletons) consolidate(sk);
  return res;
}
}  // namespace oracle

// =============================================================================================
// C ABI for ctypes (tests and the CPU baseline only).
using namespace oracle;

namespace {
thread_local std::string g_error;
template <class F> int guarded(F &&fn) {
  try { g_error.clear(); fn(); return 0; }
  catch (const std::exception &e) { g_error = e.what(); return -1; }
  catch (...) { g_error = "unknown error"; return -1; }
}
struct PackedOut {
  std::vector<int64_t> labels, v_off{0}, e_off{0};
  std::vector<float> vertices, radii;
  std::vector<uint32_t> edges;
  int64_t paths = 0, components = 0, invalidated = 0;
};
}  // namespace

extern "C" {
const char *oracle_last_error() { return g_error.c_str(); }

// labels: int64 values (Fortran order, shape sx,sy,sz); cc_out n; mapping_out n+1; roots_out n.
int64_t oracle_ccl(const int64_t *labels, int64_t sx, int64_t sy, int64_t sz, uint32_t *cc_out, int64_t *mapping_out,
                   int64_t *roots_out) {
  int64_t count = -1;
  guarded([&] {
    Shape s{sx, sy, sz};
    Components c = ccl(labels, s);
    std::copy(c.cc.begin(), c.cc.end(), cc_out);
    std::copy(c.mapping.begin(), c.mapping.end(), mapping_out);
    std::copy(c.roots.begin(), c.roots.end(), roots_out);
    count = int64_t(c.roots.size());
  });
  return count;
}

// Voxel-graph components: returns the count; keys_out holds count + 1 entries, uniform_out one.
int64_t oracle_graph_ccl(const int64_t *labels, int64_t sx, int64_t sy, int64_t sz, const uint32_t *graph, int compact_2d,
                         uint32_t *cc_out, int64_t *mapping_out, int64_t *roots_out, int64_t *keys_out,
                         int32_t *uniform_out) {
  int64_t count = -1;
  guarded([&] {
    GraphComponents c = graph_ccl(labels, graph, compact_2d != 0, Shape{sx, sy, sz});
    std::copy(c.cc.begin(), c.cc.end(), cc_out);
    std::copy(c.mapping.begin(), c.mapping.end(), mapping_out);
    std::copy(c.roots.begin(), c.roots.end(), roots_out);
    std::copy(c.keys.begin(), c.keys.end(), keys_out);
    *uniform_out = c.input_uniform ? 1 : 0;
    count = int64_t(c.roots.size());
  });
  return count;
}

int oracle_graph_edt(const uint32_t *cc, int64_t sx, int64_t sy, int64_t sz, const uint32_t *graph, const float *aniso,
                     int black_border, float *out) {
  return guarded([&] { graph_edt<uint32_t>(cc, graph, Shape{sx, sy, sz}, aniso, black_border != 0, out); });
}

int oracle_edt(const uint32_t *cc, int64_t sx, int64_t sy, int64_t sz, const float *aniso, int black_border, int dims,
               float *out) {
  return guarded([&] { edt_rule<uint32_t>(cc, Shape{sx, sy, sz}, aniso, black_border != 0, dims, out); });
}

int oracle_edt_f64(const uint32_t *cc, int64_t sx, int64_t sy, int64_t sz, const double *aniso, int black_border, int dims,
                   double *out) {
  return guarded([&] { edt_exact_f64<uint32_t>(cc, Shape{sx, sy, sz}, aniso, black_border != 0, dims, out); });
}

// CPython-set order of inserted 3-tuples: returns the iteration order as indices into `points`.
int oracle_pointset_order(const int64_t *points, int64_t count, int64_t *order_out) {
  return guarded([&] {
    PointSet st;
    std::map<std::array<int64_t, 3>, int64_t> where;
    for (int64_t i = 0; i < count; ++i) {
      std::array<int64_t, 3> p = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
      st.add(p);
      if (!where.count(p)) where[p] = i;
    }
    auto vals = st.values();
    for (size_t i = 0; i < vals.size(); ++i) order_out[i] = where.at(vals[i]);
  });
}

// Border targets of a component volume: points_out (capacity 3*cap int64), offsets_out count+2.
int64_t oracle_border_targets(const uint32_t *cc, int64_t sx, int64_t sy, int64_t sz, int64_t count, const float *aniso,
                              int64_t *points_out, int64_t cap, int64_t *offsets_out) {
  int64_t total = -1;
  guarded([&] {
    Shape s{sx, sy, sz};
    std::vector<uint32_t> v(cc, cc + s.n());
    auto b = border_targets(v, s, size_t(count), aniso);
    int64_t pos = 0;
    offsets_out[0] = 0;
    for (size_t id = 0; id < b.size(); ++id) {
      for (auto &p : b[id]) {
        if (pos >= cap) throw std::runtime_error("border point capacity");
        for (int a = 0; a < 3; ++a) points_out[3 * pos + a] = p[a];
        pos++;
      }
      offsets_out[id + 1] = pos;
    }
    total = pos;
  });
  return total;
}

// Geometric flood from sources (one label gate per source's component): dist_out n floats.
int oracle_geometric_flood(const uint32_t *cc, int64_t sx, int64_t sy, int64_t sz, const float *aniso,
                           const int64_t *sources, int64_t nsources, float *dist_out) {
  return guarded([&] {
    Shape s{sx, sy, sz};
    std::vector<uint32_t> v(cc, cc + s.n());
    std::vector<float> d(s.n(), FINF);
    float W[26];
    geometric_weights(aniso, W);
    geometric_flood(v, s, W, std::vector<i64>(sources, sources + nsources), d);
    std::copy(d.begin(), d.end(), dist_out);
  });
}

// Target orders of every nonzero id (ascending id, each the id's voxels in the order of
// sort_target_order), concatenated into order_out (one entry per nonzero voxel). Returns the count.
int64_t oracle_target_order(const uint32_t *cc, int64_t n, const float *daf, int64_t *order_out) {
  int64_t total = -1;
  guarded([&] {
    uint32_t top = 0;
    for (int64_t v = 0; v < n; ++v) top = std::max(top, cc[v]);
    std::vector<std::vector<i64>> order(size_t(top) + 1);
    for (int64_t v = 0; v < n; ++v)
      if (cc[v]) order[cc[v]].push_back(v);
    int64_t pos = 0;
    for (auto &ord : order) {
      sort_target_order(ord, daf);
      for (i64 v : ord) order_out[pos++] = v;
    }
    total = pos;
  });
  return total;
}

// One rolling-ball invalidation of `path` (voxel list) in label cc[path[0]]: alive (n bytes) is
// updated in place (claimed voxels die) and claims_out (n u64, optional) receives each claimed
// voxel's final key ((bottleneck bits << 32) | source), MAXKEY elsewhere. Returns the claim count.
int64_t oracle_invalidate(const uint32_t *cc, int64_t sx, int64_t sy, int64_t sz, const float *dbf, uint8_t *alive,
                          const int64_t *path, int64_t len, const float *aniso, float ball_scale, float ball_constant,
                          uint64_t *claims_out) {
  int64_t claimed = -1;
  guarded([&] {
    Shape s{sx, sy, sz};
    if (len <= 0) throw std::invalid_argument("empty path");
    std::vector<uint32_t> c(cc, cc + s.n());
    std::vector<float> d(dbf, dbf + s.n()), f(s.n(), 1.0f), dist(s.n(), FINF);
    std::vector<u64> key(s.n(), MAXKEY), claims(s.n(), MAXKEY);
    std::vector<uint8_t> al(alive, alive + s.n()), inq(s.n(), 0);
    const float delta = std::min({aniso[0], aniso[1], aniso[2]});
    Tracer tr{s, c, d, f, dist, key, al, inq, aniso[0], aniso[1], aniso[2], ball_scale, ball_constant, delta};
    tr.claims = &claims;
    claimed = tr.invalidate(c[path[0]], std::vector<i64>(path, path + len));
    std::copy(al.begin(), al.end(), alive);
    if (claims_out) std::copy(claims.begin(), claims.end(), claims_out);
  });
  return claimed;
}

// Per-id boxes and voxel counts of a component volume with ids 1..count: lo/hi 3*(count+1) int64
// (lo INT32_MAX and hi -1 for an absent id), counts count+1 (counts[0]: background).
int oracle_analyze(const uint32_t *cc, int64_t sx, int64_t sy, int64_t sz, int64_t count, int64_t *lo, int64_t *hi,
                   int64_t *counts) {
  return guarded([&] {
    Shape s{sx, sy, sz};
    std::vector<uint32_t> c(cc, cc + s.n());
    for (uint32_t id : c)
      if (int64_t(id) > count) throw std::invalid_argument("id above count");
    auto boxes = analyze(c, s, size_t(count));
    for (size_t i = 0; i < boxes.size(); ++i) {
      for (int a = 0; a < 3; ++a) { lo[3 * i + a] = boxes[i].lo[a]; hi[3 * i + a] = boxes[i].hi[a]; }
      counts[i] = boxes[i].count;
    }
  });
}

// One railroad on a single-label field (cc: nonzero = the label's voxels): path target-first.
int64_t oracle_railroad(const float *field, const uint32_t *cc, int64_t sx, int64_t sy, int64_t sz, int64_t target,
                        int64_t *path_out, int64_t cap) {
  int64_t len = -1;
  guarded([&] {
    Shape s{sx, sy, sz};
    std::vector<uint32_t> c(cc, cc + s.n());
    std::vector<float> f(field, field + s.n()), dist(s.n(), FINF), dbf(s.n(), 1.0f);
    std::vector<u64> key(s.n(), MAXKEY);
    std::vector<uint8_t> alive(s.n(), 1), inq(s.n(), 0);
    Tracer tr{s, c, dbf, f, dist, key, alive, inq, 1, 1, 1, 1, 1, 1};
    auto path = tr.railroad(c[target], target);
    if (int64_t(path.size()) > cap) throw std::runtime_error("path capacity");
    std::copy(path.begin(), path.end(), path_out);
    len = int64_t(path.size());
  });
  return len;
}

// Prepared fields of a volume: dbf, daf, pdrf (n floats each; pdrf with roots still unrailed),
// per-id tables (count+1): active, dbf_max, max_daf, root, target. Returns count (or -1).
int64_t oracle_prepare(const int64_t *labels, int64_t sx, int64_t sy, int64_t sz, const double *params /*8*/,
                       const float *aniso, double dust, int fix_borders, int threads, uint32_t *cc_out, float *dbf_out,
                       float *daf_out, float *pdrf_out, int64_t table_cap, uint8_t *active_out, float *dbf_max_out,
                       float *max_daf_out, int64_t *root_out, int64_t *target_out, uint8_t *soma_out) {
  int64_t count = -1;
  guarded([&] {
    Params p;
    p.scale = params[0]; p.constant = params[1]; p.pdrf_scale = params[2]; p.pdrf_exponent = params[3];
    p.soma_detection = params[4]; p.soma_acceptance = params[5]; p.soma_scale = params[6]; p.soma_constant = params[7];
    std::copy(aniso, aniso + 3, p.aniso);
    p.dust = dust;
    p.fix_borders = (fix_borders & 1) != 0;  // bit 0: fix_borders, bit 1: fill_holes, bit 2: fix_avocados
    p.fill_holes = (fix_borders & 2) != 0;
    p.fix_avocados = (fix_borders & 4) != 0;
    Shape s{sx, sy, sz};
    Prepared pr = prepare(labels, s, p, threads);
    size_t k = pr.comp.roots.size() + 1;
    if (int64_t(k) > table_cap) throw std::runtime_error("table capacity");
    std::copy(pr.comp.cc.begin(), pr.comp.cc.end(), cc_out);
    std::copy(pr.dbf.begin(), pr.dbf.end(), dbf_out);
    std::copy(pr.daf.begin(), pr.daf.end(), daf_out);
    std::copy(pr.pdrf.begin(), pr.pdrf.end(), pdrf_out);
    for (size_t i = 0; i < k; ++i) {
      active_out[i] = pr.active[i]; dbf_max_out[i] = pr.dbf_max[i]; max_daf_out[i] = pr.max_daf[i];
      root_out[i] = pr.root[i]; target_out[i] = pr.target[i];
      if (soma_out) soma_out[i] = pr.soma[i];
    }
    count = int64_t(k - 1);
  });
  return count;
}

// Vertex-weighted flood from `sources` over the weights `vw` and the parent field (all labels active).
int oracle_vw_parents(const int64_t *labels, int64_t sx, int64_t sy, int64_t sz, const float *vw, const int64_t *sources,
                      int64_t nsources, float *dist_out, int64_t *parent_out) {
  return guarded([&] {
    Shape s{sx, sy, sz};
    Components c = ccl(labels, s);
    std::vector<float> w(vw, vw + s.n()), dist(s.n(), FINF);
    std::vector<uint8_t> active(c.roots.size() + 1, 1);
    active[0] = 0;
    vw_flood(c.cc, s, w, std::vector<i64>(sources, sources + nsources), dist);
    std::vector<int64_t> parent;
    parent_field(c.cc, s, active, dist, parent);
    std::copy(dist.begin(), dist.end(), dist_out);
    std::copy(parent.begin(), parent.end(), parent_out);
  });
}

// Whole path. params: scale, const, pdrf_scale, pdrf_exponent, soma_detection, soma_acceptance,
// soma_scale, soma_const; max_paths < 0 = none; fix_branching 1, or 0 (parent walks with the
// STOP rule), or -1 (parent walks to the root, Kimimaro's paths; for tests). Returns an opaque
// handle (null on error).
void *oracle_skeletonize(const int64_t *labels, int64_t sx, int64_t sy, int64_t sz, const double *params,
                         int64_t max_paths, const float *aniso, double dust, int fix_branching, int fix_borders,
                         const int64_t *targets_before, int64_t ntargets, const int64_t *targets_after,
                         int64_t nafter, int threads, const uint32_t *graph, int graph_compact_2d) {
  PackedOut *out = nullptr;
  guarded([&] {
    Params p;
    p.graph = graph;
    p.graph_compact_2d = graph_compact_2d != 0;
    p.scale = params[0]; p.constant = params[1]; p.pdrf_scale = params[2]; p.pdrf_exponent = params[3];
    p.soma_detection = params[4]; p.soma_acceptance = params[5]; p.soma_scale = params[6]; p.soma_constant = params[7];
    p.max_paths = max_paths;
    std::copy(aniso, aniso + 3, p.aniso);
    p.dust = dust;
    p.fix_branching = fix_branching > 0;  // 1 fix, 0 parent walks with the STOP rule, -1 full walks
    p.stop_emitted = fix_branching == 0;
    p.fix_borders = (fix_borders & 1) != 0;  // bit 0: fix_borders, bit 1: fill_holes, bit 2: fix_avocados
    p.fill_holes = (fix_borders & 2) != 0;
    p.fix_avocados = (fix_borders & 4) != 0;
    std::vector<std::array<i64, 3>> before;
    for (int64_t i = 0; i < ntargets; ++i) before.push_back({targets_before[3 * i], targets_before[3 * i + 1], targets_before[3 * i + 2]});
    std::vector<std::array<i64, 3>> after;
    for (int64_t i = 0; i < nafter; ++i) after.push_back({targets_after[3 * i], targets_after[3 * i + 1], targets_after[3 * i + 2]});
    Result r = skeletonize(labels, Shape{sx, sy, sz}, p, before, after, threads);
    auto o = std::make_unique<PackedOut>();
    for (auto &sk : r.skeletons) {
      o->labels.push_back(sk.label);
      for (auto &v : sk.vertices) o->vertices.insert(o->vertices.end(), v.begin(), v.end());
      o->radii.insert(o->radii.end(), sk.radii.begin(), sk.radii.end());
      for (auto &e : sk.edges) o->edges.insert(o->edges.end(), e.begin(), e.end());
      o->v_off.push_back(int64_t(o->radii.size()));
      o->e_off.push_back(int64_t(o->edges.size() / 2));
    }
    o->paths = r.paths; o->components = r.components; o->invalidated = r.invalidated;
    out = o.release();
  });
  return out;
}

// sizes[0..5] = skeletons, vertices, edges, paths, components, invalidated voxels
void oracle_result_sizes(void *h, int64_t *sizes) {
  auto *o = static_cast<PackedOut *>(h);
  sizes[0] = int64_t(o->labels.size()); sizes[1] = int64_t(o->radii.size()); sizes[2] = int64_t(o->edges.size() / 2);
  sizes[3] = o->paths; sizes[4] = o->components; sizes[5] = o->invalidated;
}
void oracle_result_copy(void *h, int64_t *labels, int64_t *v_off, int64_t *e_off, float *vertices, uint32_t *edges,
                        float *radii) {
  auto *o = static_cast<PackedOut *>(h);
  std::copy(o->labels.begin(), o->labels.end(), labels);
  std::copy(o->v_off.begin(), o->v_off.end(), v_off);
  std::copy(o->e_off.begin(), o->e_off.end(), e_off);
  std::copy(o->vertices.begin(), o->vertices.end(), vertices);
  std::copy(o->edges.begin(), o->edges.end(), edges);
  std::copy(o->radii.begin(), o->radii.end(), radii);
}
void oracle_result_free(void *h) { delete static_cast<PackedOut *>(h); }
}
