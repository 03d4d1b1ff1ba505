// Synthetic label volumes for the benchmark workloads (host C++17, no GPU).
//
//   tubes: each label is a tube along a random 3-D polyline (5-10 segments of 10-40 voxels,
//          radius U[3,8]) with 1-3 branches that leave from random polyline points.
//   trees: each label is a random branching tree grown under a voxel budget: segments are
//          random walks of U[20,200] unit steps (direction <- normalize(direction + N(0, 0.35^2 I))),
//          each segment end spawns a branch with probability 0.1, radius U[2,10] per segment;
//          10% of the labels carry a soma ball of radius U[15,30] at the root.
// Walks reflect at the volume walls; balls are painted at every step (0.5 voxel along tube
// segments, 1 voxel along tree walks); the first writer of a voxel keeps it, so later labels
// never cut earlier ones. Output is Fortran order (x fastest), uint32 labels 1..L.
//
//   cells (the stress workload): Poisson(mean) cell seeds uniform in the volume; a voxel belongs to
//          its nearest seed under the physical metric (anisotropy); a voxel with a 6-neighbour in
//          another cell is dropped (erosion by one voxel); inside each cell 1-4 tubes of radius
//          U[2,6] (2-4 segments of U[10,30] voxels from a point near the seed), painted only on
//          the cell's own voxels, uint64 ids cell << 32 | tube (cells from 1, tubes 1..4).
//          Cells are painted by worker threads; a cell writes only voxels it owns, so the result
//          does not depend on the thread count.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (auto &v : s) {  // splitmix64 seeding
      seed += 0x9E3779B97F4A7C15ull;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      v = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {  // xoshiro256**
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return double(next() >> 11) * 0x1.0p-53; }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  int64_t integer(int64_t a, int64_t b) { return a + int64_t(next() % uint64_t(b - a + 1)); }  // inclusive
  double normal() {
    double u1 = std::max(uniform(), 1e-300), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  void direction(double d[3]) {
    double n = 0;
    do {
      for (int a = 0; a < 3; ++a) d[a] = normal();
      n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    } while (n < 1e-9);
    for (int a = 0; a < 3; ++a) d[a] /= n;
  }
};

struct Volume {
  int64_t dim[3];
  uint32_t *data;
  int64_t painted = 0;
  void ball(const double c[3], double r, uint32_t label) {
    int64_t lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::max<int64_t>(0, int64_t(std::ceil(c[a] - r)));
      hi[a] = std::min<int64_t>(dim[a] - 1, int64_t(std::floor(c[a] + r)));
      if (lo[a] > hi[a]) return;
    }
    double r2 = r * r;
    for (int64_t z = lo[2]; z <= hi[2]; ++z) {
      double dz = double(z) - c[2];
      for (int64_t y = lo[1]; y <= hi[1]; ++y) {
        double dy = double(y) - c[1], dyz = dy * dy + dz * dz;
        if (dyz > r2) continue;
        uint32_t *row = data + dim[0] * (y + dim[1] * z);
        for (int64_t x = lo[0]; x <= hi[0]; ++x) {
          double dx = double(x) - c[0];
          if (dx * dx + dyz <= r2 && row[x] == 0) { row[x] = label; ++painted; }
        }
      }
    }
  }
  void step(double p[3], double d[3], double h) {
    for (int a = 0; a < 3; ++a) {
      p[a] += h * d[a];
      double top = double(dim[a] - 1);
      if (p[a] < 0) { p[a] = -p[a]; d[a] = -d[a]; }
      if (p[a] > top) { p[a] = 2 * top - p[a]; d[a] = -d[a]; }
      p[a] = std::min(std::max(p[a], 0.0), top);
    }
  }
};

// A tube label: polyline with branches, painted every half voxel.
void tube(Volume &vol, Rng &rng, uint32_t label) {
  double r = rng.uniform(3.0, 8.0);
  std::vector<std::array<double, 3>> points;
  auto polyline = [&](double start[3]) {
    double p[3] = {start[0], start[1], start[2]};
    int64_t segments = rng.integer(5, 10);
    for (int64_t s = 0; s < segments; ++s) {
      double d[3];
      rng.direction(d);
      double length = rng.uniform(10.0, 40.0);
      for (double t = 0; t < length; t += 0.5) {
        vol.ball(p, r, label);
        points.push_back({p[0], p[1], p[2]});
        vol.step(p, d, 0.5);
      }
      vol.ball(p, r, label);
    }
  };
  double start[3];
  for (int a = 0; a < 3; ++a) start[a] = rng.uniform(0.0, double(vol.dim[a] - 1));
  polyline(start);
  int64_t branches = rng.integer(1, 3);
  for (int64_t b = 0; b < branches && !points.empty(); ++b) {
    auto q = points[size_t(rng.integer(0, int64_t(points.size()) - 1))];
    double from[3] = {q[0], q[1], q[2]};
    polyline(from);
  }
}

// A tree label grown under a voxel budget.
void tree(Volume &vol, Rng &rng, uint32_t label, int64_t budget) {
  int64_t begin = vol.painted;
  double root[3];
  for (int a = 0; a < 3; ++a) root[a] = rng.uniform(0.0, double(vol.dim[a] - 1));
  if (rng.uniform() < 0.1) vol.ball(root, rng.uniform(15.0, 30.0), label);
  struct Tip { double p[3], d[3]; };
  std::vector<Tip> tips;
  Tip first;
  std::memcpy(first.p, root, sizeof(root));
  rng.direction(first.d);
  tips.push_back(first);
  int64_t steps = 0, max_steps = std::max<int64_t>(2000, budget / 2);
  while (!tips.empty() && vol.painted - begin < budget && steps < max_steps) {
    Tip t = tips.front();
    tips.erase(tips.begin());
    double r = rng.uniform(2.0, 10.0);
    int64_t walk = rng.integer(20, 200);
    for (int64_t s = 0; s < walk && vol.painted - begin < budget; ++s, ++steps) {
      vol.ball(t.p, r, label);
      double n = 0;
      for (int a = 0; a < 3; ++a) { t.d[a] += 0.35 * rng.normal(); n += t.d[a] * t.d[a]; }
      n = std::sqrt(std::max(n, 1e-12));
      for (int a = 0; a < 3; ++a) t.d[a] /= n;
      vol.step(t.p, t.d, 1.0);
    }
    tips.push_back(t);  // the walk continues from its end
    if (rng.uniform() < 0.1) {
      Tip b = t;
      rng.direction(b.d);
      tips.push_back(b);
    }
  }
}
// Cell seeds bucketed on a grid; nearest() searches growing shells of buckets until no closer seed
// can remain (exact nearest, ties to the lower seed index).
struct Cells {
  double w[3];              // anisotropy
  int64_t dim[3], nb[3];    // volume and bucket grid
  double bsize[3];          // bucket edge in voxels
  std::vector<std::array<double, 3>> seeds;
  std::vector<std::vector<int>> buckets;
  int64_t bucket_of(int a, double v) const { return std::min<int64_t>(nb[a] - 1, std::max<int64_t>(0, int64_t(v / bsize[a]))); }
  void build() {
    const double per = double(dim[0]) * dim[1] * dim[2] / std::max<size_t>(seeds.size(), 1);
    for (int a = 0; a < 3; ++a) {  // buckets about one cell across in physical units
      const double phys = std::cbrt(per * w[0] * w[1] * w[2]);
      bsize[a] = std::max(1.0, phys / w[a]);
      nb[a] = std::max<int64_t>(1, int64_t(std::ceil(double(dim[a]) / bsize[a])));
    }
    buckets.assign(size_t(nb[0] * nb[1] * nb[2]), {});
    for (size_t i = 0; i < seeds.size(); ++i) {
      const auto &s = seeds[i];
      buckets[size_t(bucket_of(0, s[0]) + nb[0] * (bucket_of(1, s[1]) + nb[1] * bucket_of(2, s[2])))].push_back(int(i));
    }
  }
  int nearest(int64_t x, int64_t y, int64_t z) const {
    const double p[3] = {double(x), double(y), double(z)};
    const int64_t c[3] = {bucket_of(0, p[0]), bucket_of(1, p[1]), bucket_of(2, p[2])};
    int best = -1;
    double bd = 1e300;
    for (int64_t ring = 0;; ++ring) {
      for (int64_t k = c[2] - ring; k <= c[2] + ring; ++k)
        for (int64_t j = c[1] - ring; j <= c[1] + ring; ++j)
          for (int64_t i = c[0] - ring; i <= c[0] + ring; ++i) {
            if (std::max({std::abs(i - c[0]), std::abs(j - c[1]), std::abs(k - c[2])}) != ring) continue;
            if (i < 0 || j < 0 || k < 0 || i >= nb[0] || j >= nb[1] || k >= nb[2]) continue;
            for (int s : buckets[size_t(i + nb[0] * (j + nb[1] * k))]) {
              double d = 0;
              for (int a = 0; a < 3; ++a) {
                const double e = w[a] * (seeds[size_t(s)][a] - p[a]);
                d += e * e;
              }
              if (d < bd || (d == bd && s < best)) { bd = d; best = s; }
            }
          }
      // every bucket outside the searched shells is at least this far (physical units)
      double clear = 1e300;
      for (int a = 0; a < 3; ++a) {
        const double lo = double(c[a] - ring) * bsize[a], hi = double(c[a] + ring + 1) * bsize[a];
        clear = std::min({clear, w[a] * (p[a] - lo), w[a] * (hi - p[a])});
      }
      const bool all = c[0] - ring <= 0 && c[1] - ring <= 0 && c[2] - ring <= 0 && c[0] + ring >= nb[0] - 1 &&
                       c[1] + ring >= nb[1] - 1 && c[2] + ring >= nb[2] - 1;
      if (all || (best >= 0 && bd <= clear * clear)) return best;
    }
  }
  // Inside cell s and not on its boundary (every 6-neighbour inside the volume is in cell s too).
  bool interior(int s, int64_t x, int64_t y, int64_t z) const {
    if (nearest(x, y, z) != s) return false;
    const int64_t q[6][3] = {{x - 1, y, z}, {x + 1, y, z}, {x, y - 1, z}, {x, y + 1, z}, {x, y, z - 1}, {x, y, z + 1}};
    for (const auto &n : q) {
      if (n[0] < 0 || n[1] < 0 || n[2] < 0 || n[0] >= dim[0] || n[1] >= dim[1] || n[2] >= dim[2]) continue;
      if (nearest(n[0], n[1], n[2]) != s) return false;
    }
    return true;
  }
};

// The tubes of cell s: balls painted every half voxel on the cell's interior voxels only.
int64_t cell_tubes(const Cells &cells, int s, uint64_t seed, uint64_t *out) {
  Rng rng(seed);
  const int64_t *dim = cells.dim;
  int64_t painted = 0;
  std::unordered_set<int64_t> outside;  // voxels this cell found not to own (balls overlap a lot)
  const int64_t tubes = rng.integer(1, 4);
  for (int64_t t = 1; t <= tubes; ++t) {
    const uint64_t id = (uint64_t(s + 1) << 32) | uint64_t(t);
    const double r = rng.uniform(2.0, 6.0), r2 = r * r;
    double p[3];
    for (int a = 0; a < 3; ++a)
      p[a] = std::min(std::max(cells.seeds[size_t(s)][a] + rng.uniform(-0.25, 0.25) * cells.bsize[a], 0.0), double(dim[a] - 1));
    const int64_t segments = rng.integer(2, 4);
    for (int64_t g = 0; g < segments; ++g) {
      double d[3];
      rng.direction(d);
      const double length = rng.uniform(10.0, 30.0);
      for (double h = 0; h <= length; h += 0.5) {
        int64_t lo[3], hi[3];
        bool empty = false;
        for (int a = 0; a < 3; ++a) {
          lo[a] = std::max<int64_t>(0, int64_t(std::ceil(p[a] - r)));
          hi[a] = std::min<int64_t>(dim[a] - 1, int64_t(std::floor(p[a] + r)));
          empty |= lo[a] > hi[a];
        }
        for (int64_t z = lo[2]; !empty && z <= hi[2]; ++z)
          for (int64_t y = lo[1]; y <= hi[1]; ++y)
            for (int64_t x = lo[0]; x <= hi[0]; ++x) {
              const double dx = x - p[0], dy = y - p[1], dz = z - p[2];
              if (dx * dx + dy * dy + dz * dz > r2) continue;
              const int64_t i = x + dim[0] * (y + dim[1] * z);
              uint64_t &v = out[size_t(i)];
              if (v != 0 || outside.count(i)) continue;
              if (cells.interior(s, x, y, z)) { v = id; ++painted; }
              else outside.insert(i);
            }
        for (int a = 0; a < 3; ++a) {  // step, reflecting at the walls
          p[a] += 0.5 * d[a];
          const double top = double(dim[a] - 1);
          if (p[a] < 0) { p[a] = -p[a]; d[a] = -d[a]; }
          if (p[a] > top) { p[a] = 2 * top - p[a]; d[a] = -d[a]; }
          p[a] = std::min(std::max(p[a], 0.0), top);
        }
      }
    }
  }
  return painted;
}
}  // namespace

extern "C" {
// The stress volume (see "cells" above) into a zeroed-by-this-call Fortran-ordered uint64 array.
// Returns the number of cells; *painted receives the painted voxel count.
int64_t brook_synth_cells(uint64_t seed, int64_t sx, int64_t sy, int64_t sz, double mean_cells, double ax, double ay,
                          double az, int threads, uint64_t *out, int64_t *painted) {
  std::memset(out, 0, size_t(sx * sy * sz) * sizeof(uint64_t));
  Rng rng(seed);
  // Poisson(mean) by the normal approximation (mean in the thousands), at least one cell.
  const int64_t count = std::max<int64_t>(1, int64_t(std::llround(mean_cells + std::sqrt(mean_cells) * rng.normal())));
  Cells cells;
  cells.w[0] = ax; cells.w[1] = ay; cells.w[2] = az;
  cells.dim[0] = sx; cells.dim[1] = sy; cells.dim[2] = sz;
  for (int64_t i = 0; i < count; ++i)
    cells.seeds.push_back({rng.uniform(0.0, double(sx - 1)), rng.uniform(0.0, double(sy - 1)), rng.uniform(0.0, double(sz - 1))});
  cells.build();
  threads = std::max(1, threads);
  std::vector<int64_t> done(size_t(threads), 0);
  std::vector<std::thread> pool;
  for (int w = 0; w < threads; ++w)
    pool.emplace_back([&, w]() {
      for (int64_t s = w; s < count; s += threads) done[size_t(w)] += cell_tubes(cells, int(s), seed * 1000003ull + uint64_t(s), out);
    });
  for (auto &t : pool) t.join();
  int64_t total = 0;
  for (int64_t v : done) total += v;
  *painted = total;
  return count;
}

// kind 0: tubes, kind 1: trees (budget = occupancy * voxels / labels). Returns painted voxels.
int64_t brook_synth_volume(int kind, uint64_t seed, int64_t sx, int64_t sy, int64_t sz, int labels, double occupancy,
                           uint32_t *out) {
  Volume vol{{sx, sy, sz}, out};
  std::memset(out, 0, size_t(sx * sy * sz) * sizeof(uint32_t));
  Rng rng(seed);
  int64_t budget = int64_t(occupancy * double(sx * sy * sz) / std::max(labels, 1));
  for (int l = 1; l <= labels; ++l) {
    if (kind == 0) tube(vol, rng, uint32_t(l));
    else tree(vol, rng, uint32_t(l), budget);
  }
  return vol.painted;
}
}
