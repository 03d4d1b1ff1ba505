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
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
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
}  // namespace

extern "C" {
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
