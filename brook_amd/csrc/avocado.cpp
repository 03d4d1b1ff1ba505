// Avocado protection (fix_avocados; avocado.cpp:83-172 of the reference, which restates Kimimaro
// 5.8.1's engage_avocado_protection): a label whose DBF exceeds detection / 2.5 somewhere is a
// candidate; its box mask (faces' 2-D holes filled) gives the voxel of largest DBF, and six rays
// from there decide whether the label is the pit of a surrounding fruit label; a pit merges into
// its fruit, and every candidate's box then has its holes filled with the owner. Up to 20 passes,
// the DBF refreshed after each pass that changed a label. Host work, as in the reference: the
// candidate order is the iteration order of a CPython set of the ids.
#include "avocado.hpp"

#include <algorithm>
#include <cmath>
#include <map>

namespace brook {
namespace {

// A CPython 3.12 set of non-negative integers (hash(k) == k): open addressing over a power-of-two
// table, 9 linear probes then perturbed jumps, deleted slots kept as dummies, growth to 4x (2x
// above 50000) of the used count when 3/5 of the table is filled. Only the iteration order
// matters here; it depends on these rules.
class PySet {
 public:
  bool contains(size_t key) const {
    const Slot &s = slots_[lookup(key)];
    return s.state == kLive && s.key == key;
  }
  void add(size_t key) {
    size_t at = lookup(key, true);
    Slot &s = slots_[at];
    if (s.state == kLive) return;
    const bool fresh = s.state == kEmpty;
    s.key = key;
    s.state = kLive;
    ++used_;
    if (!fresh) return;
    ++fill_;
    if (fill_ * 5 >= mask() * 3) rebuild(used_ > 50000 ? used_ * 2 : used_ * 4);
  }
  void discard(size_t key) {
    Slot &s = slots_[lookup(key)];
    if (s.state == kLive && s.key == key) {
      s.state = kDummy;
      --used_;
    }
  }
  void difference_update(const PySet &other) {
    if (this == &other) { *this = PySet(); return; }
    for (const Slot &s : other.slots_)
      if (s.state == kLive) discard(s.key);
    if (fill_ - used_ > mask() / 4) rebuild(used_ > 50000 ? used_ * 2 : used_ * 4);
  }
  void update(const PySet &other) {
    if (this == &other || other.used_ == 0) return;
    if ((fill_ + other.used_) * 5 >= mask() * 3) rebuild((used_ + other.used_) * 2);
    if (fill_ == 0 && slots_.size() == other.slots_.size() && other.fill_ == other.used_) {
      *this = other;
      return;
    }
    if (fill_ == 0) {
      for (const Slot &s : other.slots_)
        if (s.state == kLive) slots_[free_slot(s.key)] = s;
      fill_ = used_ = other.used_;
      return;
    }
    for (const Slot &s : other.slots_)
      if (s.state == kLive) add(s.key);
  }
  size_t size() const { return used_; }
  std::vector<size_t> items() const {
    std::vector<size_t> out;
    for (const Slot &s : slots_)
      if (s.state == kLive) out.push_back(s.key);
    return out;
  }

 private:
  enum : uint8_t { kEmpty = 0, kLive = 1, kDummy = 2 };
  struct Slot {
    size_t key = 0;
    uint8_t state = kEmpty;
  };
  std::vector<Slot> slots_ = std::vector<Slot>(8);
  size_t used_ = 0, fill_ = 0;

  size_t mask() const { return slots_.size() - 1; }
  // The slot holding `key`, else the empty slot ending its probe sequence (for an insertion: the
  // last dummy met on the way, when there is one).
  size_t lookup(size_t key, bool insert = false) const {
    const size_t m = mask();
    size_t i = key & m, perturb = key, dummy = SIZE_MAX;
    for (;;) {
      const size_t run = i + 9 <= m ? 9 : 0;
      for (size_t j = 0; j <= run; ++j) {
        const Slot &s = slots_[i + j];
        if (s.state == kEmpty) return insert && dummy != SIZE_MAX ? dummy : i + j;
        if (s.state == kLive && s.key == key) return i + j;
        if (s.state == kDummy) dummy = i + j;
      }
      perturb >>= 5;
      i = (i * 5 + 1 + perturb) & m;
    }
  }
  size_t free_slot(size_t key) const {  // the first empty slot of key's probe sequence
    const size_t m = mask();
    size_t i = key & m, perturb = key;
    for (;;) {
      const size_t run = i + 9 <= m ? 9 : 0;
      for (size_t j = 0; j <= run; ++j)
        if (slots_[i + j].state == kEmpty) return i + j;
      perturb >>= 5;
      i = (i * 5 + 1 + perturb) & m;
    }
  }
  void rebuild(size_t minused) {
    size_t size = 8;
    while (size <= minused) size <<= 1;
    if (size == 8 && slots_.size() == 8 && fill_ == used_) return;  // nothing to purge
    std::vector<Slot> old(size);
    old.swap(slots_);
    fill_ = used_;
    for (const Slot &s : old)
      if (s.state == kLive) slots_[free_slot(s.key)] = s;
  }
};

using P3 = std::array<int64_t, 3>;

struct Grid {
  int64_t s[3];
  size_t at(const P3 &p) const { return size_t(p[0] + s[0] * (p[1] + s[1] * p[2])); }
  size_t size() const { return size_t(s[0] * s[1] * s[2]); }
  P3 point(size_t i) const {
    const int64_t v = int64_t(i);
    return P3{v % s[0], (v / s[0]) % s[1], v / (s[0] * s[1])};
  }
};

// Background fill of a mask (1 = object): what no face-connected background path joins to the
// border becomes object. In 3-D the seeds follow fill_voids 2.1.2: on each face, scanned as rows
// with the background state carried from one row to the next, only the first voxel of each
// background run is seeded; in 2-D every border pixel of the background is.
void fill_background(std::vector<uint8_t> &mask, const Grid &g, int dims) {
  std::vector<size_t> queue;
  if (dims == 3) {
    const int axes[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
    for (const auto &ax : axes)
      for (int side = 0; side < 2; ++side) {
        const int a = ax[0], b = ax[1], c = ax[2];
        bool background = false;
        for (int64_t y = 0; y < g.s[b]; ++y)
          for (int64_t x = 0; x < g.s[a]; ++x) {
            P3 p{};
            p[a] = x;
            p[b] = y;
            p[c] = side ? g.s[c] - 1 : 0;
            const size_t i = g.at(p);
            if (!mask[i] && !background) queue.push_back(i);
            background = !mask[i];
          }
      }
  } else {
    for (int64_t y = 0; y < g.s[1]; ++y)
      for (int64_t x = 0; x < g.s[0]; ++x)
        if (x == 0 || y == 0 || x == g.s[0] - 1 || y == g.s[1] - 1) {
          const size_t i = g.at(P3{x, y, 0});
          if (!mask[i]) queue.push_back(i);
        }
  }
  for (size_t i : queue) mask[i] = 2;
  for (size_t head = 0; head < queue.size(); ++head) {
    const P3 p = g.point(queue[head]);
    for (int a = 0; a < dims; ++a)
      for (int step : {-1, 1}) {
        P3 q = p;
        q[a] += step;
        if (q[a] < 0 || q[a] >= g.s[a]) continue;
        const size_t i = g.at(q);
        if (mask[i] == 0) {
          mask[i] = 2;
          queue.push_back(i);
        }
      }
  }
  for (uint8_t &v : mask) v = v == 2 ? 0 : 1;
}

// The 2-D holes of each of the six faces of the box mask become object.
void fill_faces(std::vector<uint8_t> &mask, const Grid &g) {
  const int axes[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
  for (const auto &ax : axes)
    for (int side = 0; side < 2; ++side) {
      const int a = ax[0], b = ax[1], c = ax[2];
      const Grid pg{{g.s[a], g.s[b], 1}};
      std::vector<uint8_t> plane(pg.size());
      auto at = [&](int64_t x, int64_t y) {
        P3 p{};
        p[a] = x;
        p[b] = y;
        p[c] = side ? g.s[c] - 1 : 0;
        return g.at(p);
      };
      for (int64_t y = 0; y < g.s[b]; ++y)
        for (int64_t x = 0; x < g.s[a]; ++x) plane[size_t(x + g.s[a] * y)] = mask[at(x, y)];
      fill_background(plane, pg, 2);
      for (int64_t y = 0; y < g.s[b]; ++y)
        for (int64_t x = 0; x < g.s[a]; ++x) mask[at(x, y)] = plane[size_t(x + g.s[a] * y)];
    }
}

// The six rays from `centre` (a positive ray runs to the last voxel, a negative one stops before
// coordinate 0): each ends at background or at the first other label it meets (a hit). With at
// least three hits and all but at most one (none when exactly three) on one label, that label is
// the fruit; else the centre's label stays itself.
std::pair<uint32_t, uint32_t> fruit_of(const std::vector<uint32_t> &labels, const Grid &g, const P3 &centre) {
  const uint32_t pit = labels[g.at(centre)];
  std::map<uint32_t, int> hits;
  int rays = 0;
  for (int a = 0; a < 3; ++a)
    for (int dir : {1, -1}) {
      for (P3 p = centre; dir > 0 ? p[a] < g.s[a] : p[a] > 0; p[a] += dir) {
        const uint32_t v = labels[g.at(p)];
        if (!v) break;
        if (v != pit) {
          ++hits[v];
          ++rays;
          break;
        }
      }
    }
  if (rays < 3) return {pit, pit};
  uint32_t best = pit;
  int most = 0;
  for (const auto &h : hits)
    if (h.second > most) {
      best = h.first;
      most = h.second;
    }
  if (rays - most > (rays > 3 ? 1 : 0)) return {pit, pit};
  return {pit, best};
}

}  // namespace

int avocado_correct(std::vector<uint32_t> &labels, std::vector<float> &distance, Dims d, size_t table, double detection,
                    const std::function<void(const std::vector<uint32_t> &, std::vector<float> &)> &refresh,
                    const std::vector<int64_t> &keys) {
  const Grid g{{d.sx, d.sy, d.sz}};
  // the sets hold keys; labels and boxes are indexed by compact id
  auto key = [&](size_t id) { return keys.empty() ? id : size_t(keys[id]); };
  auto id_of = [&](size_t k) {
    return keys.empty() ? k : size_t(std::lower_bound(keys.begin(), keys.end(), int64_t(k)) - keys.begin());
  };
  PySet unchanged;
  int passes = 0;
  for (; passes < 20; ++passes) {
    const float cutoff = float(detection / 2.5);
    std::vector<uint8_t> present(table, 0);
    for (size_t i = 0; i < labels.size(); ++i) present[distance[i] > cutoff ? labels[i] : 0] = 1;
    PySet candidates;
    for (size_t id = 0; id < table; ++id)
      if (present[id]) candidates.add(key(id));
    candidates.difference_update(unchanged);
    candidates.discard(0);
    if (!candidates.size()) break;
    struct Span {
      P3 lo{INT64_MAX, INT64_MAX, INT64_MAX}, hi{-1, -1, -1};
    };
    std::vector<Span> boxes(table);  // of the labels as the pass starts
    for (size_t i = 0; i < labels.size(); ++i) {
      if (!labels[i]) continue;
      const P3 p = g.point(i);
      Span &b = boxes[labels[i]];
      for (int a = 0; a < 3; ++a) {
        b.lo[a] = std::min(b.lo[a], p[a]);
        b.hi[a] = std::max(b.hi[a], p[a]);
      }
    }
    PySet unchanged_now, changed;
    for (size_t candidate : candidates.items()) {
      const size_t label = id_of(candidate);
      const Span &b = boxes[label];
      const Grid box{{b.hi[0] - b.lo[0] + 1, b.hi[1] - b.lo[1] + 1, b.hi[2] - b.lo[2] + 1}};
      auto global = [&](size_t i) {
        const P3 p = box.point(i);
        return g.at(P3{p[0] + b.lo[0], p[1] + b.lo[1], p[2] + b.lo[2]});
      };
      std::vector<uint8_t> mask(box.size());
      for (size_t i = 0; i < mask.size(); ++i) mask[i] = labels[global(i)] == uint32_t(label);
      fill_faces(mask, box);
      // numpy argmax of mask * distance over the box in x-fastest order: the first maximum, a NaN
      // (0 * inf) winning for good
      size_t best = 0;
      float top = float(mask[0]) * distance[global(0)];
      for (size_t i = 0; i < mask.size(); ++i) {
        const float v = float(mask[i]) * distance[global(i)];
        if (!(v <= top) && !std::isnan(top)) {
          top = v;
          best = i;
        }
      }
      P3 centre = box.point(best);
      for (int a = 0; a < 3; ++a) centre[a] += b.lo[a];
      const auto po = fruit_of(labels, g, centre);
      const uint32_t pit = po.first, owner = po.second;
      if (pit == owner && !changed.contains(key(pit))) {
        unchanged_now.add(key(pit));
      } else {
        unchanged_now.discard(key(pit));
        unchanged_now.discard(key(owner));
        changed.add(key(pit));
        changed.add(key(owner));
        for (size_t i = 0; i < mask.size(); ++i) mask[i] |= labels[global(i)] == owner;
      }
      fill_background(mask, box, 3);
      for (size_t i = 0; i < mask.size(); ++i)
        if (mask[i]) labels[global(i)] = owner;
    }
    unchanged.update(unchanged_now);
    if (!changed.size()) {
      ++passes;
      break;
    }
    refresh(labels, distance);
  }
  return passes;
}

size_t avocado_renumber(std::vector<uint32_t> &labels, const std::vector<uint32_t> &original,
                        const std::vector<int64_t> &through, size_t table, std::vector<int64_t> &mapping,
                        std::vector<int64_t> &roots) {
  std::vector<uint32_t> fresh(table, 0);
  mapping.assign(1, 0);
  roots.clear();
  uint32_t previous = UINT32_MAX;
  for (size_t i = 0; i < labels.size(); ++i) {
    uint32_t &id = fresh[labels[i]];
    if (labels[i] && !id) {
      id = uint32_t(roots.size() + 1);
      roots.push_back(int64_t(i));
      mapping.push_back(0);
    }
    labels[i] = id;
    if (id != previous) {  // a run start: the input label under it (fastremap's get_mapping)
      if (id) mapping[id] = original[i] ? through[original[i]] : 0;
      previous = id;
    }
  }
  return roots.size();
}

}  // namespace brook
