// TEST INFRASTRUCTURE ONLY: the oracle's restatement of the avocado correction (the reference's
// avocado.cpp:14-172 and integer_set.hpp, restating Kimimaro 5.8.1 engage_avocado_protection).
// Written apart from the product's (brook_amd/csrc/avocado.cpp) so that a slip in either shows.
#pragma once
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <vector>

namespace oracle_avocado {

using i64 = int64_t;

// CPython 3.12 setobject.c for keys with hash == key >= 0: entries carry (key, hash) with hash -1
// for a dummy and key < 0 for an empty entry.
struct CSet {
  struct Entry { i64 key = -1, hash = 0; };
  std::vector<Entry> table = std::vector<Entry>(8);
  size_t fill = 0, used = 0;
  size_t mask() const { return table.size() - 1; }
  static bool empty(const Entry &e) { return e.key < 0 && e.hash == 0; }
  static bool dummy(const Entry &e) { return e.hash == -1; }
  static bool active(const Entry &e) { return e.key >= 0 && e.hash != -1; }

  // set_lookkey: the entry of key, or the empty entry ending the search
  size_t lookkey(i64 key) const {
    size_t m = mask(), i = size_t(key) & m, perturb = size_t(key);
    while (true) {
      size_t probes = (i + 9 <= m) ? 9 : 0;
      for (size_t e = i; e <= i + probes; ++e) {
        if (empty(table[e])) return e;
        if (active(table[e]) && table[e].key == key) return e;
      }
      perturb >>= 5;
      i = (i * 5 + 1 + perturb) & m;
    }
  }
  static void insert_clean(std::vector<Entry> &t, i64 key) {
    size_t m = t.size() - 1, i = size_t(key) & m, perturb = size_t(key);
    while (true) {
      size_t probes = (i + 9 <= m) ? 9 : 0;
      for (size_t e = i; e <= i + probes; ++e)
        if (empty(t[e])) { t[e].key = key; t[e].hash = key; return; }
      perturb >>= 5;
      i = (i * 5 + 1 + perturb) & m;
    }
  }
  void table_resize(size_t minused) {
    size_t newsize = 8;
    while (newsize <= minused) newsize <<= 1;
    if (newsize == 8 && table.size() == 8 && fill == used) return;
    std::vector<Entry> old = table;
    table.assign(newsize, Entry());
    fill = used;
    for (const Entry &e : old) if (active(e)) insert_clean(table, e.key);
  }
  // set_add_entry
  void add(i64 key) {
    size_t m = mask(), i = size_t(key) & m, perturb = size_t(key);
    size_t freeslot = SIZE_MAX;
    while (true) {
      size_t probes = (i + 9 <= m) ? 9 : 0;
      for (size_t e = i; e <= i + probes; ++e) {
        Entry &x = table[e];
        if (empty(x)) {
          if (freeslot != SIZE_MAX) { table[freeslot] = {key, key}; used++; return; }
          x = {key, key};
          fill++; used++;
          if (fill * 5 < m * 3) return;
          table_resize(used > 50000 ? used * 2 : used * 4);
          return;
        }
        if (active(x) && x.key == key) return;
        if (dummy(x)) freeslot = e;
      }
      perturb >>= 5;
      i = (i * 5 + 1 + perturb) & m;
    }
  }
  void discard(i64 key) {
    Entry &x = table[lookkey(key)];
    if (active(x) && x.key == key) { x.key = -2; x.hash = -1; used--; }
  }
  bool contains(i64 key) const { const Entry &x = table[lookkey(key)]; return active(x) && x.key == key; }
  // set_difference_update with a set
  void difference_update(const CSet &o) {
    if (this == &o) { *this = CSet(); return; }
    for (const Entry &e : o.table) if (active(e)) discard(e.key);
    if ((fill - used) <= mask() / 4) return;
    table_resize(used > 50000 ? used * 2 : used * 4);
  }
  // set_merge
  void update(const CSet &o) {
    if (this == &o || o.used == 0) return;
    if ((fill + o.used) * 5 >= mask() * 3) table_resize((used + o.used) * 2);
    if (fill == 0 && mask() == o.mask() && o.fill == o.used) { table = o.table; fill = o.fill; used = o.used; return; }
    if (fill == 0) {
      fill = used = o.used;
      for (const Entry &e : o.table) if (active(e)) insert_clean(table, e.key);
      return;
    }
    for (const Entry &e : o.table) if (active(e)) add(e.key);
  }
  std::vector<i64> iterate() const {
    std::vector<i64> out;
    for (const Entry &e : table) if (active(e)) out.push_back(e.key);
    return out;
  }
};

struct Vol {
  i64 nx, ny, nz;
  i64 idx(i64 x, i64 y, i64 z) const { return x + nx * (y + ny * z); }
};

// fill_voids on a box mask (values 0/1): 3-D with fill_voids 2.1.2's seeds (the first voxel of each
// background run on each face, the run state carried across rows), 2-D with every border pixel.
inline void fill_mask(std::vector<uint8_t> &m, i64 sx, i64 sy, i64 sz, bool three) {
  std::vector<i64> q;
  auto at = [&](i64 x, i64 y, i64 z) { return x + sx * (y + sy * z); };
  i64 ext[3] = {sx, sy, sz};
  if (three) {
    int order[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
    for (auto &o : order)
      for (int side = 0; side < 2; ++side) {
        bool bg = false;
        for (i64 v = 0; v < ext[o[1]]; ++v)
          for (i64 u = 0; u < ext[o[0]]; ++u) {
            i64 c[3];
            c[o[0]] = u; c[o[1]] = v; c[o[2]] = side ? ext[o[2]] - 1 : 0;
            i64 i = at(c[0], c[1], c[2]);
            if (m[i] == 0 && !bg) q.push_back(i);
            bg = m[i] == 0;
          }
      }
  } else {
    for (i64 y = 0; y < sy; ++y)
      for (i64 x = 0; x < sx; ++x)
        if ((x == 0 || y == 0 || x == sx - 1 || y == sy - 1) && m[at(x, y, 0)] == 0) q.push_back(at(x, y, 0));
  }
  for (i64 i : q) m[size_t(i)] = 2;
  for (size_t h = 0; h < q.size(); ++h) {
    i64 i = q[h], x = i % sx, y = (i / sx) % sy, z = i / (sx * sy);
    i64 c[3] = {x, y, z};
    for (int a = 0; a < (three ? 3 : 2); ++a)
      for (int dlt = -1; dlt <= 1; dlt += 2) {
        i64 n[3] = {c[0], c[1], c[2]};
        n[a] += dlt;
        if (n[a] < 0 || n[a] >= ext[a]) continue;
        i64 j = at(n[0], n[1], n[2]);
        if (m[size_t(j)] == 0) { m[size_t(j)] = 2; q.push_back(j); }
      }
  }
  for (auto &v : m) v = (v == 2) ? 0 : 1;
}

// The correction; returns the passes. refresh(labels, distance) recomputes the field. keys: with a
// voxel graph, Kimimaro's id of each compact id (the rank of its graph component among all of them,
// background ones included; ascending), which its candidate and change sets hold (avocado.cpp:85-91);
// empty: the ids themselves.
inline int correct(std::vector<uint32_t> &lab, std::vector<float> &dist, const Vol &V, size_t table, double threshold,
                   const std::function<void(const std::vector<uint32_t> &, std::vector<float> &)> &refresh,
                   const std::vector<i64> &keys = {}) {
  auto key = [&](uint32_t id) { return keys.empty() ? i64(id) : keys[id]; };
  auto id_of = [&](i64 k) {
    return keys.empty() ? uint32_t(k) : uint32_t(std::lower_bound(keys.begin(), keys.end(), k) - keys.begin());
  };
  CSet unchanged;
  int passes = 0;
  while (passes < 20) {
    const float cut = float(threshold / 2.5);
    std::vector<char> seen(table, 0);
    for (size_t i = 0; i < lab.size(); ++i) seen[dist[i] > cut ? lab[i] : 0] = 1;
    CSet cand;
    for (size_t id = 0; id < table; ++id) if (seen[id]) cand.add(key(uint32_t(id)));
    cand.difference_update(unchanged);
    cand.discard(0);
    if (cand.used == 0) break;
    std::vector<std::array<i64, 6>> bb(table, std::array<i64, 6>{INT64_MAX, INT64_MAX, INT64_MAX, -1, -1, -1});
    for (i64 z = 0; z < V.nz; ++z)
      for (i64 y = 0; y < V.ny; ++y)
        for (i64 x = 0; x < V.nx; ++x) {
          uint32_t l = lab[size_t(V.idx(x, y, z))];
          if (!l) continue;
          auto &b = bb[l];
          b[0] = std::min(b[0], x); b[1] = std::min(b[1], y); b[2] = std::min(b[2], z);
          b[3] = std::max(b[3], x); b[4] = std::max(b[4], y); b[5] = std::max(b[5], z);
        }
    CSet keep_now, changed;
    for (i64 K : cand.iterate()) {
      const uint32_t L = id_of(K);
      const auto &b = bb[size_t(L)];
      const i64 sx = b[3] - b[0] + 1, sy = b[4] - b[1] + 1, sz = b[5] - b[2] + 1;
      auto g = [&](i64 x, i64 y, i64 z) { return size_t(V.idx(x + b[0], y + b[1], z + b[2])); };
      std::vector<uint8_t> m(size_t(sx * sy * sz));
      for (i64 z = 0; z < sz; ++z)
        for (i64 y = 0; y < sy; ++y)
          for (i64 x = 0; x < sx; ++x) m[size_t(x + sx * (y + sy * z))] = lab[g(x, y, z)] == uint32_t(L);
      // the six faces' 2-D holes
      i64 ext[3] = {sx, sy, sz};
      int order[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
      for (auto &o : order)
        for (int side = 0; side < 2; ++side) {
          std::vector<uint8_t> plane(size_t(ext[o[0]] * ext[o[1]]));
          auto cell = [&](i64 u, i64 v) {
            i64 c[3];
            c[o[0]] = u; c[o[1]] = v; c[o[2]] = side ? ext[o[2]] - 1 : 0;
            return size_t(c[0] + sx * (c[1] + sy * c[2]));
          };
          for (i64 v = 0; v < ext[o[1]]; ++v)
            for (i64 u = 0; u < ext[o[0]]; ++u) plane[size_t(u + ext[o[0]] * v)] = m[cell(u, v)];
          fill_mask(plane, ext[o[0]], ext[o[1]], 1, false);
          for (i64 v = 0; v < ext[o[1]]; ++v)
            for (i64 u = 0; u < ext[o[0]]; ++u) m[cell(u, v)] = plane[size_t(u + ext[o[0]] * v)];
        }
      // first maximum of mask * distance (a NaN sticks)
      i64 arg = 0;
      float mx = float(m[0]) * dist[g(0, 0, 0)];
      for (i64 z = 0; z < sz; ++z)
        for (i64 y = 0; y < sy; ++y)
          for (i64 x = 0; x < sx; ++x) {
            const i64 i = x + sx * (y + sy * z);
            const float val = float(m[size_t(i)]) * dist[g(x, y, z)];
            if (!(val <= mx) && !std::isnan(mx)) { mx = val; arg = i; }
          }
      const i64 cx = arg % sx + b[0], cy = (arg / sx) % sy + b[1], cz = arg / (sx * sy) + b[2];
      // the rays
      const uint32_t pit = lab[size_t(V.idx(cx, cy, cz))];
      std::map<uint32_t, int> hit;
      int rays = 0;
      i64 lim[3] = {V.nx, V.ny, V.nz};
      for (int a = 0; a < 3; ++a) {
        for (i64 t = (a == 0 ? cx : a == 1 ? cy : cz); t < lim[a]; ++t) {
          i64 c[3] = {cx, cy, cz};
          c[a] = t;
          uint32_t w = lab[size_t(V.idx(c[0], c[1], c[2]))];
          if (w == 0) break;
          if (w != pit) { hit[w]++; rays++; break; }
        }
        for (i64 t = (a == 0 ? cx : a == 1 ? cy : cz); t > 0; --t) {
          i64 c[3] = {cx, cy, cz};
          c[a] = t;
          uint32_t w = lab[size_t(V.idx(c[0], c[1], c[2]))];
          if (w == 0) break;
          if (w != pit) { hit[w]++; rays++; break; }
        }
      }
      uint32_t owner = pit;
      if (rays >= 3) {
        uint32_t top = pit;
        int most = 0;
        for (auto &h : hit) if (h.second > most) { most = h.second; top = h.first; }
        if (!(rays - most > (rays > 3 ? 1 : 0))) owner = top;
      }
      if (pit == owner && !changed.contains(key(pit))) {
        keep_now.add(key(pit));
      } else {
        keep_now.discard(key(pit));
        keep_now.discard(key(owner));
        changed.add(key(pit));
        changed.add(key(owner));
        for (i64 z = 0; z < sz; ++z)
          for (i64 y = 0; y < sy; ++y)
            for (i64 x = 0; x < sx; ++x)
              if (lab[g(x, y, z)] == owner) m[size_t(x + sx * (y + sy * z))] = 1;
      }
      fill_mask(m, sx, sy, sz, true);
      for (i64 z = 0; z < sz; ++z)
        for (i64 y = 0; y < sy; ++y)
          for (i64 x = 0; x < sx; ++x)
            if (m[size_t(x + sx * (y + sy * z))]) lab[g(x, y, z)] = owner;
    }
    unchanged.update(keep_now);
    passes++;
    if (changed.used == 0) break;
    refresh(lab, dist);
  }
  return passes;
}

}  // namespace oracle_avocado
