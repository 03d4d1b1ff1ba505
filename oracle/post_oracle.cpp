// CPU restatement of the reference's canonical postprocess of packed skeletons (test
// infrastructure: only the tests, the smoke check and the bench's CPU leg load it).
//
// Follows cpp/src/post.cu (graph 45-67, select 70-117, prune_unused 433-442, remove_dust 115-137 +
// 443-475, remove_loops 476-491, remove_ticks 492-520, join_components 521-648, postprocess 649-651)
// and cpp/src/post_kernels.cuh (remove_loops_k 22-199, join_scan 210-254, join_fuse 258-306,
// remove_ticks_k 309-405). Those kernels follow Kimimaro 5.8.1's remove_loops,
// join_close_components and remove_ticks with the reference's own tie rules; each rule is stated
// where it is used. This restatement keeps the state the rules depend on (row order, neighbour
// lists in insertion order, cluster positions) and none of the reference's incremental shortcuts
// (leaf stripping, heap, cell grid): the cycle search walks the whole component, the shortest
// super-edge and the join candidates are found by full scans.
//
// Pins (tests/test_post.py): the reference's own checks (cpp_smoke.cpp:44-48: merge then
// postprocess(0, 0) returns the skeleton unchanged; test_package.py:220), hand-worked answers
// for each rule (a ring, a lollipop, two and three branch points on a cycle, the far-branch snip,
// a T with a short and a long arm, two chains within and beyond reach, blocked pairs, cluster
// order of the greedy fusion), NumPy's pairwise summation against numpy.sum on float32, and
// properties (a forest after remove_loops, trees untouched, idempotence, dust by cable length).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace {

using i64 = int64_t;

struct Packed {
  std::vector<i64> labels, voff{0}, eoff{0};
  std::vector<float> vertices;  // x, y, z per vertex
  std::vector<uint32_t> edges;  // local pairs
  std::vector<float> radii;
  size_t n() const { return radii.size(); }
  size_t m() const { return edges.size() / 2; }
};

// The group (skeleton index) of element i: the last group whose offset is <= i (post.cu:28-32).
int group_of(const std::vector<i64> &off, i64 i) {
  auto it = std::upper_bound(off.begin(), off.end() - 1, i);
  return int(it - off.begin()) - 1;
}

struct Graph {
  int n = 0, m = 0, k = 0;
  std::vector<int> vlabel, elabel, root, degree;
  std::vector<std::array<int, 2>> e;  // global vertex ids
};

// post.cu:45-67: validation, global edges, degrees (a self-loop counts twice), components as
// roots = least vertex index.
Graph graph(const Packed &p) {
  const size_t k = p.labels.size();
  bool ok = p.voff.size() == k + 1 && p.eoff.size() == k + 1 && p.voff.front() == 0 && p.eoff.front() == 0 &&
            std::is_sorted(p.voff.begin(), p.voff.end()) && std::is_sorted(p.eoff.begin(), p.eoff.end()) &&
            size_t(p.voff.back()) == p.radii.size() && p.vertices.size() == 3 * p.radii.size() &&
            size_t(p.eoff.back()) * 2 == p.edges.size();
  if (!ok) throw std::invalid_argument("invalid packed graph");
  if (p.voff.back() > INT32_MAX || p.eoff.back() > INT32_MAX || k > size_t(INT32_MAX))
    throw std::invalid_argument("packed graph exceeds int32 work capacity");
  Graph g;
  g.n = int(p.voff.back());
  g.m = int(p.eoff.back());
  g.k = int(k);
  g.vlabel.resize(g.n);
  g.elabel.resize(g.m);
  g.degree.assign(g.n, 0);
  g.root.resize(g.n);
  for (int i = 0; i < g.n; ++i) { g.vlabel[i] = group_of(p.voff, i); g.root[i] = i; }
  std::vector<int> &up = g.root;
  auto find = [&](int a) { while (up[a] != a) a = up[a] = up[up[a]]; return a; };
  g.e.resize(g.m);
  for (int i = 0; i < g.m; ++i) {
    const int l = group_of(p.eoff, i);
    g.elabel[i] = l;
    const i64 size = p.voff[l + 1] - p.voff[l];
    const uint32_t a = p.edges[2 * size_t(i)], b = p.edges[2 * size_t(i) + 1];
    if (a >= size || b >= size) throw std::invalid_argument("packed edge is out of bounds");
    const int ga = int(p.voff[l] + a), gb = int(p.voff[l] + b);
    g.e[i] = {ga, gb};
    g.degree[ga]++;
    g.degree[gb]++;
    int ra = find(ga), rb = find(gb);
    if (ra != rb) up[std::max(ra, rb)] = std::min(ra, rb);
  }
  for (int i = 0; i < g.n; ++i) g.root[i] = find(i);
  return g;
}

// post.cu:86-109: the kept vertices in order, the edges with both ends kept in order, per-group
// counts; labels and empty groups stay.
Packed select(const Packed &p, const std::vector<uint8_t> &keep, const Graph &g) {
  if (!g.n) return p;
  Packed out;
  out.labels = p.labels;
  std::vector<i64> rank(g.n, -1);
  std::vector<i64> vc(g.k, 0), ec(g.k, 0);
  for (int i = 0; i < g.n; ++i)
    if (keep[i]) {
      rank[i] = i64(out.radii.size());
      out.vertices.insert(out.vertices.end(), p.vertices.begin() + 3 * i64(i), p.vertices.begin() + 3 * i64(i) + 3);
      out.radii.push_back(p.radii[i]);
      vc[g.vlabel[i]]++;
    }
  for (int l = 0; l < g.k; ++l) out.voff.push_back(out.voff.back() + vc[l]);
  for (int i = 0; i < g.m; ++i) {
    const auto e = g.e[i];
    if (!keep[e[0]] || !keep[e[1]]) continue;
    const i64 base = out.voff[g.elabel[i]];
    out.edges.push_back(uint32_t(rank[e[0]] - base));
    out.edges.push_back(uint32_t(rank[e[1]] - base));
    ec[g.elabel[i]]++;
  }
  for (int l = 0; l < g.k; ++l) out.eoff.push_back(out.eoff.back() + ec[l]);
  return out;
}

Packed prune_unused(const Packed &p) {  // post.cu:433-442
  Graph g = graph(p);
  if (!g.n) return p;
  if (*std::min_element(g.degree.begin(), g.degree.end()) > 0) return p;
  std::vector<uint8_t> keep(g.n);
  for (int i = 0; i < g.n; ++i) keep[i] = g.degree[i] != 0;
  return select(p, keep, g);
}

// NumPy's pairwise float32 summation (numpy/_core/src/umath/loops_utils.h.src, pairwise_sum):
// fewer than 8 values in order from -0; up to 128 values in eight interleaved partial sums
// combined as ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7)) and the tail added in order; above
// 128 the halves (the first rounded down to a multiple of 8) summed separately (post.cu:140-149).
float pairwise(const float *v, size_t n) {
  if (n < 8) {
    float s = -0.0f;
    for (size_t i = 0; i < n; ++i) s += v[i];
    return s;
  }
  if (n <= 128) {
    float r[8];
    for (int j = 0; j < 8; ++j) r[j] = v[j];
    size_t i = 8;
    for (; i + 8 <= n; i += 8)
      for (int j = 0; j < 8; ++j) r[j] += v[i + j];
    float s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) s += v[i];
    return s;
  }
  size_t half = n / 2;
  half -= half % 8;
  return pairwise(v, half) + pairwise(v + half, n - half);
}

float edge_length32(const Packed &p, int a, int b) {  // post.cu:463-464
  float x = p.vertices[3 * i64(b)] - p.vertices[3 * i64(a)], y = p.vertices[3 * i64(b) + 1] - p.vertices[3 * i64(a) + 1],
        z = p.vertices[3 * i64(b) + 2] - p.vertices[3 * i64(a) + 2];
  x = x * x;
  y = y * y;
  z = z * z;
  return std::sqrt((x + y) + z);
}

// remove_dust (post.cu:115-137, 443-475): a component stays when its cable length exceeds the
// threshold. The length is the double sum of double edge lengths; within count * 2^-23 *
// threshold of the threshold it is NumPy's pairwise float32 sum of float32 lengths over the
// component's edges in edge order instead.
Packed remove_dust(const Packed &p, double threshold) {
  if (threshold == 0 || p.n() == 0) return p;
  Graph g = graph(p);
  std::vector<double> total(g.n, 0.0);
  std::vector<int> count(g.n, 0);
  std::vector<std::vector<int>> edges_of(g.n);
  for (int i = 0; i < g.m; ++i) {
    const auto e = g.e[i];
    double x = double(p.vertices[3 * i64(e[1])] - p.vertices[3 * i64(e[0])]);
    double y = double(p.vertices[3 * i64(e[1]) + 1] - p.vertices[3 * i64(e[0]) + 1]);
    double z = double(p.vertices[3 * i64(e[1]) + 2] - p.vertices[3 * i64(e[0]) + 2]);
    const int r = g.root[e[0]];
    total[r] += std::sqrt((x * x + y * y) + z * z);
    count[r]++;
    edges_of[r].push_back(i);
  }
  std::vector<uint8_t> keep_root(g.n, 0);
  for (int r = 0; r < g.n; ++r) {
    keep_root[r] = total[r] > threshold;
    if (count[r] > 0 && std::fabs(total[r] - threshold) <= count[r] * (1.0 / 8388608.0) * threshold) {
      std::vector<float> lengths;
      for (int i : edges_of[r]) lengths.push_back(edge_length32(p, g.e[i][0], g.e[i][1]));
      keep_root[r] = double(pairwise(lengths.data(), lengths.size())) > threshold;
    }
  }
  std::vector<uint8_t> keep(g.n);
  for (int i = 0; i < g.n; ++i) keep[i] = keep_root[g.root[i]];
  return select(p, keep, g);
}

// Sorted unique (min, max) global pairs -> per-group local edges; vertices unchanged
// (post.cu:273-292).
Packed edges_from_keys(const Packed &p, const Graph &g, std::vector<uint64_t> keys) {
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  Packed out = p;
  out.edges.clear();
  std::vector<i64> count(g.k, 0);
  for (uint64_t key : keys) {
    const int a = int(key >> 32), b = int(key & 0xffffffffu), l = g.vlabel[a];
    out.edges.push_back(uint32_t(a - p.voff[l]));
    out.edges.push_back(uint32_t(b - p.voff[l]));
    count[l]++;
  }
  out.eoff.assign(1, 0);
  for (int l = 0; l < g.k; ++l) out.eoff.push_back(out.eoff.back() + count[l]);
  return out;
}

uint64_t pair_key(i64 a, i64 b) { return (uint64_t(std::min(a, b)) << 32) | uint64_t(std::max(a, b)); }

// A component's vertices (ascending index) and edges (edge order) as local ids (post.cu:214-258).
struct Piece {
  std::vector<int> vertex, edge;  // global ids
};
std::vector<Piece> pieces_of(const Graph &g, const std::vector<uint8_t> &active, std::vector<int> &slot) {
  slot.assign(g.n, -1);
  std::vector<Piece> out;
  for (int r = 0; r < g.n; ++r)
    if (active[r]) { slot[r] = int(out.size()); out.emplace_back(); }
  for (int v = 0; v < g.n; ++v)
    if (slot[g.root[v]] >= 0) out[slot[g.root[v]]].vertex.push_back(v);
  for (int i = 0; i < g.m; ++i)
    if (slot[g.root[g.e[i][0]]] >= 0) out[slot[g.root[g.e[i][0]]]].edge.push_back(i);
  return out;
}

// ---- remove_loops (post_kernels.cuh:22-199) ----
// One component, local ids. Rows in order (dead rows stay in place, new rows are appended); each
// node lists its neighbours in insertion order, one entry per unordered pair with the pair's row
// count and its first row; cnt = live rows at a node with multiplicity.
struct LoopState {
  struct Entry { int nbr, mul, rid; };
  std::vector<std::array<int, 2>> rows;
  std::vector<uint8_t> live;
  std::vector<std::vector<Entry>> adj;
  std::vector<int> cnt;
  i64 nlive = 0;

  int find(int a, int b) const {
    for (size_t j = 0; j < adj[a].size(); ++j)
      if (adj[a][j].nbr == b) return int(j);
    return -1;
  }
  void add_row(int x, int y) {
    const int r = int(rows.size());
    rows.push_back({x, y});
    live.push_back(1);
    nlive++;
    cnt[x]++;
    cnt[y]++;
    const int f = find(x, y);
    if (f >= 0) {  // a repeated pair: both entries count it (the first match of each list)
      adj[x][f].mul++;
      adj[y][find(y, x)].mul++;
    } else {
      adj[x].push_back({y, 1, r});
      adj[y].push_back({x, 1, r});
    }
  }
};

// Returns the component's surviving rows in row order. Throws when the cycles do not resolve
// within a bound (the reference has no bound; a component that never resolves would not end).
std::vector<std::array<int, 2>> loops_one(const Packed &p, const Piece &c, const Graph &g) {
  const int nv = int(c.vertex.size());
  std::vector<int> local(g.n, -1);
  for (int j = 0; j < nv; ++j) local[c.vertex[j]] = j;
  auto P = [&](int v, int a) { return p.vertices[3 * i64(c.vertex[v]) + a]; };
  auto R = [&](int v) { return p.radii[c.vertex[v]]; };
  LoopState s;
  s.adj.resize(nv);
  s.cnt.assign(nv, 0);
  for (int i : c.edge) s.add_row(local[g.e[i][0]], local[g.e[i][1]]);
  std::vector<int> appended;  // rows appended by the previous iteration
  const i64 budget = 64 * (i64(c.edge.size()) + nv) + 64;
  for (i64 iteration = 0;; ++iteration) {
    if (iteration > budget) throw std::runtime_error("remove_loops did not converge");
    if (s.nlive == 0) break;
    size_t first = 0;
    while (!s.live[first]) first++;
    // depth-first search from the first live row's first end (Kimimaro's find_cycle)
    struct Frame { int node, parent, depth; };
    std::vector<Frame> stack{{s.rows[first][0], -1, 0}};
    std::vector<int> path;
    std::vector<uint8_t> visited(nv, 0);
    int node = -1;
    bool found = false;
    while (!stack.empty()) {
      const Frame f = stack.back();
      stack.pop_back();
      node = f.node;
      if (int(path.size()) > f.depth) path.resize(f.depth);
      path.push_back(node);
      if (visited[node]) { found = true; break; }
      visited[node] = 1;
      for (const auto &e : s.adj[node])
        if (e.nbr != f.parent) stack.push_back({e.nbr, node, f.depth + 1});
    }
    const int plen = int(path.size());
    if (!found || plen <= 1) break;
    int i0 = 0;
    while (i0 < plen - 1 && path[i0] != node) i0++;
    if (plen - i0 < 3) break;
    const std::vector<int> cy(path.begin() + i0, path.end());  // cy.front() == cy.back()
    const int cl = int(cy.size());
    std::vector<int> branch;  // cycle nodes with >= 3 live rows, ascending
    for (int j = 0; j + 1 < cl; ++j)
      if (s.cnt[cy[j]] >= 3) branch.push_back(cy[j]);
    std::sort(branch.begin(), branch.end());
    const int nb = int(branch.size());
    std::vector<uint8_t> drop(cl, 0);  // drop[j]: the cycle edge (cy[j], cy[j + 1]) goes
    std::vector<std::array<int, 2>> add;
    auto sq = [&](int v, float x, float y, float z) {
      const float dx = P(v, 0) - x, dy = P(v, 1) - y, dz = P(v, 2) - z;
      return (dx * dx + dy * dy) + dz * dz;
    };
    if (nb == 0) {
      std::fill(drop.begin(), drop.end() - 1, 1);
    } else if (nb == 1) {  // the cycle goes; the branch node joins the cycle node farthest from it
      const int bn = branch[0];
      int end = -1;
      float best = -1.0f;
      for (int j = 0; j + 1 < cl; ++j) {  // the first maximum in ascending node order
        const int v = cy[j];
        const float d = sq(v, P(bn, 0), P(bn, 1), P(bn, 2));
        if (end < 0 || d > best || (d == best && v < end)) { best = d; end = v; }
      }
      std::fill(drop.begin(), drop.end() - 1, 1);
      add.push_back({bn, end});
    } else if (nb == 2) {  // the side with fewer nodes between the two branch nodes stays
      const int L = cl - 1;
      int p0 = -1, p1 = -1;
      for (int j = 0; j < L; ++j)
        if (s.cnt[cy[j + 1]] >= 3) (p0 < 0 ? p0 : p1) = j;
      const bool inner = double(p1 - p0) < double(L) / 2.0;
      for (int j = 0; j + 1 < cl; ++j) {
        // edge j joins path position j - 1 and j of cy[1..]; edge 0 closes the cycle
        const bool on_inner = j - 1 >= 0 && j - 1 >= p0 && j <= p1;
        if (!(inner ? on_inner : !on_inner)) drop[j] = 1;
      }
    } else {  // three or more: the node nearest the branch nodes' centroid becomes a hub
      float cx = 0.0f, cyy = 0.0f, cz = 0.0f;
      for (int v : branch) { cx += P(v, 0); cyy += P(v, 1); cz += P(v, 2); }
      cx /= float(nb);
      cyy /= float(nb);
      cz /= float(nb);
      int hub = -1;
      float best = 0.0f;
      for (int v = 0; v < nv; ++v) {  // every node of the component, the first minimum
        const float d = sq(v, cx, cyy, cz);
        if (hub < 0 || d < best) { best = d; hub = v; }
      }
      float far = 0.0f;
      for (int i = 0; i < nb; ++i) {
        const float d = sq(branch[i], P(hub, 0), P(hub, 1), P(hub, 2));
        if (i == 0 || d > far) far = d;
      }
      if (std::sqrt(far) > R(hub)) {
        drop[0] = 1;  // a snip of the cycle's first edge only
      } else {
        std::fill(drop.begin(), drop.end() - 1, 1);
        for (int v : branch)
          if (v != hub) add.push_back({v, hub});
      }
    }
    // last iteration's appended rows become (lo, hi); then every dropped cycle edge goes with all
    // of its rows (a single row by its first-row record; several: the live rows equal to (lo, hi))
    for (int r : appended)
      if (s.rows[r][0] > s.rows[r][1]) std::swap(s.rows[r][0], s.rows[r][1]);
    for (int j = 0; j + 1 < cl; ++j) {
      if (!drop[j]) continue;
      const int a = cy[j], b = cy[j + 1];
      const int e = s.find(a, b);
      if (e < 0) continue;
      const int mul = s.adj[a][e].mul;
      if (mul == 1) {
        s.live[s.adj[a][e].rid] = 0;
      } else {
        const int lo = std::min(a, b), hi = std::max(a, b);
        for (size_t r = first; r < s.rows.size(); ++r)
          if (s.live[r] && s.rows[r][0] == lo && s.rows[r][1] == hi) s.live[r] = 0;
      }
      s.nlive -= mul;
      s.cnt[a] -= mul;
      s.cnt[b] -= mul;
      const int back = s.find(b, a);
      s.adj[a].erase(s.adj[a].begin() + e);
      s.adj[b].erase(s.adj[b].begin() + back);
    }
    appended.clear();
    for (const auto &r : add) {
      appended.push_back(int(s.rows.size()));
      s.add_row(r[0], r[1]);
    }
  }
  std::vector<std::array<int, 2>> out;
  for (size_t r = 0; r < s.rows.size(); ++r)
    if (s.live[r]) out.push_back(s.rows[r]);
  return out;
}

// remove_loops (post.cu:476-491): components with a cycle (edges >= vertices) lose their cycles;
// every edge of the result becomes a sorted unique (min, max) pair without self-loops; a cyclic
// component left with no edge loses its vertices.
Packed remove_loops(const Packed &p) {
  if (!p.n() || !p.m()) return p;
  Graph g = graph(p);
  std::vector<int> nv(g.n, 0), ne(g.n, 0);
  for (int v = 0; v < g.n; ++v) nv[g.root[v]]++;
  for (int i = 0; i < g.m; ++i) ne[g.root[g.e[i][0]]]++;
  std::vector<uint8_t> cyclic(g.n, 0);
  bool any = false;
  for (int r = 0; r < g.n; ++r) any |= (cyclic[r] = ne[r] > 0 && ne[r] - nv[r] + 1 > 0);
  if (!any) return p;
  std::vector<int> slot;
  std::vector<Piece> pieces = pieces_of(g, cyclic, slot);
  std::vector<uint64_t> keys;
  for (int i = 0; i < g.m; ++i)
    if (slot[g.root[g.e[i][0]]] < 0 && g.e[i][0] != g.e[i][1]) keys.push_back(pair_key(g.e[i][0], g.e[i][1]));
  std::vector<uint8_t> emptied(pieces.size(), 0);
  for (size_t c = 0; c < pieces.size(); ++c) {
    const auto rows = loops_one(p, pieces[c], g);
    emptied[c] = rows.empty();
    for (const auto &r : rows) {
      const int a = pieces[c].vertex[r[0]], b = pieces[c].vertex[r[1]];
      if (a != b) keys.push_back(pair_key(a, b));
    }
  }
  Packed out = edges_from_keys(p, g, keys);
  if (std::find(emptied.begin(), emptied.end(), 1) == emptied.end()) return out;
  std::vector<uint8_t> keep(g.n);
  for (int v = 0; v < g.n; ++v) keep[v] = slot[g.root[v]] < 0 || !emptied[slot[g.root[v]]];
  return select(out, keep, graph(out));
}

float P_(const Packed &p, const Piece &c, int v, int a) { return p.vertices[3 * i64(c.vertex[v]) + a]; }

// ---- remove_ticks (post_kernels.cuh:309-405) ----
// One tree with a branch point, local ids; neighbours of a node: the rows that start at it, then
// the rows that end at it, each in row order. Super-edges join critical nodes (degree 1 or >= 3)
// along chains of the depth-first search from the first leaf; a chain's length is the float sum of
// float edge lengths in search order.
std::vector<uint8_t> ticks_one(const Packed &p, const Piece &c, const Graph &g, double threshold) {
  const int nv = int(c.vertex.size()), ne = int(c.edge.size());
  std::vector<int> local(g.n, -1);
  for (int j = 0; j < nv; ++j) local[c.vertex[j]] = j;
  std::vector<std::vector<std::array<int, 2>>> nbr(nv);  // (node, row)
  for (int r = 0; r < ne; ++r) nbr[local[g.e[c.edge[r]][0]]].push_back({local[g.e[c.edge[r]][1]], r});
  for (int r = 0; r < ne; ++r) nbr[local[g.e[c.edge[r]][1]]].push_back({local[g.e[c.edge[r]][0]], r});
  auto deg = [&](int v) { return int(nbr[v].size()); };
  std::vector<uint8_t> keep(ne, 1);
  int start = -1;
  std::vector<int> bc(nv);
  for (int v = 0; v < nv; ++v) {
    bc[v] = deg(v) >= 3 ? deg(v) : 0;
    if (start < 0 && deg(v) == 1) start = v;
  }
  if (start < 0) return keep;
  struct Super { int u, v; double len; bool alive, eligible; std::vector<int> chains; };
  std::vector<Super> se;  // indexed by chain id; alive ones are super-edges
  std::vector<int> chain_of(ne, -1);
  std::vector<std::vector<int>> slots(nv);  // per node: its super-edge ends, most recent first
  std::vector<int> owner;                   // per slot (2 * id + end): its super-edge
  struct Frame { int node, parent, root, chain; float dist; };
  std::vector<Frame> stack{{start, -1, start, -1, 0.0f}};
  while (!stack.empty()) {
    Frame f = stack.back();
    stack.pop_back();
    const int d = deg(f.node);
    if ((d == 1 || d >= 3) && f.node != f.root) {
      Super &s = se[f.chain];
      s.u = f.root;
      s.v = f.node;
      s.len = double(f.dist);
      s.alive = true;
      s.eligible = deg(f.root) == 1 || d == 1;
      slots[f.root].insert(slots[f.root].begin(), 2 * f.chain);
      slots[f.node].insert(slots[f.node].begin(), 2 * f.chain + 1);
      owner[2 * f.chain] = owner[2 * f.chain + 1] = f.chain;
      f.dist = 0.0f;
      f.root = f.node;
    }
    for (const auto &q : nbr[f.node]) {
      if (q[0] == f.parent) continue;
      float dx = P_(p, c, f.node, 0) - P_(p, c, q[0], 0), dy = P_(p, c, f.node, 1) - P_(p, c, q[0], 1),
            dz = P_(p, c, f.node, 2) - P_(p, c, q[0], 2);
      dx *= dx;
      dy *= dy;
      dz *= dz;
      int ch = f.chain;
      if (f.node == f.root) {  // a chain leaves a critical node
        ch = int(se.size());
        se.push_back(Super{-1, -1, 0.0, false, false, {ch}});
        owner.resize(2 * se.size(), -1);
      }
      chain_of[q[1]] = ch;
      stack.push_back({q[0], f.node, f.root, ch, f.dist + std::sqrt((dx + dy) + dz)});
    }
  }
  int nalive = 0;
  for (const auto &s : se) nalive += s.alive;
  std::vector<uint8_t> dead_chain(se.size(), 0);
  while (nalive > 1) {
    int best = -1;  // the shortest eligible super-edge, ties by (lower end, higher end)
    for (int s = 0; s < int(se.size()); ++s) {
      if (!se[s].alive || !se[s].eligible) continue;
      if (best < 0) { best = s; continue; }
      const auto key = [&](int t) { return std::make_tuple(se[t].len, std::min(se[t].u, se[t].v), std::max(se[t].u, se[t].v)); };
      if (key(s) < key(best)) best = s;
    }
    if (best < 0) break;
    const int e1 = se[best].u, e2 = se[best].v;
    if (bc[e1] == 1 && bc[e2] == 1) break;
    if (se[best].len >= threshold) break;
    for (int ch : se[best].chains) dead_chain[ch] = 1;
    se[best].alive = false;
    nalive--;
    bc[e1]--;
    bc[e2]--;
    for (int v : {e1, e2}) {  // a branch node left with two super-edges joins them
      if (bc[v] != 2) continue;
      int a = -1, b = -1;
      for (int slot : slots[v]) {
        const int o = owner[slot];
        if (!se[o].alive || (se[o].u != v && se[o].v != v) || o == a) continue;
        if (a < 0) a = o;
        else { b = o; break; }
      }
      if (a < 0 || b < 0) { bc[v] = 0; continue; }
      const int ea = se[a].u == v ? se[a].v : se[a].u, eb = se[b].u == v ? se[b].v : se[b].u;
      se[a].len = (0.0 + se[a].len) + se[b].len;
      se[a].u = ea;
      se[a].v = eb;
      se[a].eligible = true;
      se[a].chains.insert(se[a].chains.end(), se[b].chains.begin(), se[b].chains.end());
      se[b].alive = false;
      nalive--;
      for (int slot : slots[eb])
        if (owner[slot] == b) owner[slot] = a;
      bc[v] = 0;
    }
  }
  for (int r = 0; r < ne; ++r) keep[r] = !dead_chain[chain_of[r]];
  return keep;
}

// remove_ticks (post.cu:492-520): trees with a branch point lose their short terminal chains; the
// edges keep their order, every vertex stays.
Packed remove_ticks(const Packed &p, double threshold) {
  if (threshold == 0 || !p.n() || !p.m()) return p;
  Graph g = graph(p);
  std::vector<int> nv(g.n, 0), ne(g.n, 0);
  std::vector<uint8_t> branch(g.n, 0), active(g.n, 0);
  for (int v = 0; v < g.n; ++v) { nv[g.root[v]]++; branch[g.root[v]] |= g.degree[v] >= 3; }
  for (int i = 0; i < g.m; ++i) ne[g.root[g.e[i][0]]]++;
  bool any = false;
  for (int r = 0; r < g.n; ++r) any |= (active[r] = branch[r] && ne[r] == nv[r] - 1);
  if (!any) return p;
  std::vector<int> slot;
  std::vector<Piece> pieces = pieces_of(g, active, slot);
  std::vector<uint8_t> live(g.m, 1);
  for (const Piece &c : pieces) {
    const auto keep = ticks_one(p, c, g, threshold);
    for (size_t r = 0; r < keep.size(); ++r) live[c.edge[r]] = keep[r];
  }
  Packed out = p;
  out.edges.clear();
  out.eoff.assign(1, 0);
  std::vector<i64> count(g.k, 0);
  for (int i = 0; i < g.m; ++i)
    if (live[i]) {
      out.edges.push_back(p.edges[2 * size_t(i)]);
      out.edges.push_back(p.edges[2 * size_t(i) + 1]);
      count[g.elabel[i]]++;
    }
  for (int l = 0; l < g.k; ++l) out.eoff.push_back(out.eoff.back() + count[l]);
  return out;
}

// ---- join_components (post.cu:521-648, post_kernels.cuh:210-306) ----
// Per label with several components: reach = max(2 * largest positive radius, 0). For every pair
// of components (A < B by root) the nearest vertex pair within reach + 1e-6 (double distance of
// the float coordinates; ties by the lower then the higher vertex index) is a candidate. Greedy
// fusion: among live candidates whose distance is at most the float sum of their vertices' radii,
// the one of least float distance, ties by the clusters' list positions (a fused cluster moves to
// the front, the newest first); stop when none is left or its distance exceeds the reach. After
// a fusion the candidates of the fused cluster to each other cluster keep the nearer one.
Packed join(const Packed &input) {
  Packed p = prune_unused(input);
  if (!p.n() || !p.m()) return p;
  Graph g = graph(p);
  std::vector<int> ncomp(g.k, 0);
  std::vector<float> rmax(g.k, 0.0f);
  for (int v = 0; v < g.n; ++v) {
    const int l = g.vlabel[v];
    if (g.root[v] == v) ncomp[l]++;
    if (p.radii[v] > 0) rmax[l] = std::max(rmax[l], p.radii[v]);
  }
  std::vector<uint64_t> keys;
  for (int i = 0; i < g.m; ++i) keys.push_back(pair_key(g.e[i][0], g.e[i][1]));
  bool joined = false;
  for (int l = 0; l < g.k; ++l) {
    if (ncomp[l] <= 1) continue;
    const double reach = std::max(2.0 * double(rmax[l]), 0.0), bound = reach + 0.000001;
    std::vector<int> roots;
    for (i64 v = p.voff[l]; v < p.voff[l + 1]; ++v)
      if (g.root[v] == v) roots.push_back(int(v));
    const int nc = int(roots.size());
    auto index_of = [&](int root) { return int(std::lower_bound(roots.begin(), roots.end(), root) - roots.begin()); };
    struct Pair { int x, y; double d; i64 va, vb; bool alive; };
    std::vector<Pair> pairs;  // by (A, B)
    {
      std::vector<std::vector<Pair>> best(nc, std::vector<Pair>(nc, Pair{-1, -1, 0, -1, -1, false}));
      float b32 = float(bound);
      b32 = b32 * b32 * 1.0001f + 1e-30f;
      for (i64 i = p.voff[l]; i < p.voff[l + 1]; ++i)
        for (i64 j = p.voff[l]; j < p.voff[l + 1]; ++j) {
          if (g.root[j] <= g.root[i]) continue;
          const float *a = &p.vertices[3 * i], *b = &p.vertices[3 * j];
          const float fx = b[0] - a[0], fy = b[1] - a[1], fz = b[2] - a[2];
          if ((fx * fx + fy * fy) + fz * fz > b32) continue;
          const double ex = double(b[0]) - double(a[0]), ey = double(b[1]) - double(a[1]), ez = double(b[2]) - double(a[2]);
          const double d = std::sqrt((ex * ex + ey * ey) + ez * ez);
          if (!(d < bound)) continue;
          Pair &q = best[index_of(g.root[i])][index_of(g.root[j])];
          const i64 lo = std::min(i, j), hi = std::max(i, j);
          if (!q.alive || d < q.d || (d == q.d && (lo < std::min(q.va, q.vb) || (lo == std::min(q.va, q.vb) && hi < std::max(q.va, q.vb)))))
            q = Pair{index_of(g.root[i]), index_of(g.root[j]), d, i, j, true};
        }
      for (int x = 0; x < nc; ++x)
        for (int y = 0; y < nc; ++y)
          if (best[x][y].alive) pairs.push_back(best[x][y]);
    }
    std::vector<int> pos(nc);
    for (int c = 0; c < nc; ++c) pos[c] = c;
    int fusions = 0;
    for (;;) {
      int pick = -1;
      float bd = 0.0f;
      int bi = 0, bj = 0;
      for (int q = 0; q < int(pairs.size()); ++q) {
        const Pair &c = pairs[q];
        if (!c.alive) continue;
        if (c.d > double(float(p.radii[c.va] + p.radii[c.vb]))) continue;  // blocked by the radii
        const float d32 = float(c.d);
        int ia = pos[c.x], ib = pos[c.y];
        if (ia > ib) std::swap(ia, ib);
        if (pick < 0 || d32 < bd || (d32 == bd && (ia < bi || (ia == bi && ib < bj)))) { pick = q; bd = d32; bi = ia; bj = ib; }
      }
      if (pick < 0 || double(bd) > reach) break;
      const int X = pairs[pick].x, Y = pairs[pick].y;
      keys.push_back(pair_key(pairs[pick].va, pairs[pick].vb));
      joined = true;
      pairs[pick].alive = false;
      fusions++;
      std::vector<int> kept(nc, -1);  // per other cluster: its candidate with X
      for (int q = 0; q < int(pairs.size()); ++q) {
        Pair &c = pairs[q];
        if (!c.alive) continue;
        if (c.x == Y) c.x = X;
        if (c.y == Y) c.y = X;
        if (c.x == c.y) { c.alive = false; continue; }
        if (c.x != X && c.y != X) continue;
        const int other = c.x == X ? c.y : c.x;
        if (kept[other] < 0) { kept[other] = q; continue; }
        Pair &o = pairs[kept[other]];
        const auto order = [](const Pair &t) { return std::make_tuple(t.d, std::min(t.va, t.vb), std::max(t.va, t.vb)); };
        if (order(c) < order(o)) { o.alive = false; kept[other] = q; }
        else c.alive = false;
      }
      pos[X] = -fusions;
    }
  }
  if (!joined) return p;
  return edges_from_keys(p, g, keys);
}

Packed postprocess(const Packed &p, double dust, double ticks) {  // post.cu:649-651
  Packed out = remove_dust(p, dust);
  out = remove_loops(out);
  out = join(out);
  out = remove_ticks(out, ticks);
  return prune_unused(out);
}

thread_local std::string last_error;

}  // namespace

extern "C" {

// op: 0 postprocess(a = dust, b = ticks), 1 remove_dust(a), 2 remove_loops, 3 join, 4 remove_ticks(a),
// 5 prune_unused. Returns a result handle, or null with the message in oracle_post_error().
void *oracle_post(int op, size_t k, const int64_t *labels, const int64_t *voff, const int64_t *eoff, const float *vertices,
                  const uint32_t *edges, const float *radii, double a, double b) {
  try {
    Packed p;
    p.labels.assign(labels, labels + k);
    p.voff.assign(voff, voff + k + 1);
    p.eoff.assign(eoff, eoff + k + 1);
    const size_t n = size_t(std::max<int64_t>(p.voff.back(), 0)), m = size_t(std::max<int64_t>(p.eoff.back(), 0));
    p.vertices.assign(vertices, vertices + 3 * n);
    p.radii.assign(radii, radii + n);
    p.edges.assign(edges, edges + 2 * m);
    Packed out;
    switch (op) {
      case 0: out = postprocess(p, a, b); break;
      case 1: out = remove_dust(p, a); break;
      case 2: out = remove_loops(p); break;
      case 3: out = join(p); break;
      case 4: out = remove_ticks(p, a); break;
      case 5: out = prune_unused(p); break;
      default: throw std::invalid_argument("unknown postprocess step");
    }
    return new Packed(std::move(out));
  } catch (const std::exception &e) {
    last_error = e.what();
    return nullptr;
  }
}
const char *oracle_post_error() { return last_error.c_str(); }
void oracle_post_sizes(void *h, int64_t *sizes) {
  const Packed &p = *static_cast<Packed *>(h);
  sizes[0] = int64_t(p.labels.size());
  sizes[1] = int64_t(p.n());
  sizes[2] = int64_t(p.m());
}
void oracle_post_copy(void *h, int64_t *labels, int64_t *voff, int64_t *eoff, float *vertices, uint32_t *edges, float *radii) {
  const Packed &p = *static_cast<Packed *>(h);
  std::copy(p.labels.begin(), p.labels.end(), labels);
  std::copy(p.voff.begin(), p.voff.end(), voff);
  std::copy(p.eoff.begin(), p.eoff.end(), eoff);
  std::copy(p.vertices.begin(), p.vertices.end(), vertices);
  std::copy(p.edges.begin(), p.edges.end(), edges);
  std::copy(p.radii.begin(), p.radii.end(), radii);
}
void oracle_post_free(void *h) { delete static_cast<Packed *>(h); }

// Per vertex: its component's root (the least vertex index of the component).
int oracle_post_components(size_t k, const int64_t *labels, const int64_t *voff, const int64_t *eoff, const float *vertices,
                           const uint32_t *edges, const float *radii, int64_t *out) {
  try {
    Packed p;
    p.labels.assign(labels, labels + k);
    p.voff.assign(voff, voff + k + 1);
    p.eoff.assign(eoff, eoff + k + 1);
    const size_t n = size_t(std::max<int64_t>(p.voff.back(), 0)), m = size_t(std::max<int64_t>(p.eoff.back(), 0));
    p.vertices.assign(vertices, vertices + 3 * n);
    p.radii.assign(radii, radii + n);
    p.edges.assign(edges, edges + 2 * m);
    Graph g = graph(p);
    for (int i = 0; i < g.n; ++i) out[i] = g.root[i];
    return 0;
  } catch (const std::exception &e) {
    last_error = e.what();
    return -1;
  }
}

float oracle_pairwise_sum(const float *values, size_t n) { return pairwise(values, n); }

}  // extern "C"
