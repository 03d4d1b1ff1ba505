"""CPU checks of the oracle: golden vectors of the reference's suite, known answers, and
independent brute-force restatements chosen so a dropped term or a wrong index fails."""
import itertools
import json
import os
import random
from collections import deque

import numpy as np
import pytest

import oracle
from tests.golden import shapes

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden_digests.json")))


@pytest.mark.parametrize("name", [c[0] for c in shapes.golden_cases()])
def test_golden_digests(name):
    # Kimimaro 5.8.1 results held by the reference's tests (test_package.py:2079-2106).
    case = {c[0]: c for c in shapes.golden_cases()}[name]
    options = dict(shapes.GOLDEN_OPTIONS, **case[2])
    for layout in (np.ascontiguousarray, np.asfortranarray):
        got = shapes.summary_from_packed(oracle.skeletonize(layout(case[1]), **options))
        assert got == {int(k): v for k, v in GOLDEN[name].items()}


def test_ccl_known_answer():
    # c_smoke.c:27-34: mapping and roots of an 8-voxel line.
    labels = np.array([1, 1, 0, 0, 0, 0, 2, 2], np.uint8)[:, None, None]
    cc, mapping, roots = oracle.ccl(labels)
    assert list(mapping) == [0, 1, 2] and list(roots) == [0, 6]
    assert cc.ravel(order="F").tolist() == [1, 1, 0, 0, 0, 0, 2, 2]


def _bfs_components(labels):
    """Brute-force 26-connected labelling by BFS in Fortran flat order (ids by first voxel)."""
    lab = np.asfortranarray(labels)
    sx, sy, sz = lab.shape
    out = np.zeros(lab.shape, np.int64, order="F")
    nid = 0
    for z in range(sz):
        for y in range(sy):
            for x in range(sx):
                if lab[x, y, z] == 0 or out[x, y, z]:
                    continue
                nid += 1
                out[x, y, z] = nid
                q = deque([(x, y, z)])
                while q:
                    a, b, c = q.popleft()
                    for da, db, dc in itertools.product((-1, 0, 1), repeat=3):
                        p = (a + da, b + db, c + dc)
                        if 0 <= p[0] < sx and 0 <= p[1] < sy and 0 <= p[2] < sz and not out[p] and lab[p] == lab[x, y, z]:
                            out[p] = nid
                            q.append(p)
    return out


@pytest.mark.parametrize("seed", range(4))
def test_ccl_matches_bfs(seed):
    rng = np.random.default_rng(seed)
    labels = rng.integers(0, 4, size=(9, 7, 6)) * (rng.random((9, 7, 6)) < 0.6)
    cc, mapping, roots = oracle.ccl(labels)
    ref = _bfs_components(labels)
    assert np.array_equal(cc.astype(np.int64), ref)
    flat = labels.ravel(order="F")
    assert [flat[r] for r in roots] == list(mapping[1:])
    assert all(np.flatnonzero(ref.ravel(order="F") == i + 1)[0] == r for i, r in enumerate(roots))


def test_edt_known_answer():
    # c_smoke.c:53-56: {2,1,0,0,0,0,1,2} for the 8-voxel line, no black border.
    labels = np.array([1, 1, 0, 0, 0, 0, 2, 2], np.uint8)[:, None, None]
    d = oracle.edt(labels, (1, 1, 1), False)
    assert d.ravel().tolist() == [2, 1, 0, 0, 0, 0, 1, 2]


def _brute_edt(labels, aniso, black_border):
    lab = np.asarray(labels)
    pts = np.argwhere(np.ones(lab.shape, bool))
    vals = lab.reshape(-1)
    out = np.zeros(lab.shape)
    a = np.asarray(aniso, np.float64)
    shape = np.array(lab.shape)
    for i, p in enumerate(pts):
        v = lab[tuple(p)]
        if v == 0:
            continue
        other = pts[vals != v]
        best = np.inf if len(other) == 0 else np.min(np.sum(((other - p) * a) ** 2, axis=1))
        if black_border:  # the outside is a wall one voxel beyond each face
            for ax in range(3):
                best = min(best, ((p[ax] + 1) * a[ax]) ** 2, ((shape[ax] - p[ax]) * a[ax]) ** 2)
        out[tuple(p)] = np.sqrt(best)
    return out


@pytest.mark.parametrize("seed,bb", [(0, False), (1, True), (2, False)])
def test_edt_f64_matches_brute_force(seed, bb):
    rng = np.random.default_rng(seed)
    labels = (rng.integers(1, 4, size=(7, 6, 5)) * (rng.random((7, 6, 5)) < 0.8)).astype(np.uint32)
    aniso = (1.5, 2.0, 3.25)
    got = oracle.edt_f64(labels, aniso, bb)
    ref = _brute_edt(labels, [float(np.float32(a)) for a in aniso], bb)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("seed", range(3))
def test_edt_rule_within_tolerance_of_f64(seed):
    rng = np.random.default_rng(seed)
    labels = np.zeros((40, 36, 30), np.uint32)
    for lab in range(1, 6):
        c = rng.integers(0, 30, 3)
        r = rng.integers(3, 12)
        x, y, z = np.ogrid[:40, :36, :30]
        labels[((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r) & (labels == 0)] = lab
    aniso = (4.0, 4.0, 40.0)
    d32 = oracle.edt(labels, aniso).astype(np.float64)
    d64 = oracle.edt_f64(labels, aniso)
    fg = labels > 0
    rel = np.abs(d32[fg] - d64[fg]) / d64[fg]
    assert rel.max() <= 1e-6


def test_pointset_matches_cpython_set():
    rng = random.Random(11)
    for _ in range(100):
        pts = [tuple(rng.randint(-700, 700) for _ in range(3)) for _ in range(rng.randint(1, 400))]
        s = set()
        for p in pts:
            s.add(p)
        assert [pts[i] for i in oracle.pointset_order(pts)] == list(s)


def test_geometric_flood_matches_bellman_ford():
    rng = np.random.default_rng(3)
    labels = (rng.random((8, 7, 6)) < 0.7).astype(np.uint32) * rng.integers(1, 3, (8, 7, 6)).astype(np.uint32)
    cc, mapping, roots = oracle.ccl(labels)
    aniso = (1.0, 2.0, 3.0)
    got = oracle.geometric_flood(cc, aniso, roots)
    # Bellman-Ford to the fixpoint with float32 additions
    x, y, z = aniso
    w = [np.float32(v) for v in (x, x, y, y, z, z)]
    offs = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
    for d, pair in ((np.sqrt(x * x + y * y), [(-1, -1, 0), (-1, 1, 0), (1, -1, 0), (1, 1, 0)]),
                    (np.sqrt(y * y + z * z), [(0, -1, -1), (0, -1, 1), (0, 1, -1), (0, 1, 1)]),
                    (np.sqrt(x * x + z * z), [(-1, 0, -1), (-1, 0, 1), (1, 0, -1), (1, 0, 1)])):
        w += [np.float32(d)] * 4
        offs += pair
    for o in itertools.product((-1, 1), repeat=3):
        w.append(np.float32(np.sqrt(x * x + y * y + z * z)))
        offs.append(o)
    dist = np.full(cc.shape, np.inf, np.float32)
    for r in roots:
        dist[np.unravel_index(r, cc.shape, order="F")] = 0
    changed = True
    while changed:
        changed = False
        for p in np.argwhere(dist < np.inf):
            p = tuple(p)
            for o, wk in zip(offs, w):
                q = tuple(a + b for a, b in zip(p, o))
                if all(0 <= q[i] < cc.shape[i] for i in range(3)) and cc[q] == cc[p]:
                    c = np.float32(dist[p] + wk)
                    if c < dist[q]:
                        dist[q] = c
                        changed = True
    assert np.array_equal(got, dist)


def _vw_step(d, f):
    """brook_vw_step (common.cuh:123-127): fl(d + f), one ulp up where a positive f adds nothing."""
    c = np.float32(d + f)
    if c == d and f > 0 and d < np.inf:
        c = np.nextafter(np.float32(d), np.float32(np.inf))
    return np.float32(c)


def test_vw_parents_match_bellman_ford_and_chain_rule():
    """fix_branching=False field: the vertex-weighted flood equals a brute-force float32 Bellman-Ford
    fixpoint bit for bit (weights with ties and one-ulp steps), and every parent is the first
    same-label 26-neighbour (table order) of smallest distance, whose step gives the voxel's
    distance; chains reach a source."""
    rng = np.random.default_rng(5)
    labels = (rng.random((8, 7, 6)) < 0.75).astype(np.uint32) * rng.integers(1, 3, (8, 7, 6)).astype(np.uint32)
    cc, _, roots = oracle.ccl(labels)
    vw = rng.choice(np.float32([0.5, 1.0, 1.0, 2.0, 1e-8]), size=cc.shape).astype(np.float32)
    dist, parent = oracle.vw_parents(labels, vw, roots)
    offs = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1), (-1, -1, 0), (-1, 1, 0), (1, -1, 0),
            (1, 1, 0), (0, -1, -1), (0, -1, 1), (0, 1, -1), (0, 1, 1), (-1, 0, -1), (-1, 0, 1), (1, 0, -1), (1, 0, 1)]
    offs += list(itertools.product((-1, 1), repeat=3))
    want = np.full(cc.shape, np.inf, np.float32)
    sources = {np.unravel_index(r, cc.shape, order="F") for r in roots}
    for p in sources:
        want[p] = 0
    nbrs = lambda p: [q for q in (tuple(a + b for a, b in zip(p, o)) for o in offs)
                      if all(0 <= q[i] < cc.shape[i] for i in range(3)) and cc[q] == cc[p]]
    changed = True
    while changed:
        changed = False
        for p in np.argwhere(want < np.inf):
            p = tuple(p)
            for q in nbrs(p):
                c = _vw_step(want[p], vw[q])
                if c < want[q]:
                    want[q] = c
                    changed = True
    assert np.array_equal(dist, want)
    flat = lambda q: int(np.ravel_multi_index(q, cc.shape, order="F"))
    for v in np.argwhere(cc > 0):
        v = tuple(v)
        if v in sources:
            assert parent[v] == -1
            continue
        best = None
        for q in nbrs(v):
            if best is None or dist[q] < dist[best]:
                best = q
        assert parent[v] == flat(best)
        assert dist[v] == _vw_step(dist[best], vw[v])
        steps, cur = 0, v
        while parent[cur] >= 0:
            nxt = np.unravel_index(int(parent[cur]), cc.shape, order="F")
            assert dist[nxt] < dist[cur]
            cur, steps = nxt, steps + 1
        assert cur in sources and steps < cc.size


@pytest.mark.parametrize("seed", [1001, 1002])
def test_parent_walk_stop_rule_equals_full_walks(seed):
    """The STOP rule of the parent walks (lockstep_regular.inc:160-165) gives the same skeletons as
    Kimimaro's fix_branching=False paths, every walk running to the root."""
    from brook_amd import synth
    labels = synth.volume("tubes", (96, 96, 64), 8, seed)
    params = dict(scale=1.5, const=300, pdrf_scale=100000, pdrf_exponent=4)
    stop = oracle.skeletonize(labels, params, (16, 16, 40), 1000, fix_branching=False, threads=8)
    full = oracle.skeletonize(labels, params, (16, 16, 40), 1000, fix_branching=False, stop_emitted=False, threads=8)
    assert stop["stats"] == full["stats"] and len(stop["labels"]) > 3
    for name in ("labels", "v_off", "e_off", "vertices", "edges", "radii"):
        assert np.array_equal(stop[name], full[name]), name


def test_railroad_road_known_answer():
    # test_package.py:751-770: a 9^3 field whose z = 0 plane is rail.
    field = np.ones((9, 9, 9), np.float32, order="F")
    field[:, :, 0] = 0
    flat = lambda p: int(np.ravel_multi_index(p, field.shape, order="F"))
    road = [(4, 4, 4), (3, 3, 3), (2, 2, 2), (1, 1, 1), (1, 1, 0)]
    assert oracle.railroad(field, flat((4, 4, 4))) == [flat(p) for p in road]


D3_ORDER = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1),
            (-1, -1, 0), (-1, 1, 0), (1, -1, 0), (1, 1, 0), (0, -1, -1), (0, -1, 1), (0, 1, -1), (0, 1, 1),
            (-1, 0, -1), (-1, 0, 1), (1, 0, -1), (1, 0, 1),
            (-1, -1, -1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1), (1, 1, -1), (1, -1, 1), (-1, 1, 1), (1, 1, 1)]


def test_railroad_join_order_known_answers():
    # test_package.py:783-811: dijkstra3d joins the first rail in its neighbour order.
    shape, centre = (5, 5, 5), (2, 2, 2)
    flat = lambda p: int(np.ravel_multi_index(p, shape, order="F"))
    near = [tuple(c + d for c, d in zip(centre, o)) for o in D3_ORDER]
    for rails in [(i,) for i in range(26)] + list(itertools.combinations(range(26), 2)) + [tuple(range(26))]:
        field = np.ones(shape, np.float32, order="F")
        for i in rails:
            field[near[i]] = 0
        assert oracle.railroad(field, flat(centre)) == [flat(centre), flat(near[rails[0]])], rails


def test_target_order_and_pdrf_rules():
    # PDRF closed form (rule book A.8) and the order key (A.9) on a small two-label volume.
    labels = np.zeros((30, 20, 10), np.uint32)
    labels[2:28, 5:12, 2:8] = 1
    labels[5:25, 14:19, 3:7] = 2
    tp = dict(scale=1.5, const=3, pdrf_scale=1000, pdrf_exponent=4)
    pr = oracle.prepare(labels, tp, (2, 2, 3), dust_threshold=0)
    dbf, daf, pdrf, cc = pr["dbf"], pr["daf"], pr["pdrf"], pr["cc"]
    for lab in (1, 2):
        m = cc == lab
        dmax = np.float32(dbf[m].max())
        assert pr["dbf_max"][lab] == dmax
        M = np.float32(1.0 / np.power(np.float64(dmax), 1.01))
        inv = np.float32(1.0 / np.float64(pr["max_daf"][lab]))
        p = (np.float32(1) - dbf[m] * M)
        p = (p * p) * (p * p)
        expect = p * np.float32(1000) + daf[m] * inv
        assert np.array_equal(pdrf[m], expect.astype(np.float32))
        # the target is the lowest flat index among the DAF maxima
        flat_daf = daf.ravel(order="F")
        flat_cc = cc.ravel(order="F")
        idx = np.flatnonzero(flat_cc == lab)
        best = idx[flat_daf[idx] == flat_daf[idx].max()].min()
        assert pr["target"][lab] == best
