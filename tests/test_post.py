"""The canonical postprocess of packed skeletons (post.cu, post_kernels.cuh): remove_dust,
remove_loops, join_components, remove_ticks, prune_unused.

CPU tests pin the oracle (oracle/post_oracle.cpp): the reference's own checks (cpp_smoke.cpp:44-48:
a merged result postprocessed without thresholds is unchanged; test_package.py:220), NumPy's
pairwise float32 sum against numpy.sum, and hand-worked answers for every rule: a ring (no branch
node: the cycle and then its vertices go), a lollipop (one branch node joins the farthest cycle
node), two branch nodes (the shorter arc stays), three branch nodes (a hub near their centroid,
or a single snip when a branch node lies beyond the hub's radius), ticks on a T and on an H
(short arms go shortest first, a worn-down branch point fuses its two chains), dust by cable
length with the float32 rule inside the uncertainty band, joins within and beyond reach, blocked
by the radii, chained over three pieces and never across labels. Properties on random graphs: a
forest after remove_loops, idempotence, untouched trees. GPU tests compare brook_amd with the
oracle bit for bit on the same inputs.
"""
import numpy as np
import pytest

import oracle
from brook_amd._packed import PackedSkeletons


def packed(*skeletons):
    """(label, vertices, edges, radii) tuples -> a packed dict."""
    labels, voff, eoff, vs, es, rs = [], [0], [0], [], [], []
    for label, v, e, r in skeletons:
        v = np.asarray(v, np.float32).reshape(-1, 3)
        e = np.asarray(e, np.uint32).reshape(-1, 2)
        labels.append(label)
        vs.append(v)
        es.append(e)
        rs.append(np.broadcast_to(np.asarray(r, np.float32), (len(v),)).copy())
        voff.append(voff[-1] + len(v))
        eoff.append(eoff[-1] + len(e))
    return dict(labels=np.array(labels, np.int64), v_off=np.array(voff, np.int64), e_off=np.array(eoff, np.int64),
                vertices=np.concatenate(vs) if vs else np.zeros((0, 3), np.float32),
                edges=np.concatenate(es) if es else np.zeros((0, 2), np.uint32),
                radii=np.concatenate(rs) if rs else np.zeros(0, np.float32))


def ring(n, radius=10.0):
    t = np.arange(n) * (2 * np.pi / n)
    v = np.stack([radius * np.cos(t), radius * np.sin(t), np.zeros(n)], 1)
    e = [(min(i, (i + 1) % n), max(i, (i + 1) % n)) for i in range(n)]
    return v, e


def with_tails(v, e, nodes, length=2):
    """Radial tails of `length` vertices at the given ring nodes."""
    v, e = list(map(tuple, v)), list(e)
    for node in nodes:
        x, y, z = v[node]
        norm = np.hypot(x, y)
        prev = node
        for k in range(1, length + 1):
            v.append((x + k * x / norm, y + k * y / norm, z))
            e.append((prev, len(v) - 1))
            prev = len(v) - 1
    return np.array(v), e


def edge_set(p, k=0):
    e0, e1 = int(p["e_off"][k]), int(p["e_off"][k + 1])
    return {tuple(sorted(map(int, x))) for x in p["edges"][e0:e1]}


def same(a, b):
    for name in ("labels", "v_off", "e_off", "vertices", "edges", "radii"):
        assert np.array_equal(np.asarray(a[name]), np.asarray(b[name])), name


def chain(x0, n, step=1.0, axis=0, start=(0.0, 0.0, 0.0)):
    v = np.zeros((n, 3))
    v[:] = start
    v[:, axis] = x0 + step * np.arange(n)
    return v, [(i, i + 1) for i in range(n - 1)]


# ---- oracle pins ----

def test_pairwise_sum_is_numpys():
    rng = np.random.default_rng(3)
    for n in (0, 1, 7, 8, 9, 15, 16, 127, 128, 129, 255, 256, 1000, 4099):
        v = (rng.random(n) * 7).astype(np.float32)
        assert oracle.pairwise_sum(v) == np.sum(v), n


def test_skeleton_postprocess_without_thresholds_is_unchanged():
    """cpp_smoke.cpp:44-48: a result merged from two copies and postprocessed with dust 0 and
    ticks 0 is the result itself (trees of one component per label)."""
    from brook_amd import synth
    labels = synth.volume("tubes", (48, 40, 32), 5, 7)
    sk = oracle.skeletonize(labels, dict(scale=1.5, const=2, pdrf_scale=100000, pdrf_exponent=4), (1, 1, 1), 0)
    merged = oracle.merge_fragments([sk, sk], [(0, 0, 0), (0, 0, 0)])
    same(oracle.post("postprocess", merged, 0, 0), merged)
    same(oracle.post("postprocess", sk, 0, 0), sk)


def test_ring_without_branch_nodes_disappears():
    v, e = ring(6)
    p = packed((7, v, e, 1.0))
    out = oracle.post("remove_loops", p)
    assert out["labels"].tolist() == [7] and out["v_off"].tolist() == [0, 0] and len(out["edges"]) == 0


def test_lollipop_joins_the_farthest_cycle_node():
    v, e = with_tails(*ring(6), [0], length=3)
    out = oracle.post("remove_loops", packed((1, v, e, 1.0)))
    assert edge_set(out) == {(0, 3), (0, 6), (6, 7), (7, 8)}
    assert len(out["vertices"]) == len(v)  # the ring's other nodes stay until prune_unused
    assert list(map(tuple, out["edges"].tolist())) == sorted(edge_set(out))  # sorted (min, max) rows
    pruned = oracle.post("prune_unused", out)
    assert len(pruned["vertices"]) == 5  # 0, 3 and the tail


def test_two_branch_nodes_keep_the_shorter_arc():
    v, e = with_tails(*ring(8), [0, 2])
    out = oracle.post("remove_loops", packed((1, v, e, 1.0)))
    tails = {(0, 8), (8, 9), (2, 10), (10, 11)}
    assert edge_set(out) == tails | {(0, 1), (1, 2)}


def test_three_branch_nodes_hub_or_snip():
    v, e = with_tails(*ring(8), [0, 1, 2])
    tails = {(0, 8), (8, 9), (1, 10), (10, 11), (2, 12), (12, 13)}
    star = oracle.post("remove_loops", packed((1, v, e, 100.0)))  # hub = node 1, every branch node within its radius
    assert edge_set(star) == tails | {(0, 1), (1, 2)}
    snip = oracle.post("remove_loops", packed((1, v, e, 1.0)))  # node 0 lies 7.65 from the hub: one edge goes
    kept = edge_set(snip)
    assert tails <= kept and len(kept) == len(e) - 1 and len(kept - tails) == 7


def _forest(p):
    roots = oracle.post_components(p)
    for k in range(len(p["labels"])):
        v0, v1 = int(p["v_off"][k]), int(p["v_off"][k + 1])
        e0, e1 = int(p["e_off"][k]), int(p["e_off"][k + 1])
        used = set(np.asarray(p["edges"][e0:e1]).ravel().tolist())
        comps = {int(roots[v0 + i]) for i in used}
        if e1 > e0:
            assert e1 - e0 == len(used) - len(comps), k


def random_packed(seed, labels=3, cycles=True, isolated=False):
    """Random graphs on unique integer positions: per label 1-3 trees of 4-40 vertices, extra edges
    closing cycles when `cycles`, radii in [0.5, 4)."""
    rng = np.random.default_rng(seed)
    parts = []
    for label in range(labels):
        v, e, r = [], [], []
        for _ in range(int(rng.integers(1, 4))):
            n = int(rng.integers(4, 40))
            base = len(v)
            start = rng.integers(0, 40, 3).astype(float)
            pts = [start]
            for i in range(1, n):
                parent = int(rng.integers(max(0, i - 3), i))
                pts.append(pts[parent] + rng.integers(-2, 3, 3))
                e.append((base + parent, base + i))
            v.extend(pts)
            r.extend(rng.uniform(0.5, 4.0, n))
            if cycles:
                for _ in range(int(rng.integers(0, 4))):
                    a, b = sorted(rng.integers(0, n, 2).tolist())
                    if a != b and (base + a, base + b) not in e:
                        e.append((base + a, base + b))
        v = np.array(v, np.float64)
        _, first = np.unique(v, axis=0, return_index=True)  # unique positions: drop repeats and their edges
        keep = np.zeros(len(v), bool)
        keep[first] = True
        remap = np.cumsum(keep) - 1
        e = [(int(remap[a]), int(remap[b])) for a, b in e if keep[a] and keep[b]]
        e = [(min(a, b), max(a, b)) for a, b in e if a != b]
        e = list(dict.fromkeys(e))
        rng.shuffle(e)
        parts.append((int(label * 3 + 1), v[keep] * np.array([1.0, 1.5, 2.0]), e, np.array(r)[keep]))
    return packed(*parts)


@pytest.mark.parametrize("seed", range(6))
def test_remove_loops_leaves_forests(seed):
    p = random_packed(seed)
    out = oracle.post("remove_loops", p)
    _forest(out)
    same(oracle.post("remove_loops", out), out)  # no cycle left: returned as is
    trees = random_packed(seed, cycles=False)
    same(oracle.post("remove_loops", trees), trees)


def test_ticks_on_a_T():
    a, ea = chain(1, 10)            # +x arm, 10 long from the branch node at the origin
    b, eb = chain(-1, 10, step=-1)  # -x arm
    c, ec = chain(1, 2, axis=1)     # +y arm, 2 long
    v = np.concatenate([[[0, 0, 0]], a, b, c])
    e = [(0, 1)] + [(x + 1, y + 1) for x, y in ea] + [(0, 11)] + [(x + 11, y + 11) for x, y in eb] + [(0, 21)] + \
        [(x + 21, y + 21) for x, y in ec]
    p = packed((3, v, e, 1.0))
    out = oracle.post("remove_ticks", p, 5.0)
    assert edge_set(out) == {tuple(sorted(x)) for x in e[:20]}
    assert out["edges"].tolist() == np.asarray(e[:20]).tolist()  # the edges keep their order
    same(oracle.post("remove_ticks", p, 1.5), p)


def test_ticks_on_an_H():
    bridge, eb = chain(1, 4)                                  # b1 = 0 at x 0, b2 at x 5
    v = [[0, 0, 0]] + bridge.tolist() + [[5, 0, 0]]
    e = [(0, 1)] + [(x + 1, y + 1) for x, y in eb] + [(4, 5)]

    def arm(at, n, direction):
        x0 = v[at]
        prev = at
        for k in range(1, n + 1):
            v.append([x0[0], x0[1] + direction * k, 0])
            e.append((prev, len(v) - 1))
            prev = len(v) - 1
        return set(range(len(e) - n, len(e)))

    short1, long1 = arm(0, 2, 1), arm(0, 10, -1)
    short2, long2 = arm(5, 3, 1), arm(5, 10, -1)
    p = packed((1, np.array(v, float), e, 1.0))
    rows = lambda out: {tuple(x) for x in out["edges"].tolist()}
    everything = {tuple(x) for x in e}
    gone = lambda idx: {tuple(e[i]) for i in idx}
    assert rows(oracle.post("remove_ticks", p, 4.0)) == everything - gone(short1) - gone(short2)
    assert rows(oracle.post("remove_ticks", p, 2.5)) == everything - gone(short1)


def test_dust_by_cable_length():
    a, ea = chain(0, 5)               # 4 long
    b, eb = chain(0, 11, start=(0, 5, 0))  # 10 long
    p = packed((5, np.concatenate([a, b]), ea + [(x + 5, y + 5) for x, y in eb], 1.0))
    assert oracle.post("remove_dust", p, 4.0)["v_off"].tolist() == [0, 11]
    same(oracle.post("remove_dust", p, 3.999), p)
    assert oracle.post("remove_dust", p, 10.0)["v_off"].tolist() == [0, 0]


def test_dust_float32_rule_inside_the_band():
    """A random 60-edge chain: at a threshold halfway between its double cable length and NumPy's
    float32 pairwise sum of float32 edge lengths, the two disagree, and the float32 rule decides."""
    rng = np.random.default_rng(11)
    verts = np.cumsum(rng.uniform(0.1, 1.7, (61, 3)), axis=0).astype(np.float32)
    e = [(i, i + 1) for i in range(60)]
    d = (verts[1:] - verts[:-1]).astype(np.float32)
    sq = (d * d).astype(np.float32)
    lengths = np.sqrt((sq[:, 0] + sq[:, 1]) + sq[:, 2]).astype(np.float32)
    f = float(np.sum(lengths))
    total = float(np.sum(np.sqrt(np.sum((verts[1:].astype(np.float64) - verts[:-1].astype(np.float64)) ** 2, 1))))
    assert total != f
    threshold = (f + total) / 2
    p = packed((1, verts, e, 1.0))
    got = oracle.post("remove_dust", p, threshold)
    if f > threshold:
        same(got, p)
    else:
        assert got["v_off"].tolist() == [0, 0]


def test_joins():
    a, ea = chain(0, 3)
    b, eb = chain(5, 3)
    near = packed((1, np.concatenate([a, b]), ea + [(x + 3, y + 3) for x, y in eb], 2.0))
    out = oracle.post("join_components", near)
    assert edge_set(out) == {(0, 1), (1, 2), (3, 4), (4, 5), (2, 3)}
    c, ec = chain(7, 3)
    far = packed((1, np.concatenate([a, c]), ea + [(x + 3, y + 3) for x, y in ec], 2.0))
    same(oracle.post("join_components", far), far)  # gap 5 beyond the reach 2 * 2
    radii = np.full(6, 0.5, np.float32)
    radii[0] = 3.0  # reach 6, but the nearest pair's radii sum to 1 < 3: blocked
    blocked = packed((1, np.concatenate([a, b]), ea + [(x + 3, y + 3) for x, y in eb], radii))
    same(oracle.post("join_components", blocked), blocked)
    labelled = packed((1, a, ea, 2.0), (2, b, eb, 2.0))
    same(oracle.post("join_components", labelled), labelled)
    b2, _ = chain(4, 3)
    c2, _ = chain(9, 3)
    three = packed((1, np.concatenate([a, b2, c2]), ea + [(x + 3, y + 3) for x, y in ea] + [(x + 6, y + 6) for x, y in ea], 2.0))
    out = oracle.post("join_components", three)
    assert edge_set(out) - edge_set(three) == {(2, 3), (5, 6)}


@pytest.mark.parametrize("seed", range(4))
def test_postprocess_properties(seed):
    p = random_packed(seed)
    out = oracle.post("postprocess", p, 3.0, 2.0)
    assert out["labels"].tolist() == p["labels"].tolist()
    _forest(out)
    degree = np.zeros(len(out["radii"]), int)
    for k in range(len(out["labels"])):
        v0, e0, e1 = int(out["v_off"][k]), int(out["e_off"][k]), int(out["e_off"][k + 1])
        for x in np.asarray(out["edges"][e0:e1]).ravel():
            degree[v0 + int(x)] += 1
    assert (degree > 0).all()  # prune_unused
    with pytest.raises(ValueError, match="out of bounds"):
        bad = dict(p)
        bad["edges"] = p["edges"].copy()
        bad["edges"][0, 0] = 10 ** 6
        oracle.post("postprocess", bad, 1, 1)


# ---- GPU ----

@pytest.fixture(scope="module")
def post():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import brook_amd
    return brook_amd.device_post


def as_packed(p, device=None):
    out = PackedSkeletons(p["labels"], p["v_off"], p["e_off"], np.asarray(p["vertices"], np.float32).reshape(-1, 3),
                          np.asarray(p["edges"], np.uint32).reshape(-1, 2), np.asarray(p["radii"], np.float32), (1.0, 1.0, 1.0))
    return out.to_device(device) if device else out


def gpu_same(got, want, what):
    g = got.as_dict()
    for name in ("labels", "v_off", "e_off", "vertices", "edges", "radii"):
        assert np.array_equal(np.asarray(g[name]).reshape(np.asarray(want[name]).shape), np.asarray(want[name])), (what, name)


def _cases():
    out = []
    v, e = ring(6)
    out.append(("ring", packed((7, v, e, 1.0))))
    v, e = with_tails(*ring(6), [0], length=3)
    out.append(("lollipop", packed((1, v, e, 1.0))))
    v, e = with_tails(*ring(8), [0, 2])
    out.append(("two", packed((1, v, e, 1.0))))
    v, e = with_tails(*ring(8), [0, 1, 2])
    out.append(("star", packed((1, v, e, 100.0))))
    out.append(("snip", packed((1, v, e, 1.0))))
    for seed in range(8):
        out.append((f"random{seed}", random_packed(seed, labels=4)))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("step,a,b", [("remove_dust", 6.0, 0), ("remove_loops", 0, 0), ("join_components", 0, 0),
                                      ("remove_ticks", 3.0, 0), ("prune_unused", 0, 0), ("postprocess", 6.0, 3.0),
                                      ("postprocess", 0, 0)])
def test_gpu_post_equals_oracle(post, step, a, b):
    fn = {"remove_dust": lambda p: post.remove_dust(p, a), "remove_loops": post.remove_loops,
          "join_components": post.join_close_components, "remove_ticks": lambda p: post.remove_ticks(p, a),
          "prune_unused": post.prune_unused, "postprocess": lambda p: post.postprocess(p, a, b)}[step]
    for name, p in _cases():
        print(f"post {step} {name}", flush=True)
        want = oracle.post(step, p, a, b)
        gpu_same(fn(as_packed(p)), want, (step, name))
    p = _cases()[-1][1]
    gpu_same(fn(as_packed(p, "cuda")), oracle.post(step, p, a, b), (step, "device input"))


@pytest.mark.gpu
def test_gpu_post_components_and_errors(post):
    for name, p in _cases():
        assert np.array_equal(post.components(as_packed(p)), oracle.post_components(p)), name
    bad = dict(_cases()[-1][1])
    bad["edges"] = bad["edges"].copy()
    bad["edges"][0, 0] = 10 ** 6
    with pytest.raises(ValueError, match="out of bounds"):
        post.postprocess(as_packed(bad), 1, 1)


@pytest.mark.gpu
def test_gpu_post_on_skeletons_and_merges(post):
    """Skeletons of the tube workload merged from overlapping chunks: the postprocess equals the
    oracle's, and without thresholds leaves a one-piece result unchanged (cpp_smoke.cpp:44-48)."""
    import brook_amd
    from brook_amd import packed as bp
    from brook_amd import synth
    params = dict(scale=1.5, const=2, pdrf_scale=100000, pdrf_exponent=4)
    labels = synth.volume("tubes", (70, 50, 40), 6, 11)
    whole = brook_amd.skeletonize(labels, params, dust_threshold=0, output="packed", progress=False)
    same_input = post.postprocess(whole, 0, 0)
    gpu_same(same_input, whole.as_dict(), "unchanged")
    half = labels.shape[0] // 2
    left = brook_amd.skeletonize(np.asfortranarray(labels[:half + 4]), params, dust_threshold=0, output="packed", progress=False)
    right = brook_amd.skeletonize(np.asfortranarray(labels[half - 4:]), params, dust_threshold=0, output="packed", progress=False)
    merged = bp.merge_fragments([left, right], [(0, 0, 0), (half - 4, 0, 0)])
    for dust, ticks in ((0, 0), (5, 3), (20, 8)):
        want = oracle.post("postprocess", merged.as_dict(), dust, ticks)
        gpu_same(post.postprocess(merged, dust, ticks), want, ("merged", dust, ticks))


@pytest.mark.gpu
def test_gpu_dust_band_uses_the_float32_rule(post):
    """Chains of 60, 200 and 1000 edges (the last two split NumPy's pairwise blocks) at thresholds
    halfway between the double and float32 cable lengths: the device's pairwise sum decides."""
    rng = np.random.default_rng(5)
    for n in (60, 200, 1000):
        verts = np.cumsum(rng.uniform(0.1, 1.7, (n + 1, 3)), axis=0).astype(np.float32)
        d = (verts[1:] - verts[:-1]).astype(np.float32)
        sq = (d * d).astype(np.float32)
        f = float(np.sum(np.sqrt((sq[:, 0] + sq[:, 1]) + sq[:, 2]).astype(np.float32)))
        total = float(np.sum(np.sqrt(np.sum((verts[1:].astype(np.float64) - verts[:-1].astype(np.float64)) ** 2, 1))))
        if total == f:
            continue
        threshold = (f + total) / 2
        p = packed((1, verts, [(i, i + 1) for i in range(n)], 1.0), (2, verts[:3] + 100, [(0, 1), (1, 2)], 1.0))
        want = oracle.post("remove_dust", p, threshold)
        assert (want["v_off"][1] > 0) == (f > threshold) != (total > threshold)
        gpu_same(post.remove_dust(as_packed(p), threshold), want, n)
