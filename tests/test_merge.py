"""Packed merge (the reference's merge_fragments, packed.cpp:78-86 over packed_merge.cu:88-166).

CPU tests pin the oracle restatement (oracle/merge.py): a hand-worked known answer (shared and
duplicated vertices keep the first radius, -0 meets +0, self-loops, duplicate and reversed edges,
an unused vertex, a label left with no edge), NaN vertices that never merge, and properties on
real skeletons (split-and-merge equals the whole, idempotence, translation by the origin). The
reference's own tests hold no merge vectors (tests/python/test_package.py:204-219 only checks a
one-fragment merge's length). GPU tests compare brook_amd.packed.merge_fragments with the oracle
bit for bit.
"""
import numpy as np
import pytest

import oracle
from tests.golden import shapes

F32 = np.float32


def _part(skeletons):
    """Packed dict from [(label, vertices, radii, edges)]."""
    labels, verts, radii, edges, v_off, e_off = [], [], [], [], [0], [0]
    for label, v, r, e in skeletons:
        labels.append(label)
        verts.append(np.asarray(v, F32).reshape(-1, 3))
        radii.append(np.asarray(r, F32).reshape(-1))
        edges.append(np.asarray(e, np.uint32).reshape(-1, 2))
        v_off.append(v_off[-1] + len(verts[-1]))
        e_off.append(e_off[-1] + len(edges[-1]))
    return dict(labels=np.asarray(labels, np.int64), v_off=np.asarray(v_off, np.int64), e_off=np.asarray(e_off, np.int64),
                vertices=np.concatenate(verts) if verts else np.zeros((0, 3), F32),
                edges=np.concatenate(edges) if edges else np.zeros((0, 2), np.uint32),
                radii=np.concatenate(radii) if radii else np.zeros(0, F32))


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _assert_same(got, want):
    g = got.as_dict() if hasattr(got, "as_dict") else got
    assert g["labels"].tolist() == want["labels"].tolist()
    assert g["v_off"].tolist() == want["v_off"].tolist()
    assert g["e_off"].tolist() == want["e_off"].tolist()
    assert np.array_equal(_bits(np.asarray(g["vertices"], F32)), _bits(want["vertices"]))
    assert np.array_equal(_bits(np.asarray(g["radii"], F32)), _bits(want["radii"]))
    assert np.array_equal(np.asarray(g["edges"], np.uint32), want["edges"])


def known_case():
    """Two fragments of label 5 meeting at x = 2 and 0, label 3 with one vertex and no edge, label 9
    empty; worked by hand below."""
    a = _part([(5, [(0, 0, 0), (1, 0, 0), (2, 0, 0), (7, 7, 7)], [1, 2, 3, 4], [(0, 1), (1, 2), (1, 0)]),
               (3, [(4, 4, 4)], [9], []),
               (9, [], [], [])])
    # origin (2, -0, 0): vertex 0 lands on a's (2,0,0), vertex 2 on a's (0,0,0) as (0,-0,0)
    b = _part([(5, [(0, 0, 0), (1, 0, 0), (-2, -0.0, 0)], [5, 6, 7], [(0, 1), (2, 2), (0, 2)])])
    return [a, b], [(0, 0, 0), (2, -0.0, 0)]


def known_answer():
    # merged ids: label 3 -> (4,4,4) id 0; label 5 -> (0,0,0) 1, (1,0,0) 2, (2,0,0) 3, (3,0,0) 4, (7,7,7) 5
    # edges: a (1,2) (2,3) (1,2); b (3,4), self-loop (1,1), (1,3) -> unique (1,2) (1,3) (2,3) (3,4)
    # ids 0 and 5 are on no edge; label 3 stays with no vertex
    return dict(labels=np.array([3, 5]), v_off=np.array([0, 0, 4]), e_off=np.array([0, 0, 4]),
                vertices=np.array([(0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 0, 0)], F32),
                radii=np.array([1, 2, 3, 6], F32), edges=np.array([(0, 1), (0, 2), (1, 2), (2, 3)], np.uint32))


def nan_case():
    nan = np.nan
    p = _part([(1, [(nan, 0, 0), (nan, 0, 0), (0, 0, 0)], [1, 2, 3], [(0, 2), (1, 2)])])
    want = dict(labels=np.array([1]), v_off=np.array([0, 3]), e_off=np.array([0, 2]),
                vertices=np.array([(0, 0, 0), (nan, 0, 0), (nan, 0, 0)], F32), radii=np.array([3, 1, 2], F32),
                edges=np.array([(0, 1), (0, 2)], np.uint32))
    return [p], [(0, 0, 0)], want


def _skeleton_fragments():
    labels = shapes.sine_bands()
    whole = oracle.skeletonize(labels, **shapes.GOLDEN_OPTIONS)
    assert len(whole["labels"]) >= 2 and len(whole["edges"]) > 20
    return whole


def _split(whole):
    """Two fragments holding every vertex of each skeleton and alternate edges."""
    halves = []
    for parity in (0, 1):
        rows = []
        for k, label in enumerate(whole["labels"].tolist()):
            v0, v1 = whole["v_off"][k], whole["v_off"][k + 1]
            e0, e1 = whole["e_off"][k], whole["e_off"][k + 1]
            rows.append((label, whole["vertices"][v0:v1], whole["radii"][v0:v1], whole["edges"][e0:e1][parity::2]))
        halves.append(_part(rows))
    return halves


def test_oracle_known_answer():
    parts, origins = known_case()
    _assert_same(oracle.merge_fragments(parts, origins), known_answer())


def test_oracle_nan_vertices_never_merge():
    parts, origins, want = nan_case()
    _assert_same(oracle.merge_fragments(parts, origins), want)


def test_oracle_errors_and_empty():
    with pytest.raises(ValueError):
        oracle.merge_fragments([_part([(1, [(0, 0, 0)], [1], [(0, 1)])])], [(0, 0, 0)])
    with pytest.raises(ValueError):
        oracle.merge_fragments([_part([(1, [], [], [(0, 0)])])], [(0, 0, 0)])
    out = oracle.merge_fragments([_part([(1, [], [], [])])], [(0, 0, 0)])
    assert out["labels"].size == 0 and out["v_off"].tolist() == [0]


def test_oracle_split_idempotent_and_translation():
    whole = _skeleton_fragments()
    once = oracle.merge_fragments([whole], [(0, 0, 0)])
    # a skeleton is a tree on distinct positions: the merge only reorders it
    assert once["labels"].tolist() == whole["labels"].tolist()
    assert np.diff(once["e_off"]).tolist() == np.diff(whole["e_off"]).tolist()
    assert np.diff(once["v_off"]).tolist() == np.diff(whole["v_off"]).tolist()
    for k in range(len(once["labels"])):
        a = whole["vertices"][whole["v_off"][k]:whole["v_off"][k + 1]]
        b = once["vertices"][once["v_off"][k]:once["v_off"][k + 1]]
        assert np.array_equal(np.unique(a, axis=0), b)  # sorted (x, y, z), one per position
        ra = dict(zip(map(tuple, a.tolist()), whole["radii"][whole["v_off"][k]:whole["v_off"][k + 1]].tolist()))
        rb = once["radii"][once["v_off"][k]:once["v_off"][k + 1]].tolist()
        assert [ra[tuple(v)] for v in b.tolist()] == rb
    _assert_same(oracle.merge_fragments(_split(whole), [(0, 0, 0), (0, 0, 0)]), once)
    _assert_same(oracle.merge_fragments([once], [(0, 0, 0)]), once)
    moved = oracle.merge_fragments([whole], [(16, -8, 40)])
    assert np.array_equal(moved["vertices"], once["vertices"] + np.array([16, -8, 40], F32))
    assert np.array_equal(moved["edges"], once["edges"])


def random_fragments(seed, fragments=4, labels=6, grid=6):
    """Fragments of random skeleton pieces on a small integer grid (many shared positions), with
    self-loops, duplicate edges, -0 coordinates and labels without vertices."""
    rng = np.random.default_rng(seed)
    parts, origins = [], []
    for _ in range(fragments):
        rows = []
        for label in rng.permutation(labels)[: rng.integers(1, labels + 1)]:
            n = int(rng.integers(0, 40))
            v = rng.integers(-grid, grid, size=(n, 3)).astype(F32)
            v[rng.random((n, 3)) < 0.1] = -0.0
            r = rng.random(n).astype(F32)
            m = int(rng.integers(0, 2 * n + 1)) if n else 0
            e = rng.integers(0, max(n, 1), size=(m, 2)).astype(np.uint32)
            rows.append((int(label) * 7 + 1, v, r, e))
        parts.append(_part(rows))
        origins.append(tuple(float(x) for x in rng.integers(-2, 3, size=3)))
    return parts, origins


def _with_vertices(d):
    """The labels of a packed dict that have vertices (a merge lists no other)."""
    keep = np.flatnonzero(np.diff(d["v_off"]) > 0)
    return dict(d, labels=d["labels"][keep], v_off=np.concatenate([[0], d["v_off"][keep + 1]]),
                e_off=np.concatenate([[0], d["e_off"][keep + 1]]))


def test_oracle_random_fragments_idempotent_and_preshifted():
    saw_empty = False
    for seed in range(5):
        parts, origins = random_fragments(seed)
        once = oracle.merge_fragments(parts, origins)
        saw_empty |= len(_with_vertices(once)["labels"]) < len(once["labels"])
        _assert_same(oracle.merge_fragments([once], [(0, 0, 0)]), _with_vertices(once))
        shifted = [dict(p, vertices=p["vertices"] + np.asarray(o, F32)) for p, o in zip(parts, origins)]
        _assert_same(oracle.merge_fragments(shifted, [(0, 0, 0)] * len(parts)), once)
    assert saw_empty  # some label kept with no vertex


# ---- GPU ----

@pytest.fixture(scope="module")
def packed_api():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from brook_amd import packed
    return packed


def _packed(d):
    from brook_amd import PackedSkeletons
    return PackedSkeletons(d["labels"], d["v_off"], d["e_off"], d["vertices"], d["edges"], d["radii"], (1.0, 1.0, 1.0))


@pytest.mark.gpu
def test_gpu_merge_known_answer_and_nan(packed_api):
    parts, origins = known_case()
    _assert_same(packed_api.merge_fragments([_packed(p) for p in parts], origins), known_answer())
    parts, origins, want = nan_case()
    _assert_same(packed_api.merge_fragments([_packed(p) for p in parts], origins), want)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(6))
def test_gpu_merge_random_fragments(packed_api, seed):
    parts, origins = random_fragments(100 + seed, fragments=8, labels=12, grid=5)
    want = oracle.merge_fragments(parts, origins)
    _assert_same(packed_api.merge_fragments([_packed(p) for p in parts], origins), want)


@pytest.mark.gpu
def test_gpu_merge_device_parts_and_errors(packed_api):
    import torch
    parts, origins = random_fragments(7, fragments=3)
    want = oracle.merge_fragments(parts, origins)
    got = packed_api.merge_fragments([_packed(p).to_device("cuda:0") for p in parts], origins, on_device=True)
    assert got.on_device
    _assert_same(got.to_host(), want)
    with pytest.raises(ValueError):
        packed_api.merge_fragments([_packed(parts[0])], [])
    bad = _part([(1, [(0, 0, 0)], [1], [(0, 1)])])
    with pytest.raises(Exception, match="out of bounds"):
        packed_api.merge_fragments([_packed(bad)], [(0, 0, 0)])
    assert len(packed_api.merge_fragments([_packed(_part([(1, [], [], [])]))], [(0, 0, 0)])) == 0
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_gpu_merge_chunked_skeletons(packed_api):
    """Skeletons of the four x-halves of a volume (overlapping by one plane), merged at their
    origins, equal the oracle's merge of the same fragments."""
    import brook_amd
    from brook_amd import synth
    labels = synth.tubes_128(1001)
    aniso = (16.0, 16.0, 40.0)
    params = dict(scale=1.5, const=300, pdrf_scale=100000, pdrf_exponent=4)
    cuts = [(0, 40), (39, 80), (79, 128)]
    parts, origins = [], []
    for x0, x1 in cuts:
        p = brook_amd.skeletonize(np.asfortranarray(labels[x0:x1]), params, aniso, output="packed", progress=False)
        parts.append(p)
        origins.append((x0 * aniso[0], 0.0, 0.0))
    want = oracle.merge_fragments([p.as_dict() for p in parts], origins)
    got = packed_api.merge_fragments(parts, origins)
    _assert_same(got, want)
    assert len(got) >= 10
