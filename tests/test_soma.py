"""Somata (the batched reference path for soma labels without enclosed voids: batched.cu:58-108,
186-231; lockstep.cu:113-206; lockstep_regular.inc:35-202).

CPU tests pin the oracle: the reference's own soma test (test_package.py:347-367: a ball with six
processes -- every process traced, the soma root a vertex, one connected skeleton) in both
branching modes, the root rule on a plateau of maxima with a known centroid and tie, the seeded
DAF against a brute-force float32 Bellman-Ford from the free-space seeds and the seed formula
against its closed forms, trimming (no vertex within the soma radius but the root), and filled
voids: a soma whose box is the whole volume skeletonizes like the volume with its cavity filled
(a label inside the cavity disappears); a soma whose voids hold part of another label is traced in
a private arena (soma.cu:54-72, 207-215): its skeleton is the filled soma's alone, the other label's
is the one it has without any soma, and the arena moves with its box bit for bit. GPU tests compare
brook_amd with the oracle bit for bit.
"""
import numpy as np
import pytest

import oracle
from tests.golden import shapes

P = shapes.SOMA_PARAMS


def _components(vertices, edges):
    parent = list(range(len(vertices)))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for a, b in edges:
        parent[find(int(a))] = find(int(b))
    return len({find(i) for i in range(len(vertices))})


@pytest.mark.parametrize("fix_branching", [True, False])
def test_oracle_soma_ball_reference_properties(fix_branching):
    labels = shapes.soma_ball()
    r = oracle.skeletonize(labels, P, (1, 1, 1), 0, fix_branching=fix_branching, fix_borders=False)
    assert r["labels"].tolist() == [1]
    v, e = r["vertices"], r["edges"]
    offsets = np.abs(v - 36)
    assert (offsets.max(axis=1) > 30).sum() >= 6, "every process is traced"
    assert np.any(np.all(v == 36, axis=1)), "the soma root is a vertex"
    assert _components(v, e) == 1
    radius = 12.041595 * P["soma_invalidation_scale"] + P["soma_invalidation_const"]  # dbf_max = sqrt(145)
    d = np.sqrt(((v - 36.0) ** 2).sum(axis=1))
    assert np.all((d > radius) | (d == 0)), "only the root lies within the soma radius"
    assert r["stats"]["paths"] == 6


def test_oracle_soma_root_on_a_plateau():
    """A 40 x 9 x 9 bar (DBF maximum 5 on a segment of the axis x = 4..35, 32 voxels) with a
    process: the centroid of the maxima is x = 19.5, two candidates tie at 19 and 20, and the first
    in (x, y, z) order wins."""
    labels = np.zeros((48, 24, 24), np.uint8)
    labels[4:44, 8:17, 8:17] = 1
    labels[20:23, 16:23, 11:14] = 1  # a process on one side
    params = dict(P, soma_detection_threshold=3, soma_acceptance_threshold=4)
    p = oracle.prepare(labels, params, (1, 1, 1), 0, fix_borders=False)
    assert p["soma"][1] == 1 and p["dbf_max"][1] == 5.0
    dbf = p["dbf"]
    maxima = np.argwhere((dbf == 5.0) & (p["cc"] == 1))
    lo = np.argwhere(labels).min(axis=0)
    rel = maxima - lo
    centre = rel.mean(axis=0).astype(np.float32).astype(np.float64)
    cost = ((rel - centre) ** 2).sum(axis=1)
    order = np.lexsort((rel[:, 2], rel[:, 1], rel[:, 0]))
    best = order[np.argmin(cost[order])]  # first minimum in sorted order
    x, y, z = maxima[best]
    assert p["root"][1] == x + labels.shape[0] * (y + labels.shape[1] * z)
    assert np.count_nonzero(cost == cost.min()) >= 2  # the tie is exercised


def _freespace(shape, root, w, radius):
    """The free-space seed in float32 (freespace_seed_lab's expression) and the mask of the ball."""
    f32 = np.float32
    w = [f32(a) for a in w]
    X, Y, Z = np.indices(shape)
    dxi, dyi, dzi = X - root[0], Y - root[1], Z - root[2]
    rdx, rdy, rdz = w[0] * dxi.astype(f32), w[1] * dyi.astype(f32), w[2] * dzi.astype(f32)
    inside = rdx * rdx + rdy * rdy + rdz * rdz <= f32(radius) * f32(radius)
    dx, dy, dz = (np.abs(a.astype(f32)) for a in (dxi, dyi, dzi))
    m3 = np.minimum(np.minimum(dx, dy), dz)
    mxy, myz, mxz = np.minimum(dx, dy), np.minimum(dy, dz), np.minimum(dx, dz)
    val = (m3 * np.sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) + w[0] * (dx - m3) + w[1] * (dy - m3) + w[2] * (dz - m3)
           + (mxy - m3) * (np.sqrt(w[0] * w[0] + w[1] * w[1]) - w[0] - w[1])
           + (mxz - m3) * (np.sqrt(w[0] * w[0] + w[2] * w[2]) - w[0] - w[2])
           + (myz - m3) * (np.sqrt(w[1] * w[1] + w[2] * w[2]) - w[1] - w[2]))
    return val.astype(f32), inside


def _bellman_ford(mask, start, w):
    """Float32 Bellman-Ford to the fixpoint over 26-neighbours inside `mask` from initial values."""
    f32 = np.float32
    d = start.copy()
    steps = []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                if dx or dy or dz:
                    steps.append(((dx, dy, dz), f32(np.sqrt(float(w[0]) ** 2 * dx * dx + float(w[1]) ** 2 * dy * dy
                                                             + float(w[2]) ** 2 * dz * dz))))
    pad = np.full(tuple(s + 2 for s in d.shape), np.inf, f32)
    for _ in range(10000):
        pad[1:-1, 1:-1, 1:-1] = d
        new = d.copy()
        for (dx, dy, dz), c in steps:
            nb = pad[1 + dx:pad.shape[0] - 1 + dx, 1 + dy:pad.shape[1] - 1 + dy, 1 + dz:pad.shape[2] - 1 + dz]
            new = np.minimum(new, np.where(mask, nb + c, np.inf).astype(f32))
        new[~mask] = np.inf
        if np.array_equal(new, d):
            return d
        d = new
    raise AssertionError("no fixpoint")


@pytest.mark.parametrize("aniso", [(1, 1, 1), (2, 3, 5)])
def test_oracle_soma_daf_is_the_seeded_fixpoint(aniso):
    labels = shapes.soma_ball(size=40, centre=20, radius=9)
    p = oracle.prepare(labels, P, aniso, 0, fix_borders=False)
    assert p["soma"][1] == 1
    cc, daf = p["cc"], p["daf"]
    r = int(p["root"][1])
    root = (r % 40, (r // 40) % 40, r // 1600)
    val, inside = _freespace(labels.shape, root, aniso, p["dbf_max"][1])
    mask = cc == 1
    start = np.full(labels.shape, np.inf, np.float32)
    seeds = inside & mask
    start[seeds] = val[seeds]
    start[root] = 0.0
    want = _bellman_ford(mask, start, aniso)
    assert np.array_equal(daf[mask], want[mask])
    # the seeds matter: some ball voxels keep a seed below the plain flood's value
    single = np.full(labels.shape, np.inf, np.float32)
    single[root] = 0.0
    plain = _bellman_ford(mask, single, aniso)
    assert np.any(daf[seeds] < plain[seeds])


def test_freespace_seed_closed_forms():
    """Along an axis the seed is w * |d|; along a face or body diagonal it is the diagonal step
    times |d|; it never undercuts the Euclidean distance (beyond rounding)."""
    w = (2.0, 3.0, 5.0)
    val, inside = _freespace((21, 21, 21), (10, 10, 10), w, 1e9)
    f32 = np.float32
    for k in range(-10, 11):
        assert val[10 + k, 10, 10] == f32(2.0) * f32(abs(k))
        assert val[10, 10, 10 + k] == f32(5.0) * f32(abs(k))
    for k in range(-10, 11):
        assert abs(val[10 + k, 10 + k, 10] - abs(k) * np.sqrt(13.0)) <= 1e-5 * max(1, abs(k)) * 4
        assert abs(val[10 + k, 10 + k, 10 + k] - abs(k) * np.sqrt(38.0)) <= 1e-5 * max(1, abs(k)) * 7
    X, Y, Z = np.indices((21, 21, 21)) - 10
    euclid = np.sqrt((2.0 * X) ** 2 + (3.0 * Y) ** 2 + (5.0 * Z) ** 2)
    assert np.all(val >= euclid * (1 - 1e-6) - 1e-6)


def test_oracle_soma_cells_trimming_and_voids():
    labels = shapes.soma_cells()
    p = oracle.prepare(labels, P, (1, 1, 1), 0, fix_borders=False)
    assert p["soma"].sum() >= 2 and ((p["dbf_max"] > 6) & (p["soma"] == 0)).sum() >= 1  # somata and a checked non-soma
    r = oracle.skeletonize(labels, P, (1, 1, 1), 0, fix_borders=False)
    for k, label in enumerate(r["labels"].tolist()):
        v = r["vertices"][r["v_off"][k]:r["v_off"][k + 1]]
        e = r["edges"][r["e_off"][k]:r["e_off"][k + 1]]
        assert _components(v, e) == 1, label
    for i in range(1, p["count"] + 1):
        if not p["soma"][i]:
            continue
        root = int(p["root"][i])
        rp = np.array([root % labels.shape[0], (root // labels.shape[0]) % labels.shape[1],
                       root // (labels.shape[0] * labels.shape[1])], np.float64)
        label = int(labels[tuple(rp.astype(int))])
        k = r["labels"].tolist().index(label)
        v = r["vertices"][r["v_off"][k]:r["v_off"][k + 1]].astype(np.float64)
        d = np.sqrt(((v - rp) ** 2).sum(axis=1))
        radius = float(p["dbf_max"][i]) * P["soma_invalidation_scale"] + P["soma_invalidation_const"]
        assert np.all((d > radius) | (d == 0)), label
        assert np.any(d == 0), label


def _prefilled(labels):
    out = labels.copy()
    c = labels.shape[0] // 2
    out[c + 3:c + 8, c - 2:c + 3, c - 2:c + 3] = 1
    return out


@pytest.mark.parametrize("fix_branching", [True, False])
@pytest.mark.parametrize("nucleus", [False, True])
def test_oracle_soma_voids_equal_a_prefilled_volume(fix_branching, nucleus):
    """The soma's box is the whole volume, so its refreshed DBF (the EDT of the filled box) is the
    EDT of the volume with the cavity filled: the skeletons equal those of that volume, and a
    label wholly inside the cavity is gone."""
    labels = shapes.soma_cavity(nucleus=nucleus)
    got = oracle.skeletonize(labels, P, (1, 1, 1), 0, fix_branching=fix_branching, fix_borders=False)
    want = oracle.skeletonize(_prefilled(labels), P, (1, 1, 1), 0, fix_branching=fix_branching, fix_borders=False)
    assert got["labels"].tolist() == [1]
    for name in ("v_off", "e_off", "vertices", "edges", "radii"):
        assert np.array_equal(got[name], want[name]), name
    p = oracle.prepare(labels, P, (1, 1, 1), 0, fix_borders=False)
    assert p["soma"][1] == 1 and (not nucleus or p["active"][2] == 0)


def _label(packed, label):
    k = packed["labels"].tolist().index(label)
    v0, v1, e0, e1 = packed["v_off"][k], packed["v_off"][k + 1], packed["e_off"][k], packed["e_off"][k + 1]
    return packed["vertices"][v0:v1], packed["edges"][e0:e1], packed["radii"][v0:v1]


def _same_label(got, want, label):
    for a, b, name in zip(_label(got, label), _label(want, label), ("vertices", "edges", "radii")):
        assert np.array_equal(a, b), (label, name)


@pytest.mark.parametrize("fix_branching", [True, False])
def test_oracle_soma_private_arena(fix_branching):
    """A staircase (label 3) runs from the soma's cavity through its wall to the outside, so filling
    the soma's voids would take part of it: the soma is traced in a private arena (soma.cu:207-215),
    its filled box alone, and the staircase keeps every voxel. The box is the whole volume here, so
    the soma's skeleton is that of the filled soma alone, and the staircase's is the one it has when
    no soma is detected at all (labels never interact)."""
    labels = shapes.soma_cavity(staircase=True)
    opts = dict(fix_branching=fix_branching, fix_borders=False)
    got = oracle.skeletonize(labels, P, (1, 1, 1), 0, **opts)
    assert got["labels"].tolist() == [1, 3]
    alone = _prefilled(np.where(labels == 1, 1, 0).astype(np.uint8))
    _same_label(got, oracle.skeletonize(alone, P, (1, 1, 1), 0, **opts), 1)
    no_soma = dict(P, soma_detection_threshold=1e9, soma_acceptance_threshold=1e9)
    _same_label(got, oracle.skeletonize(labels, no_soma, (1, 1, 1), 0, **opts), 3)
    # the arena depends on its box alone: padded with 5 background voxels and a far label, the
    # soma's skeleton moves by 5 voxels, bit for bit
    padded = np.pad(labels, ((5, 9), (5, 3), (5, 6)))
    padded[-3:, -2:, -4:] = 7
    moved = oracle.skeletonize(np.asfortranarray(padded), P, (1, 1, 1), 0, **opts)
    assert moved["labels"].tolist() == [1, 3, 7]
    v, e, r = _label(moved, 1)
    v0, e0, r0 = _label(got, 1)
    assert np.array_equal(v, v0 + np.float32(5)) and np.array_equal(e, e0) and np.array_equal(r, r0)
    r = oracle.skeletonize(shapes.soma_with_void(), P, (1, 1, 1), 0, fix_borders=False)
    assert r["labels"].tolist() == [1] and len(r["edges"]) > 0


# ---- GPU ----

@pytest.fixture(scope="module")
def brook():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import brook_amd
    return brook_amd


def _same(got, want, what=""):
    g = got.as_dict()
    assert g["labels"].tolist() == want["labels"].tolist(), what
    for name in ("v_off", "e_off", "vertices", "edges", "radii"):
        assert np.array_equal(np.asarray(g[name]), np.asarray(want[name])), (what, name)


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
@pytest.mark.parametrize("aniso", [(1, 1, 1), (2, 3, 5)])
def test_gpu_soma_ball_equals_oracle(brook, fix_branching, aniso):
    labels = shapes.soma_ball()
    got = brook.skeletonize(labels, P, aniso, dust_threshold=0, fix_branching=fix_branching, fix_borders=False,
                            output="packed", progress=False)
    want = oracle.skeletonize(labels, P, aniso, 0, fix_branching=fix_branching, fix_borders=False)
    _same(got, want, (fix_branching, aniso))


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
@pytest.mark.parametrize("fix_borders", [True, False])
def test_gpu_soma_cells_equal_oracle(brook, fix_branching, fix_borders):
    labels = shapes.soma_cells()
    got = brook.skeletonize(labels, P, (1, 1, 1), dust_threshold=0, fix_branching=fix_branching,
                            fix_borders=fix_borders, output="packed", progress=False)
    want = oracle.skeletonize(labels, P, (1, 1, 1), 0, fix_branching=fix_branching, fix_borders=fix_borders, threads=4)
    _same(got, want, (fix_branching, fix_borders))


@pytest.mark.gpu
def test_gpu_soma_fields_equal_oracle(brook):
    labels = shapes.soma_cells()
    got = brook.prepare_fields(labels, teasar_params=P, anisotropy=(1, 1, 1), dust_threshold=0, fix_borders=True)
    want = oracle.prepare(labels, P, (1, 1, 1), 0, fix_borders=True)
    act = want["active"].astype(bool)
    for name in ("root", "target", "dbf_max", "max_daf"):
        assert np.array_equal(got[name][act], want[name][act]), name
    assert np.array_equal(got["daf"], want["daf"])


@pytest.mark.gpu
def test_gpu_soma_batch_and_max_paths(brook):
    vols = [shapes.soma_cells(), shapes.soma_ball(), shapes.soma_cells(seed=11)]
    batch = brook.skeletonize_batch(vols, teasar_params=P, dust_threshold=0, fix_borders=False, output="packed",
                                    progress=False)
    for i, v in enumerate(vols):
        want = oracle.skeletonize(v, P, (1, 1, 1), 0, fix_borders=False, threads=4)
        _same(batch[i], want, i)
    capped = dict(P, max_paths=2)
    got = brook.skeletonize(shapes.soma_ball(), capped, (1, 1, 1), dust_threshold=0, fix_borders=False,
                            output="packed", progress=False)
    want = oracle.skeletonize(shapes.soma_ball(), capped, (1, 1, 1), 0, fix_borders=False)
    _same(got, want, "max_paths")


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
@pytest.mark.parametrize("case", ["void", "cavity", "nucleus"])
def test_gpu_soma_voids_equal_oracle(brook, fix_branching, case):
    labels = {"void": shapes.soma_with_void(), "cavity": shapes.soma_cavity(),
              "nucleus": shapes.soma_cavity(nucleus=True)}[case]
    got = brook.skeletonize(labels, P, (1, 1, 1), dust_threshold=0, fix_branching=fix_branching, fix_borders=True,
                            output="packed", progress=False)
    want = oracle.skeletonize(labels, P, (1, 1, 1), 0, fix_branching=fix_branching, fix_borders=True)
    _same(got, want, (case, fix_branching))
    fields = brook.prepare_fields(labels, teasar_params=P, anisotropy=(1, 1, 1), dust_threshold=0, fix_borders=True)
    ref = oracle.prepare(labels, P, (1, 1, 1), 0, fix_borders=True)
    assert np.array_equal(fields["dbf"], ref["dbf"]) and np.array_equal(fields["daf"], ref["daf"])
    assert np.array_equal(fields["active"], ref["active"])


def _arena_case(name):
    """Volumes whose soma goes to a private arena: the staircase alone; with a label 2 wholly inside
    the cavity as well (it stays, nothing is evicted); with a 2-voxel component of the soma's own
    label inside the cavity (its group joins the arena's skeleton, shared vertices keep the radius of
    the lower component id); padded with background and a far label (the arena's box is not the
    volume)."""
    labels = shapes.soma_cavity(staircase=True, nucleus=name == "nucleus")
    c = labels.shape[0] // 2
    if name == "blob":
        labels[c + 4, c - 1, c - 1:c + 1] = 1
    if name == "padded":
        labels = np.pad(labels, ((5, 9), (5, 3), (5, 6)))
        labels[-3:, -2:, -4:] = 7
    return np.asfortranarray(labels)


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
@pytest.mark.parametrize("fix_borders", [True, False])
@pytest.mark.parametrize("case", ["staircase", "nucleus", "blob", "padded"])
def test_gpu_soma_private_arena_equals_oracle(brook, fix_branching, fix_borders, case):
    labels = _arena_case(case)
    aniso = (2, 3, 5) if case == "padded" else (1, 1, 1)
    got = brook.skeletonize(labels, P, aniso, dust_threshold=0, fix_branching=fix_branching, fix_borders=fix_borders,
                            output="packed", progress=False)
    want = oracle.skeletonize(labels, P, aniso, 0, fix_branching=fix_branching, fix_borders=fix_borders)
    assert 3 in want["labels"].tolist()
    _same(got, want, (case, fix_branching, fix_borders))
    fields = brook.prepare_fields(labels, teasar_params=P, anisotropy=aniso, dust_threshold=0, fix_borders=fix_borders)
    ref = oracle.prepare(labels, P, aniso, 0, fix_borders=fix_borders)
    assert np.array_equal(fields["active"], ref["active"]) and np.array_equal(fields["dbf"], ref["dbf"])


@pytest.mark.gpu
def test_gpu_soma_private_arena_batch_targets_and_graph(brook):
    """Arenas inside a batch share its launches; manual targets on the soma move into the arena; a
    voxel graph is cropped to the arena's box (the labels' graph, and one cutting a process off)."""
    vols = [_arena_case("staircase"), shapes.soma_cells(), _arena_case("padded")]
    batch = brook.skeletonize_batch(vols, teasar_params=P, dust_threshold=0, fix_borders=True, output="packed",
                                    progress=False)
    wants = [oracle.skeletonize(v, P, (1, 1, 1), 0, fix_borders=True, threads=4) for v in vols]
    for i in range(len(vols)):
        _same(batch[i], wants[i], i)
    # nine samples (three with arenas): the per-sample stages run on the engine's worker threads
    big = brook.skeletonize_batch(vols * 3, teasar_params=P, dust_threshold=0, fix_borders=True, output="packed",
                                  progress=False)
    for i in range(len(big)):
        _same(big[i], wants[i % len(vols)], ("nine", i))
    labels = _arena_case("padded")
    before, after = [[29, 29, 6]], [[29, 6, 29], [29, 29, 50]]
    assert all(labels[tuple(p)] == 1 for p in before + after)
    got = brook.skeletonize(labels, P, (1, 1, 1), dust_threshold=0, fix_borders=False, extra_targets_before=before,
                            extra_targets_after=after, output="packed", progress=False)
    want = oracle.skeletonize(labels, P, (1, 1, 1), 0, fix_borders=False, extra_targets_before=before,
                              extra_targets_after=after)
    _same(got, want, "targets")
    from tests.golden import graphs
    c = labels.shape[1] // 2
    side = np.zeros(labels.shape, bool)
    side[:, c + 14:, :] = True
    for cut in (False, True):
        graph = np.asfortranarray(graphs.label_graph(labels, side if cut else np.zeros(labels.shape, bool)))
        got = brook.skeletonize(labels, P, (1, 1, 1), dust_threshold=0, fix_borders=True, voxel_graph=graph,
                                output="packed", progress=False)
        want = oracle.skeletonize(labels, P, (1, 1, 1), 0, fix_borders=True, voxel_graph=graph)
        _same(got, want, ("graph", cut))
