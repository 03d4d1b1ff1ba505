"""Voxel graphs (the `voxel_graph` option; graph_components.cu, graph_edt.cu, the graph gates of the
floods, railroads, backtraces and invalidations): the oracle pinned by the reference's known answers
(c_smoke.c:37-50), its Kimimaro 5.8.1 answers (tests/golden/graphs.py), brute force and the plain
path; then the GPU against the oracle through the C ABI."""
import itertools

import numpy as np
import pytest

import oracle
from tests.golden import graphs, shapes


def _summary(packed):
    return {k: tuple(v) for k, v in shapes.summary_from_packed(packed).items()}


def _full_graph(labels):
    """Every step between equal nonzero labels (no cut)."""
    return graphs.voxel_graph(labels, np.zeros(labels.shape, bool), np.zeros(labels.shape, bool))


# ---- oracle pins ---------------------------------------------------------------------------------

def test_graph_ccl_known_answer():
    # c_smoke.c:37-50: an all-zero graph isolates every voxel; the foreground voxels are components
    # 1..4 in root order, numbered without gaps for the background
    labels = np.array([1, 1, 0, 0, 0, 0, 2, 2], np.uint8)[:, None, None]
    cc, mapping, roots, keys, uniform = oracle.graph_ccl(labels, np.zeros(labels.shape, np.uint32))
    assert cc[:, 0, 0].tolist() == [1, 2, 0, 0, 0, 0, 3, 4]
    assert mapping.tolist() == [0, 1, 1, 2, 2] and roots.tolist() == [0, 1, 6, 7]
    assert keys.tolist() == [0, 1, 2, 7, 8] and not uniform
    edt = oracle.graph_edt(cc, np.zeros(labels.shape, np.uint32))
    assert edt[:, 0, 0].tolist() == [0.5, 0.5, 0, 0, 0, 0, 0.5, 0.5]


@pytest.mark.parametrize("seed", range(5))
def test_graph_ccl_of_the_labels_own_graph_is_ccl(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(x) for x in rng.integers(4, 14, 3))
    labels = (rng.integers(0, 4, shape) * (rng.random(shape) < 0.5)).astype(np.uint32)
    cc, mapping, roots = oracle.ccl(labels)
    gcc, gmapping, groots, keys, _ = oracle.graph_ccl(labels, _full_graph(labels))
    assert np.array_equal(cc, gcc) and np.array_equal(mapping, gmapping) and np.array_equal(roots, groots)
    # keys: each component's rank among all of them by lowest voxel; without steps from the
    # background every background voxel is a component of its own
    bg_before = np.concatenate([[0], np.cumsum(labels.ravel(order="F") == 0)])
    assert keys.tolist() == [0] + [int(bg_before[r] + k + 1) for k, r in enumerate(roots)]


def test_graph_ccl_joins_labels_the_graph_joins():
    labels, graph = graphs.spanning_labels()
    cc, mapping, roots, keys, _ = oracle.graph_ccl(labels, graph)
    assert len(roots) == 1 and mapping.tolist() == [0, 5]  # the label where the last run starts
    assert roots[0] == np.flatnonzero(labels.ravel(order="F"))[0]
    assert keys.tolist() == [0, 1]  # the open graph joins the background too: one component in all


def _graph_edt_brute(labels, graph, aniso, black_border):
    """Distance from each voxel centre to the nearest background position of the doubled grid."""
    s = labels.shape
    d = tuple(2 * n for n in s)
    mask = np.zeros(d, bool)
    for X, Y, Z in itertools.product(*(range(n) for n in d)):
        if black_border and (X == d[0] - 1 or Y == d[1] - 1 or Z == d[2] - 1):
            continue
        src = (X // 2, Y // 2, Z // 2)
        if not labels[src]:
            continue
        sub = (X & 1) | ((Y & 1) << 1) | ((Z & 1) << 2)
        bit = {1: 0, 2: 2, 4: 4}.get(sub)
        mask[X, Y, Z] = True if bit is None else bool((int(graph[src]) >> bit) & 1)
    zeros = np.argwhere(~mask).astype(np.float64)
    out = np.zeros(s)
    half = np.array(aniso, np.float64) / 2
    for p in itertools.product(*(range(n) for n in s)):
        if not labels[p]:
            continue
        c = 2 * np.array(p, np.float64)
        best = np.inf
        if len(zeros):
            best = np.min(np.sum(((zeros - c) * half) ** 2, axis=1))
        if black_border:  # outside the volume along any axis, one doubled position beyond each face
            for a in range(3):
                best = min(best, ((c[a] + 1) * half[a]) ** 2, ((d[a] - c[a]) * half[a]) ** 2)
        out[p] = np.sqrt(best)
    return out


@pytest.mark.parametrize("seed,black_border", [(0, False), (1, True), (2, False), (3, True)])
def test_graph_edt_is_the_doubled_grid_distance(seed, black_border):
    rng = np.random.default_rng(40 + seed)
    shape = tuple(int(x) for x in rng.integers(3, 7, 3))
    labels = (rng.random(shape) < 0.7).astype(np.uint32) * rng.integers(1, 3, shape).astype(np.uint32)
    graph = rng.integers(0, 1 << 26, shape).astype(np.uint32)
    aniso = [(1., 1., 1.), (4., 4., 40.), (1., 1.13, 1.37)][seed % 3]
    got = oracle.graph_edt(labels, graph, aniso, black_border)
    want = _graph_edt_brute(labels, graph, aniso, black_border)
    fg = labels > 0
    assert np.allclose(got[fg], want[fg], rtol=1e-6, atol=0) and not got[~fg].any()


@pytest.mark.parametrize("fix_branching", [True, False])
def test_oracle_spanning_labels_equal_kimimaro(fix_branching):
    labels, graph = graphs.spanning_labels()
    for extra in (dict(), dict(fill_holes=True, fix_avocados=True)):
        got = oracle.skeletonize(labels, graphs.HOLE_PARAMS, dust_threshold=10, fix_branching=fix_branching,
                                 voxel_graph=graph, **extra)
        assert _summary(got) == graphs.SPANNING


@pytest.mark.parametrize("fix_branching", [False, True])
def test_oracle_cut_tube_fill_holes_equals_kimimaro(fix_branching):
    labels, gs = graphs.cut_tube_graphs()
    for name, graph in gs.items():
        for i, aniso in enumerate([(1., 1., 1.), (1., 1.13, 1.37)]):
            got = _summary(oracle.skeletonize(labels, graphs.HOLE_PARAMS, aniso, 10, fix_branching=fix_branching,
                                              fill_holes=True, voxel_graph=graph))
            assert list(got) == [5] and got[5][0] - got[5][1] == 2, (name, aniso)
            assert got[5][2] == graphs.CUT_TUBE_RADIUS[name][i], (name, aniso)
            if not fix_branching and (name, aniso) in graphs.CUT_TUBE_DIGESTS:
                assert got[5] == graphs.CUT_TUBE_DIGESTS[name, aniso], (name, aniso)


def test_oracle_avocados_with_graph_equal_kimimaro():
    labels = graphs.avocado(graphs.AVOCADO_JITTER)
    graph = np.asfortranarray(graphs.label_graph(labels, np.zeros(labels.shape, bool)))
    params = dict(graphs.HOLE_PARAMS, soma_detection_threshold=10, soma_acceptance_threshold=11)
    plain = oracle.skeletonize(labels, params, dust_threshold=10, fix_branching=False, voxel_graph=graph)
    assert _summary(plain) == graphs.AVOCADO_PLAIN
    for fill_holes in (False, True):
        got = oracle.skeletonize(labels, params, dust_threshold=10, fix_branching=False, voxel_graph=graph,
                                 fill_holes=fill_holes, fix_avocados=True)
        assert _summary(got) == {3: graphs.AVOCADO_PLAIN[3]}
        branched = _summary(oracle.skeletonize(labels, params, dust_threshold=10, voxel_graph=graph,
                                               fill_holes=fill_holes, fix_avocados=True))
        assert list(branched) == [3] and branched[3][0] - branched[3][1] == 1 and branched[3][2] == 9.5


def test_oracle_avocado_order_with_graph_equals_kimimaro():
    labels, graph = graphs.avocado_order()
    aniso = (1., 1.13, 1.37)
    params = dict(graphs.HOLE_PARAMS, soma_detection_threshold=10, soma_acceptance_threshold=1000)
    plain = _summary(oracle.skeletonize(labels, params, aniso, 10, fix_branching=False, voxel_graph=graph))
    assert plain[8] == graphs.AVOCADO_ORDER_NUCLEUS and plain[3][:3] == graphs.AVOCADO_ORDER_SOMA
    corrected = oracle.skeletonize(labels, params, aniso, 10, fix_branching=False, voxel_graph=graph, fix_avocados=True)
    assert _summary(corrected) == graphs.AVOCADO_ORDER_CORRECTED
    branched = _summary(oracle.skeletonize(labels, params, aniso, 10, voxel_graph=graph, fix_avocados=True))
    assert list(branched) == [3] and branched[3][0] - branched[3][1] == 1
    assert branched[3][2] == graphs.AVOCADO_ORDER_BRANCHED_RADIUS
    assert oracle.graph_ccl(labels, graph)[3].tolist() == graphs.AVOCADO_ORDER_KEYS


@pytest.mark.parametrize("fix_branching", [False, True])
def test_oracle_soma_voids_with_graph_equal_kimimaro(fix_branching):
    labels, graph = graphs.soma_voids()
    params = dict(graphs.HOLE_PARAMS, soma_detection_threshold=3, soma_acceptance_threshold=4)
    for aniso, radius in graphs.SOMA_VOID_RADIUS.items():
        got = _summary(oracle.skeletonize(labels, params, aniso, 10, fix_branching=fix_branching, voxel_graph=graph))
        assert list(got) == [1] and got[1][0] - got[1][1] == 1 and got[1][2] == radius


@pytest.mark.parametrize("fix_branching", [False, True])
def test_oracle_painted_root_with_graph_equals_kimimaro(fix_branching):
    labels = shapes.painted_root()
    graph = np.asfortranarray(graphs.label_graph(labels, np.zeros(labels.shape, bool)))
    for fix_avocados in (False, True):
        got = _summary(oracle.skeletonize(labels, graphs.HOLE_PARAMS, dust_threshold=10, fix_borders=False,
                                          fix_branching=fix_branching, fill_holes=True, fix_avocados=fix_avocados,
                                          voxel_graph=graph))
        assert list(got) == [2, 1] and all(v[0] - v[1] == 1 for v in got.values())
        assert {label: v[2] for label, v in got.items()} == graphs.PAINTED_ROOT_GRAPH_RADIUS


# ---- GPU against the oracle -------------------------------------------------------------------------

@pytest.fixture(scope="module")
def brook():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import brook_amd
    return brook_amd


def _dev(a):
    import torch
    a = np.asarray(a)
    t = torch.from_numpy(np.ravel(a, order="F").copy()).cuda()
    return t.view(a.shape[2], a.shape[1], a.shape[0]).permute(2, 1, 0)


def _host(t):
    return t.permute(2, 1, 0).contiguous().cpu().numpy().transpose(2, 1, 0)


def _assert_packed_equal(got, want, what=""):
    g = got.as_dict() if hasattr(got, "as_dict") else got
    for name in ("labels", "v_off", "e_off", "vertices", "edges", "radii"):
        assert np.array_equal(np.asarray(g[name]), np.asarray(want[name])), (what, name)


def _random_graph_case(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(x) for x in rng.integers(8, 30, 3))
    labels = (rng.integers(0, 5, shape) * (rng.random(shape) < 0.55)).astype(np.uint32)
    graph = (rng.integers(0, 1 << 26, shape) & rng.integers(0, 1 << 26, shape)).astype(np.uint32)
    return np.asfortranarray(labels), np.asfortranarray(graph)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_gpu_graph_components_equal_oracle(brook, seed):
    from brook_amd import device
    labels, graph = _random_graph_case(seed)
    for lab, g in ((labels, graph), (labels, _full_graph(labels)), (labels.astype(np.uint16), graph)):
        ids, mapping, roots, keys = device.graph_components(lab, g)
        cc, m, r, k, _ = oracle.graph_ccl(lab, g)
        assert np.array_equal(_host(ids).astype(np.uint32), cc)
        assert np.array_equal(mapping, m) and np.array_equal(roots, r) and np.array_equal(keys, k)
    plane = (np.random.default_rng(seed).integers(0, 3, (33, 27, 1))).astype(np.uint32)
    words = np.random.default_rng(seed + 9).integers(0, 256, (33, 27, 1)).astype(np.uint8)
    ids, mapping, roots, keys = device.graph_components(plane, words)  # a uint8 plane graph
    cc, m, r, k, _ = oracle.graph_ccl(plane, words)
    assert np.array_equal(_host(ids).astype(np.uint32), cc) and np.array_equal(keys, k)


@pytest.mark.gpu
def test_gpu_graph_known_answer(brook):
    from brook_amd import device
    labels = np.array([1, 1, 0, 0, 0, 0, 2, 2], np.uint8)[:, None, None]
    ids, mapping, roots, keys = device.graph_components(labels, np.zeros(labels.shape, np.uint32))
    assert _host(ids)[:, 0, 0].tolist() == [1, 2, 0, 0, 0, 0, 3, 4]
    assert mapping.tolist() == [0, 1, 1, 2, 2] and roots.tolist() == [0, 1, 6, 7]
    d = device.graph_edt(labels, np.zeros(labels.shape, np.uint32))
    assert _host(d)[:, 0, 0].tolist() == [0.5, 0.5, 0, 0, 0, 0, 0.5, 0.5]


@pytest.mark.gpu
@pytest.mark.parametrize("seed,black_border", [(0, False), (1, True), (2, False), (3, True)])
def test_gpu_graph_edt_equals_oracle(brook, seed, black_border):
    from brook_amd import device
    labels, graph = _random_graph_case(10 + seed)
    aniso = [(1., 1., 1.), (4., 4., 40.), (1., 1.13, 1.37), (16., 16., 40.)][seed]
    got = _host(device.graph_edt(labels, graph, aniso, black_border))
    want = oracle.graph_edt(labels, graph, aniso, black_border)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
def test_gpu_graph_cases_equal_oracle(brook, fix_branching):
    for i, (labels, side, region) in enumerate(graphs.graph_cases()):
        graph = np.asfortranarray(graphs.voxel_graph(labels, side, region))
        labels = np.asfortranarray(labels)
        opts = dict(graphs.GRAPH_OPTIONS, fix_branching=fix_branching)
        want = oracle.skeletonize(labels, opts["teasar_params"], dust_threshold=opts["dust_threshold"],
                                  fix_branching=fix_branching, voxel_graph=graph)
        got = brook.skeletonize(labels, voxel_graph=graph, output="packed", progress=False, **opts)
        _assert_packed_equal(got, want, f"case {i}")
        device_got = brook.skeletonize(_dev(labels), voxel_graph=_dev(graph), output="packed", progress=False, **opts)
        _assert_packed_equal(device_got, want, f"case {i} device")
        plain = oracle.skeletonize(labels, opts["teasar_params"], dust_threshold=opts["dust_threshold"],
                                   fix_branching=fix_branching)
        if i == 0:  # the cut changes the skeleton
            assert not np.array_equal(plain["vertices"], want["vertices"]) or len(plain["labels"]) != len(want["labels"])


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
def test_gpu_kimimaro_graph_answers(brook, fix_branching):
    labels, graph = graphs.spanning_labels()
    for extra in (dict(), dict(fill_holes=True, fix_avocados=True)):
        for lab, g in ((labels, graph), (np.ascontiguousarray(labels), np.ascontiguousarray(graph)),
                       (_dev(labels), _dev(graph)), (_dev(labels), graph)):
            got = brook.skeletonize(lab, graphs.HOLE_PARAMS, dust_threshold=10, fix_branching=fix_branching,
                                    voxel_graph=g, output="packed", progress=False, **extra)
            assert _summary(got.as_dict()) == graphs.SPANNING
    tube, gs = graphs.cut_tube_graphs()
    for name, g in gs.items():
        for i, aniso in enumerate([(1., 1., 1.), (1., 1.13, 1.37)]):
            got = _summary(brook.skeletonize(tube, graphs.HOLE_PARAMS, aniso, dust_threshold=10, fix_branching=fix_branching,
                                             fill_holes=True, voxel_graph=g, output="packed", progress=False).as_dict())
            assert list(got) == [5] and got[5][0] - got[5][1] == 2 and got[5][2] == graphs.CUT_TUBE_RADIUS[name][i]
            if not fix_branching and (name, aniso) in graphs.CUT_TUBE_DIGESTS:
                assert got[5] == graphs.CUT_TUBE_DIGESTS[name, aniso]
    av = graphs.avocado(graphs.AVOCADO_JITTER)
    ag = np.asfortranarray(graphs.label_graph(av, np.zeros(av.shape, bool)))
    params = dict(graphs.HOLE_PARAMS, soma_detection_threshold=10, soma_acceptance_threshold=11)
    if not fix_branching:
        got = brook.skeletonize(av, params, dust_threshold=10, fix_branching=False, voxel_graph=ag, output="packed",
                                progress=False)
        assert _summary(got.as_dict()) == graphs.AVOCADO_PLAIN
    got = _summary(brook.skeletonize(av, params, dust_threshold=10, fix_branching=fix_branching, voxel_graph=ag,
                                     fix_avocados=True, output="packed", progress=False).as_dict())
    assert list(got) == [3] and got[3][0] - got[3][1] == 1 and got[3][2] == 9.5
    if not fix_branching:
        assert got == {3: graphs.AVOCADO_PLAIN[3]}
    lab, og = graphs.avocado_order()
    params = dict(graphs.HOLE_PARAMS, soma_detection_threshold=10, soma_acceptance_threshold=1000)
    got = _summary(brook.skeletonize(lab, params, (1., 1.13, 1.37), dust_threshold=10, fix_branching=fix_branching,
                                     voxel_graph=og, fix_avocados=True, output="packed", progress=False).as_dict())
    if fix_branching:
        assert list(got) == [3] and got[3][2] == graphs.AVOCADO_ORDER_BRANCHED_RADIUS
    else:
        assert got == graphs.AVOCADO_ORDER_CORRECTED
    sv, sg = graphs.soma_voids()
    params = dict(graphs.HOLE_PARAMS, soma_detection_threshold=3, soma_acceptance_threshold=4)
    for aniso, radius in graphs.SOMA_VOID_RADIUS.items():
        got = _summary(brook.skeletonize(sv, params, aniso, dust_threshold=10, fix_branching=fix_branching,
                                         voxel_graph=sg, output="packed", progress=False).as_dict())
        assert list(got) == [1] and got[1][0] - got[1][1] == 1 and got[1][2] == radius


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
def test_gpu_batch_with_graphs_equals_separate_calls(brook, fix_branching):
    cases = graphs.graph_cases()
    samples = [np.asfortranarray(c[0]) for c in cases]
    gs = [np.asfortranarray(graphs.voxel_graph(*c)) for c in cases]
    opts = dict(graphs.GRAPH_OPTIONS, fix_branching=fix_branching)
    single = [brook.skeletonize(s, voxel_graph=g, output="packed", progress=False, **opts) for s, g in zip(samples, gs)]
    batch = brook.skeletonize_batch(samples, voxel_graph=gs, output="packed", **opts)
    for i in range(len(samples)):
        _assert_packed_equal(batch[i], single[i].as_dict(), f"sample {i}")
    reps = -(-8 // len(samples))  # at least eight samples: the per-sample stages run on worker threads
    big = brook.skeletonize_batch(samples * reps, voxel_graph=gs * reps, output="packed", **opts)
    for i in range(len(big)):
        _assert_packed_equal(big[i], single[i % len(samples)].as_dict(), f"repeated sample {i}")
    mixed = brook.skeletonize_batch(samples, voxel_graph=[gs[0], None, gs[2]], output="packed", **opts)
    plain = brook.skeletonize(samples[1], output="packed", progress=False, **opts)
    _assert_packed_equal(mixed[1], plain.as_dict(), "no graph")
    _assert_packed_equal(mixed[2], single[2].as_dict(), "graph after a sample without one")
    with pytest.raises(ValueError):
        brook.skeletonize_batch(samples, voxel_graph=gs[:2], **opts)


@pytest.mark.gpu
def test_gpu_graph_argument_errors(brook):
    labels, graph = graphs.spanning_labels()
    with pytest.raises(TypeError):  # a device graph needs device labels
        brook.skeletonize(labels, voxel_graph=_dev(graph), dust_threshold=10, progress=False)
    with pytest.raises(ValueError):  # shape
        brook.skeletonize(labels, voxel_graph=graph[:-1], dust_threshold=10, progress=False)
    with pytest.raises(ValueError):  # a uint8 graph of a volume, not a plane
        brook.skeletonize(labels, voxel_graph=graph.astype(np.uint8), dust_threshold=10, progress=False)
