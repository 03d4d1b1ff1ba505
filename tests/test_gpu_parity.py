"""GPU parity: the HIP path against the oracle and the reference's golden vectors, through the
C ABI. Integer work (components, paths, vertex/edge sets) must be bit-exact; EDT values are
float32 and must equal the oracle's restatement of the reference's rounding rule bit for bit,
and lie within 1e-5 relative of the float64 EDT."""
import json
import os

import numpy as np
import pytest

import oracle
from tests.golden import shapes

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden_digests.json")))
CFG1_PARAMS = dict(scale=1.5, const=300, pdrf_scale=100000, pdrf_exponent=4)


@pytest.fixture(scope="module")
def brook():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import brook_amd
    return brook_amd


def _assert_packed_equal(got, want, what=""):
    g, w = got.as_dict() if hasattr(got, "as_dict") else got, want
    assert list(g["labels"]) == list(w["labels"]), what
    for name in ("v_off", "e_off"):
        assert np.array_equal(np.asarray(g[name]), np.asarray(w[name])), (what, name)
    for name in ("vertices", "edges", "radii"):
        assert np.array_equal(np.asarray(g[name]), np.asarray(w[name])), (what, name)


@pytest.mark.parametrize("name", [c[0] for c in shapes.golden_cases()])
def test_golden_digests(brook, name):
    import torch
    case = {c[0]: c for c in shapes.golden_cases()}[name]
    options = dict(shapes.GOLDEN_OPTIONS, **case[2])
    expected = {int(k): v for k, v in GOLDEN[name].items()}
    for layout in (np.ascontiguousarray, np.asfortranarray):
        packed = brook.skeletonize(layout(case[1]), output="packed", progress=False, **options)
        assert shapes.summary_from_packed(packed.as_dict()) == expected, layout.__name__
    dev = torch.from_numpy(np.ascontiguousarray(case[1])).cuda()  # C-ordered device input (gathered on the GPU)
    packed = brook.skeletonize(dev, output="packed", progress=False, **options)
    assert shapes.summary_from_packed(packed.as_dict()) == expected, "device"


def test_ccl_known_answer(brook):
    import ctypes
    from brook_amd import _native as native
    from brook_amd._intake import volume_struct
    labels = np.array([1, 1, 0, 0, 0, 0, 2, 2], np.uint8)[:, None, None]
    keep = []
    vol, _ = volume_struct(labels, keep)
    h = ctypes.c_void_p()
    native.check(native.lib().brook_connected_components(native.context().handle, ctypes.byref(vol), ctypes.byref(h)))
    mapping, roots, count = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_size_t()
    native.check(native.lib().brook_components_view(h, None, ctypes.byref(mapping), ctypes.byref(roots), ctypes.byref(count)))
    m = np.ctypeslib.as_array(ctypes.cast(mapping, ctypes.POINTER(ctypes.c_int64)), (count.value + 1,)).copy()
    r = np.ctypeslib.as_array(ctypes.cast(roots, ctypes.POINTER(ctypes.c_int64)), (count.value,)).copy()
    out = np.zeros(8, np.uint8)
    native.check(native.lib().brook_components_copy_to_host(h, out.ctypes.data, 8))
    native.lib().brook_components_destroy(h)
    assert count.value == 2 and list(m) == [0, 1, 2] and list(r) == [0, 6]
    assert out.tolist() == [1, 1, 0, 0, 0, 0, 2, 2]


def _gpu_ccl(labels):
    import ctypes
    from brook_amd import _native as native
    from brook_amd._intake import volume_struct
    keep = []
    vol, _ = volume_struct(labels, keep)
    h = ctypes.c_void_p()
    native.check(native.lib().brook_connected_components(native.context().handle, ctypes.byref(vol), ctypes.byref(h)))
    view = native.Volume()
    mapping, roots, count = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_size_t()
    native.check(native.lib().brook_components_view(h, ctypes.byref(view), ctypes.byref(mapping), ctypes.byref(roots),
                                                    ctypes.byref(count)))
    dt = {0: np.uint8, 1: np.uint16, 2: np.uint32, 3: np.uint64}[view.dtype]
    out = np.zeros(labels.shape, dt, order="F")
    native.check(native.lib().brook_components_copy_to_host(h, out.ctypes.data, out.nbytes))
    n = count.value
    m = np.ctypeslib.as_array(ctypes.cast(mapping, ctypes.POINTER(ctypes.c_int64)), (n + 1,)).copy()
    r = np.ctypeslib.as_array(ctypes.cast(roots, ctypes.POINTER(ctypes.c_int64)), (max(n, 1),))[:n].copy()
    native.lib().brook_components_destroy(h)
    return out, m, r


@pytest.mark.parametrize("seed,shape,nl", [(0, (37, 29, 23), 3), (1, (64, 64, 64), 7), (2, (130, 70, 9), 300),
                                           (3, (20, 1, 33), 2)])
def test_ccl_matches_oracle(brook, seed, shape, nl):
    rng = np.random.default_rng(seed)
    labels = np.asfortranarray(rng.integers(0, nl + 1, size=shape).astype(np.uint32) * (rng.random(shape) < 0.7))
    got, gm, gr = _gpu_ccl(labels)
    cc, m, r = oracle.ccl(labels)
    assert np.array_equal(got.astype(np.uint32), cc) and np.array_equal(gm, m) and np.array_equal(gr, r)
    assert got.dtype == (np.uint8 if len(r) <= 255 else np.uint16 if len(r) <= 65535 else np.uint32)


def test_ccl_synthetic_tubes(brook):
    from brook_amd import synth
    labels = synth.tubes_128()
    got, gm, gr = _gpu_ccl(labels)
    cc, m, r = oracle.ccl(labels)
    assert np.array_equal(got.astype(np.uint32), cc) and np.array_equal(gm, m) and np.array_equal(gr, r)


def _gpu_edt(labels, aniso, bb):
    import ctypes
    from brook_amd import _native as native
    from brook_amd._intake import volume_struct
    keep = []
    vol, _ = volume_struct(labels, keep)
    an = np.asarray(aniso, np.float32)
    h = ctypes.c_void_p()
    native.check(native.lib().brook_edt(native.context().handle, ctypes.byref(vol), an.ctypes.data, int(bb), ctypes.byref(h)))
    out = np.zeros(labels.shape, np.float32, order="F")
    native.check(native.lib().brook_array_copy_to_host(h, out.ctypes.data, out.nbytes))
    native.lib().brook_array_destroy(h)
    return out


@pytest.mark.parametrize("seed,aniso,bb", [(0, (1, 1, 1), False), (1, (16, 16, 40), False), (2, (4.5, 7.91, 40), True),
                                           (3, (8, 8, 40), False)])
def test_edt_bit_exact_and_tolerance(brook, seed, aniso, bb):
    rng = np.random.default_rng(seed)
    shape = (70, 50, 40)
    labels = np.zeros(shape, np.uint32)
    x, y, z = np.ogrid[:shape[0], :shape[1], :shape[2]]
    for lab in range(1, 9):
        c = rng.integers(0, 50, 3)
        r = rng.integers(3, 15)
        labels[((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r) & (labels == 0)] = lab
    labels = np.asfortranarray(labels)
    got = _gpu_edt(labels, aniso, bb)
    cc, _, _ = oracle.ccl(labels)
    want = oracle.edt(cc, aniso, bb)
    assert np.array_equal(got, want)
    exact = oracle.edt_f64(cc, aniso, bb)
    fg = labels > 0
    rel = np.abs(got[fg].astype(np.float64) - exact[fg]) / exact[fg]
    assert rel.max() <= 1e-5, (rel.max(), rel.mean())


def test_edt_known_answer(brook):
    labels = np.array([1, 1, 0, 0, 0, 0, 2, 2], np.uint8)[:, None, None]
    assert _gpu_edt(labels, (1, 1, 1), False).ravel().tolist() == [2, 1, 0, 0, 0, 0, 1, 2]


def test_fields_match_oracle(brook):
    from brook_amd import synth
    labels = synth.tubes_128()
    aniso = (16, 16, 40)
    got = brook.prepare_fields(labels, CFG1_PARAMS, aniso, 1000)
    want = oracle.prepare(labels, CFG1_PARAMS, aniso, 1000)
    for name in ("dbf", "daf", "pdrf"):
        assert np.array_equal(got[name], want[name]), name
    act = want["active"].astype(bool)
    assert np.array_equal(got["active"], want["active"])
    for name in ("root", "target", "dbf_max", "max_daf"):
        assert np.array_equal(got[name][act], want[name][act]), name


@pytest.mark.parametrize("seed", [1001, 1002])
def test_skeletons_match_oracle_tubes(brook, seed):
    from brook_amd import synth
    labels = synth.tubes_128(seed)
    aniso = (16, 16, 40)
    got = brook.skeletonize(labels, CFG1_PARAMS, aniso, output="packed", progress=False)
    want = oracle.skeletonize(labels, CFG1_PARAMS, aniso, 1000, threads=8)
    _assert_packed_equal(got, want, f"seed {seed}")


@pytest.mark.parametrize("seed", [1001, 1002])
def test_skeletons_match_oracle_tubes_parent_paths(brook, seed):
    """fix_branching=False: one vertex-weighted field and parent walks with the STOP rule."""
    from brook_amd import synth
    labels = synth.tubes_128(seed)
    aniso = (16, 16, 40)
    got = brook.skeletonize(labels, CFG1_PARAMS, aniso, fix_branching=False, output="packed", progress=False)
    want = oracle.skeletonize(labels, CFG1_PARAMS, aniso, 1000, fix_branching=False, threads=8)
    _assert_packed_equal(got, want, f"seed {seed}")


def test_batch_parent_paths_equal_oracle(brook):
    """fix_branching=False in the batched engine (one parent-field launch for every sample)."""
    from brook_amd import synth
    vols = [synth.volume("tubes", s, n, 170 + i) for i, (s, n) in enumerate([((64, 80, 72), 6), ((96, 64, 64), 9)])]
    opts = dict(teasar_params=CFG1_PARAMS, anisotropy=(16, 16, 40), dust_threshold=500)
    batch = brook.skeletonize_batch(vols, output="packed", fix_branching=False, **opts)
    for i, v in enumerate(vols):
        want = oracle.skeletonize(v, CFG1_PARAMS, (16, 16, 40), 500, fix_branching=False, threads=8)
        _assert_packed_equal(batch[i], want, f"sample {i}")


def test_skeletons_match_oracle_cells(brook):
    """The stress workload's shape at a small size: uint64 ids above 2^32 (cell << 32 | tube), a few
    hundred small objects clipped by eroded Voronoi cells, anisotropy (4, 4, 40)."""
    from brook_amd import synth
    labels, cells, _ = synth.cells((192, 192, 96), 120, 4004)
    assert labels.dtype == np.uint64 and cells > 100
    aniso = (4, 4, 40)
    got = brook.skeletonize(labels, CFG1_PARAMS, aniso, dust_threshold=100, output="packed", progress=False)
    want = oracle.skeletonize(labels, CFG1_PARAMS, aniso, 100, threads=8)
    assert len(want["labels"]) > 200
    _assert_packed_equal(got, want, "cells")


def test_batch_equals_separate(brook):
    from brook_amd import synth
    vols = [synth.volume("tubes", s, n, 70 + i) for i, (s, n) in enumerate([((64, 80, 72), 6), ((96, 64, 64), 9), ((64, 64, 100), 4)])]
    opts = dict(teasar_params=CFG1_PARAMS, anisotropy=(16, 16, 40), dust_threshold=500)
    batch = brook.skeletonize_batch(vols, output="packed", **opts)
    assert len(batch) == 3
    for i, v in enumerate(vols):
        single = brook.skeletonize(v, output="packed", progress=False, **opts)
        _assert_packed_equal(batch[i], single.as_dict(), f"sample {i}")


def test_served_results_equal_plain_calls(brook):
    """Continuous batching: requests that share a batch give the results of plain calls."""
    from brook_amd import synth
    shapes = [((64, 80, 72), 6), ((96, 64, 64), 9), ((64, 64, 100), 4), ((72, 72, 72), 7), ((80, 64, 96), 5)]
    vols = [synth.volume("tubes", s, n, 90 + i) for i, (s, n) in enumerate(shapes)]
    opts = dict(teasar_params=CFG1_PARAMS, anisotropy=(16, 16, 40), dust_threshold=500)
    plain = [brook.skeletonize(v, output="packed", progress=False, **opts) for v in vols]
    with brook.Server(**opts) as srv:
        futures = [srv.submit(v) for v in vols]
        got = [f.result(120) for f in futures]
    assert sum(srv.batches) == len(vols) and max(srv.batches) > 1, srv.batches
    for i, (g, p) in enumerate(zip(got, plain)):
        _assert_packed_equal(g, p.as_dict(), f"request {i} (batches {srv.batches})")


def test_edge_cases(brook):
    # empty, all background, volume at or below the dust threshold, a single voxel label
    assert brook.skeletonize(np.zeros((0, 4, 4), np.uint32), progress=False) == {}
    assert brook.skeletonize(np.zeros((20, 20, 20), np.uint32), progress=False, dust_threshold=0) == {}
    assert brook.skeletonize(np.ones((5, 5, 5), np.uint32), progress=False) == {}
    one = np.zeros((9, 9, 9), np.uint32)
    one[4, 4, 4] = 3
    assert brook.skeletonize(one, dust_threshold=0, progress=False) == {}
    # one label filling the whole volume: black border
    full = np.ones((24, 20, 18), np.uint16)
    got = brook.skeletonize(full, dict(scale=1, const=2), dust_threshold=0, output="packed", progress=False)
    want = oracle.skeletonize(full, dict(scale=1, const=2), dust_threshold=0)
    _assert_packed_equal(got, want, "full volume")


@pytest.mark.parametrize("shape", [(1, 40, 30), (40, 1, 30), (40, 30, 1), (2, 50, 20), (64, 24, 8), (30, 30, 30)])
def test_border_targets_thin_and_skewed(brook, shape):
    """fix_borders on volumes whose faces differ in size or coincide (a dimension of 1: the near and
    far faces are one plane), with an anisotropy that differs per axis: blocky labels touching every
    face, each face's border targets and the skeletons equal to the oracle's."""
    rng = np.random.default_rng(sum(shape))
    coarse = rng.integers(0, 6, size=tuple(max(1, s // 4 + 1) for s in shape)).astype(np.uint32)
    labels = np.kron(coarse, np.ones((4, 4, 4), np.uint32))[: shape[0], : shape[1], : shape[2]]
    labels = np.ascontiguousarray(labels)
    params = dict(scale=1.5, const=4)
    aniso = (4, 6, 10)
    got = brook.skeletonize(labels, params, aniso, dust_threshold=0, fix_borders=True, output="packed", progress=False)
    want = oracle.skeletonize(labels, params, aniso, 0, fix_borders=True)
    assert len(want["labels"]) > 0
    _assert_packed_equal(got, want, f"shape {shape}")


def test_unsupported_options_raise(brook):
    labels = np.zeros((20, 20, 20), np.uint32)
    labels[5:15, 5:15, 5:15] = 1
    with pytest.raises(NotImplementedError):
        brook.skeletonize(labels, streaming=True, dust_threshold=0, progress=False)
    with pytest.raises(ValueError):
        brook.skeletonize(labels, dict(pdrf_exponent=3), dust_threshold=0, progress=False)


@pytest.mark.parametrize("fix_branching", [True, False])
def test_full_workload_equals_oracle_digests(brook, fix_branching):
    """The 512^3 / 200-label tree workload, device-resident input, against the digests of the
    oracle's result (tests/golden/workload_digests.py wrote them)."""
    import torch
    from brook_amd import synth
    from tests.golden.workload_digests import PARAMS, digest
    want = json.load(open(os.path.join(HERE, "golden", "workload_digests.json")))
    want = want["fix_branching" if fix_branching else "parent_walks"]
    volume = synth.trees_512()
    flat = torch.from_numpy(np.ravel(volume, order="F")).to("cuda:0")
    device = flat.view(volume.shape[2], volume.shape[1], volume.shape[0]).permute(2, 1, 0)
    got = brook.skeletonize(device, PARAMS, (8, 8, 40), dust_threshold=1000, fix_branching=fix_branching,
                            output="packed", progress=False).as_dict()
    got = digest(got)
    for name, value in got.items():
        assert value == want[name], name


def test_device_output_views(brook):
    """BROOK_DEVICE_OUTPUT through the C ABI: the device view's arrays and get_array hold the host
    view's bytes; a result requested without the flag refuses both."""
    import ctypes
    from brook_amd import _intake, _native as native, synth
    labels = synth.tubes_128(1001)
    L = native.lib()
    ctx = native.context()
    for flags in (native.DEVICE_OUTPUT, 0):
        keep = []
        vol, _ = _intake.volume_struct(labels, keep)
        opts = _intake.options_struct(CFG1_PARAMS, (16, 16, 40), 1000, True, True, False, False, None, (), (), keep)
        opts.flags |= flags
        h = ctypes.c_void_p()
        native.check(L.brook_skeletonize(ctx.handle, ctypes.byref(vol), ctypes.byref(opts), ctypes.byref(h)))
        try:
            host = native.SkeletonsView()
            native.check(L.brook_skeletons_get_view(h, ctypes.byref(host)))
            dev = native.SkeletonsDeviceView()
            status = L.brook_skeletons_get_device_view(h, ctypes.byref(dev))
            if not flags:
                assert status == native.INVALID_ARGUMENT
                arr = ctypes.c_void_p()
                assert L.brook_skeletons_get_array(h, native.VERTICES, ctypes.byref(arr)) == native.INVALID_ARGUMENT
                continue
            native.check(status)
            nv, ne = host.vertex_count, host.edge_count
            assert (dev.skeleton_count, dev.vertex_count, dev.edge_count) == (host.skeleton_count, nv, ne) and nv > 0
            # rows, as the reference's row_view (packed.cpp:7-14): (rows, width, 1), byte strides (row, item, bytes)
            assert dev.vertices.memory == native.DEVICE and tuple(dev.vertices.shape) == (nv, 3, 1)
            assert tuple(dev.vertices.strides) == (12, 4, 12 * nv)
            assert tuple(dev.edges.shape) == (ne, 2, 1) and tuple(dev.edges.strides) == (8, 4, 8 * ne)
            assert tuple(dev.radii.shape) == (nv, 1, 1) and tuple(dev.radii.strides) == (4, 4 * nv, 4 * nv)
            for which, ptr, dtype, count in ((native.VERTICES, host.vertices, np.float32, 3 * nv),
                                             (native.EDGES, host.edges, np.uint32, 2 * ne),
                                             (native.RADII, host.radii, np.float32, nv)):
                want = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                             (count,)).copy()
                arr = ctypes.c_void_p()
                native.check(L.brook_skeletons_get_array(h, which, ctypes.byref(arr)))
                got = np.zeros(count, dtype)
                try:
                    native.check(L.brook_array_copy_to_host(arr, got.ctypes.data, got.nbytes))
                    view = native.Volume()  # the array views its rows as the device view does
                    native.check(L.brook_array_view(arr, ctypes.byref(view)))
                    expected = {native.VERTICES: dev.vertices, native.EDGES: dev.edges, native.RADII: dev.radii}[which]
                    assert tuple(view.shape) == tuple(expected.shape) and tuple(view.strides) == tuple(expected.strides)
                finally:
                    L.brook_array_destroy(arr)
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), which
            frags, m = (ctypes.c_void_p * 2)(h, h), ctypes.c_void_p()  # merges return device-output handles
            native.check(L.brook_merge_skeletons(ctx.handle, frags, None, 2, ctypes.byref(m)))
            try:
                merged = native.SkeletonsDeviceView()
                native.check(L.brook_skeletons_get_device_view(m, ctypes.byref(merged)))
                assert merged.vertex_count == nv and merged.skeleton_count == host.skeleton_count
            finally:
                L.brook_skeletons_destroy(m)
        finally:
            L.brook_skeletons_destroy(h)


@pytest.mark.parametrize("workload", ["tubes_128", "trees_512"])
def test_workload_edt_equals_oracle_digest(brook, workload):
    """The GPU EDT of the 128^3 tube and 512^3 tree workloads equals the oracle's float32 DBF bit
    for bit (tests/golden/workload_digests.py wrote its digest), so the oracle's stated error
    against the float64 EDT (max ~6e-8 relative) is the GPU's; the bound asserted here is 1e-5."""
    import hashlib

    import torch
    from brook_amd import device, synth
    want = json.load(open(os.path.join(HERE, "golden", "workload_digests.json")))["edt"][workload]
    volume = getattr(synth, workload)()
    flat = torch.from_numpy(np.ravel(volume, order="F")).to("cuda:0")
    dev = flat.view(volume.shape[2], volume.shape[1], volume.shape[0]).permute(2, 1, 0)
    dbf = device.edt(dev, tuple(want["anisotropy"]))
    host = dbf.permute(2, 1, 0).contiguous().cpu().numpy()  # Fortran order of (x, y, z)
    assert hashlib.sha256(host.tobytes()).hexdigest() == want["sha256"]
    assert want["max_rel_err_f64"] <= 1e-5


def test_batch_workload_equals_oracle_digests(brook):
    """The 64-volume batch workload (synth.batch_64) in one skeletonize_batch call, device-resident
    samples: every sample's packed result against the oracle's digest (tests/golden/batch_digests.py
    wrote them)."""
    import torch
    from brook_amd import synth
    from tests.golden.batch_digests import ANISOTROPY, sample_digest
    from tests.golden.workload_digests import PARAMS
    want = json.load(open(os.path.join(HERE, "golden", "batch_digests.json")))
    vols = synth.batch_64()
    dev = []
    for v in vols:
        flat = torch.from_numpy(np.ravel(v, order="F")).to("cuda:0")
        dev.append(flat.view(v.shape[2], v.shape[1], v.shape[0]).permute(2, 1, 0))
    batch = brook.skeletonize_batch(dev, PARAMS, ANISOTROPY, dust_threshold=1000, output="packed")
    assert len(batch) == len(want) == 64
    for i in range(64):
        got = batch[i].as_dict()
        assert len(got["labels"]) == want[i]["skeletons"], i
        assert sample_digest(got) == want[i]["sha256"], i
