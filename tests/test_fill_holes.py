"""fill_holes (holes.cu:37-62, called at skeletonize.cu:83-90): in ascending component id and with
the original boxes, each component's box voids (voxels of other ids or background not reaching a
face of the box through face neighbours, fill_voids) become the component; a component painted
over, even in part, fills nothing later.

CPU tests pin the oracle's fill against scipy.ndimage.binary_fill_holes (an independent fill_voids
with face connectivity) applied in the same order, and its skeletons against those of the volume
filled beforehand (the DBF is the EDT of the filled volume in both). GPU tests compare brook_amd
with the oracle bit for bit.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi

import oracle
from tests.golden import shapes

P = dict(scale=1.5, const=2, pdrf_scale=100000, pdrf_exponent=4)


def hollow_cells(seed=0, size=64, split=False):
    """Shells with cavities (one holding a cube label, one a staircase label that leaves through
    the wall diagonally -- on two sides when `split`, so that filling splits it), an open cup,
    and tubes."""
    rng = np.random.default_rng(seed)
    out = np.zeros((size, size, size), np.uint32)
    grid = np.indices(out.shape)

    def ball(c, r):
        return ((grid - np.reshape(c, (3, 1, 1, 1))) ** 2).sum(axis=0) <= r * r

    out[ball((18, 18, 18), 11) & ~ball((18, 18, 18), 6)] = 5   # shell with a cavity
    out[16:21, 16:21, 16:21] = 7                                  # a cube inside the cavity
    out[ball((44, 40, 20), 10) & ~ball((44, 40, 20), 5)] = 2     # shell ...
    for i in range(10):                                            # ... and a staircase through its wall
        out[44 + i, 40 + i, 20] = 8
        if split:                                                  # the filled cavity and wall cut it in two
            out[44 - i, 40 - i, 20] = 8
    cup = ball((20, 46, 46), 9) & ~ball((20, 46, 46), 5)
    cup[20:, :, :] &= False                                        # open on one side
    out[cup] = 3
    for k in range(3):
        x, z = rng.integers(4, size - 4, size=2)
        out[x - 1:x + 2, 2:size - 2, z - 1:z + 2][out[x - 1:x + 2, 2:size - 2, z - 1:z + 2] == 0] = 10 + k
    return np.asfortranarray(out)


def scipy_fill(cc):
    """fill_all_holes on component ids with scipy's fill in each original box."""
    cur = cc.copy()
    count = int(cc.max())
    boxes = {}
    for i in range(1, count + 1):
        where = np.argwhere(cc == i)
        if len(where):
            boxes[i] = (where.min(axis=0), where.max(axis=0) + 1)
    skip = set()
    for i in range(1, count + 1):
        if i not in boxes or i in skip:
            continue
        lo, hi = boxes[i]
        box = tuple(slice(a, b) for a, b in zip(lo, hi))
        mask = cur[box] == i
        voids = ndi.binary_fill_holes(mask) & ~mask
        owners = np.unique(cur[box][voids])
        skip.update(int(o) for o in owners if o and o != i)
        view = cur[box]
        view[voids] = i
    return cur


@pytest.mark.parametrize("seed,split", [(0, False), (1, True)])
def test_oracle_fill_equals_scipy(seed, split):
    labels = hollow_cells(seed, split=split)
    cc, mapping, roots = oracle.ccl(labels)
    want = scipy_fill(cc)
    got = oracle.prepare(labels, P, (1, 1, 1), 0, fix_borders=False, fill_holes=True)["cc"]
    assert np.array_equal(got, want)
    assert not np.array_equal(cc, want)  # something was filled


def test_oracle_fill_skeletons_equal_a_prefilled_volume():
    labels = hollow_cells()
    cc, mapping, _ = oracle.ccl(labels)
    prefilled = np.asfortranarray(mapping[scipy_fill(cc)].astype(labels.dtype))
    got = oracle.skeletonize(labels, P, (1, 1, 1), 0, fix_borders=True, fill_holes=True)
    want = oracle.skeletonize(prefilled, P, (1, 1, 1), 0, fix_borders=True)
    assert 7 not in got["labels"].tolist() and 8 in got["labels"].tolist()  # the cube is gone, the staircase stays

    def by_label(r):
        return {int(l): tuple(np.asarray(r[n][r["v_off" if n != "edges" else "e_off"][k]:
                                             r["v_off" if n != "edges" else "e_off"][k + 1]]).tobytes()
                              for n in ("vertices", "edges", "radii"))
                for k, l in enumerate(r["labels"].tolist())}
    assert by_label(got) == by_label(want)


# ---- GPU ----

@pytest.fixture(scope="module")
def brook():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import brook_amd
    return brook_amd


@pytest.mark.gpu
@pytest.mark.parametrize("seed,split", [(0, False), (1, True)])
@pytest.mark.parametrize("fix_branching", [True, False])
def test_gpu_fill_holes_equals_oracle(brook, seed, split, fix_branching):
    """`split`: a label that filling cuts in two; its flood reaches one piece, and a target in the
    other piece, which no path reaches, ends the label (the reference's per-component rule)."""
    labels = hollow_cells(seed, split=split)
    got = brook.skeletonize(labels, P, (2, 2, 3), dust_threshold=0, fix_branching=fix_branching, fill_holes=True,
                            output="packed", progress=False).as_dict()
    want = oracle.skeletonize(labels, P, (2, 2, 3), 0, fix_branching=fix_branching, fill_holes=True)
    assert got["labels"].tolist() == want["labels"].tolist()
    for name in ("v_off", "e_off", "vertices", "edges", "radii"):
        assert np.array_equal(np.asarray(got[name]), np.asarray(want[name])), name


@pytest.mark.gpu
def test_gpu_fill_holes_tubes_equal_oracle(brook):
    from brook_amd import synth
    labels = synth.tubes_128(1001)
    params = dict(scale=1.5, const=300, pdrf_scale=100000, pdrf_exponent=4)
    got = brook.skeletonize(labels, params, (16, 16, 40), fill_holes=True, output="packed", progress=False).as_dict()
    want = oracle.skeletonize(labels, params, (16, 16, 40), 1000, fill_holes=True, threads=8)
    assert got["labels"].tolist() == want["labels"].tolist()
    for name in ("v_off", "e_off", "vertices", "edges", "radii"):
        assert np.array_equal(np.asarray(got[name]), np.asarray(want[name])), name


@pytest.mark.gpu
def test_gpu_fill_holes_with_somata(brook):
    """Filled holes leave a soma without voids: the soma path runs on the filled volume."""
    params = shapes.SOMA_PARAMS
    for labels in (shapes.soma_cavity(nucleus=True), shapes.soma_cavity(staircase=True)):
        got = brook.skeletonize(labels, params, (1, 1, 1), dust_threshold=0, fill_holes=True, output="packed",
                                progress=False).as_dict()
        want = oracle.skeletonize(labels, params, (1, 1, 1), 0, fill_holes=True)
        assert got["labels"].tolist() == want["labels"].tolist()
        for name in ("v_off", "e_off", "vertices", "edges", "radii"):
            assert np.array_equal(np.asarray(got[name]), np.asarray(want[name])), name


def _matches_painted_root(summary, fix_branching):
    want = shapes.PAINTED_ROOT_EXPECTED[fix_branching]
    return list(summary) == [2, 1] and all(summary[label][:len(v)] == v for label, v in want.items())


@pytest.mark.parametrize("fix_branching", [False, True])
def test_oracle_painted_root_equals_reference_digests(fix_branching):
    """The reference's Kimimaro 5.8.1 summaries of its painted-root sample (test_package.py:1786-1811)."""
    labels = shapes.painted_root()
    for fix_avocados in (False, True):  # no avocado candidate exceeds the sample's thresholds
        got = oracle.skeletonize(labels, shapes.HOLE_PARAMS, (1, 1, 1), 10, fix_branching=fix_branching,
                                 fix_borders=False, fill_holes=True, fix_avocados=fix_avocados)
        assert _matches_painted_root(shapes.summary_from_packed(got), fix_branching), fix_avocados


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [False, True])
def test_gpu_painted_root_equals_reference_digests(brook, fix_branching):
    from brook_amd import _native as native
    labels = shapes.painted_root()
    for fix_avocados in (False, True):
        got = brook.skeletonize(labels, shapes.HOLE_PARAMS, dust_threshold=10, fix_branching=fix_branching,
                                fix_borders=False, fill_holes=True, fix_avocados=fix_avocados, output="packed",
                                progress=False)
        assert native.context().stat("hole_filled_voxels") == shapes.PAINTED_ROOT_FILLED
        assert _matches_painted_root(shapes.summary_from_packed(got.as_dict()), fix_branching), fix_avocados
