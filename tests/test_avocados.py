"""fix_avocados against the known answers the reference's suite holds from Kimimaro 5.8.1
(tests/python/test_package.py:843-990): which labels survive and in what order on three samples
(a pit half inside a one-voxel fruit shell, the same with x and z swapped, three labels whose
first voxels come in the order 3, 7, 5), the merged labels' refreshed distance field, and the
exact skeleton of a pit filling its fruit under anisotropy (1.1, 1.3, 2). The input arrays are the
reference test's numpy expressions restated. They pin both the product and the oracle's own
restatement; GPU tests also compare the product with the oracle on further volumes.
"""
import numpy as np
import pytest

import oracle

PARAMS = dict(scale=1.5, const=1, pdrf_scale=100000, pdrf_exponent=4, soma_detection_threshold=4.5,
              soma_acceptance_threshold=1e6, soma_invalidation_scale=1., soma_invalidation_const=0)
KEYS = [[2], [2, 1], [3, 7, 5]]
REFRESH_ANISOTROPY = (1.1, 1.3, 2.0)


def rule_samples():
    tie = np.zeros((16, 9, 16), np.uint32)
    tie[8:13, 1:6, 1:6] = 1
    tie[9:12, 2:5, 2:5] = 2
    tie[3:9, 2, 2] = 2
    tie[3, 2, 2:9] = 2
    tie[2:5, 2:5, 9:12] = 2
    order = np.zeros((16, 9, 16), np.uint32)
    order[0:3, 0:3, 10:13] = 5
    order[10:13, 0:3, 0:3] = 3
    order[5:8, 0:3, 5:8] = 7
    return [tie, np.ascontiguousarray(tie.transpose(2, 1, 0)), order]


def refresh_sample():
    sample = np.zeros((7, 7, 7), np.uint32)
    sample[1:6, 1:6, 1:6] = 1
    sample[2:5, 2:5, 2:5] = 2
    return sample


def test_oracle_avocado_known_answers():
    """The oracle's restatement against the same known answers (stage: survivors, mapping, first
    voxels, refreshed field; skeletons: label order, the refresh sample's exact skeleton)."""
    expected = [([0, 2], [168], {1: 1, 2: 1}, True), ([0, 2, 1], [329, 1169], {2: 1, 1: 2}, False),
                ([0, 3, 7, 5], [10, 725, 1440], {3: 1, 7: 2, 5: 3}, False)]
    for sample, keys, (mapping, roots, ids, refreshed) in zip(rule_samples(), KEYS, expected):
        for fix_branching in (True, False):
            got = oracle.skeletonize(np.asfortranarray(sample), PARAMS, (1, 1, 1), 0, fix_branching=fix_branching,
                                     fix_avocados=True)
            assert got["labels"].tolist() == keys
        p = oracle.prepare(np.asfortranarray(sample), PARAMS, (1, 1, 1), 0, fix_borders=False, fix_avocados=True)
        want = np.select([sample == k for k in ids], list(ids.values()), 0)
        assert np.array_equal(p["cc"], want)
        assert p["count"] == len(roots)
        cc, _, _ = oracle.ccl(np.asfortranarray(want if refreshed else sample))
        assert np.array_equal(p["dbf"], oracle.edt(cc, (1, 1, 1), False))
    vertices = np.arange(1, 6, dtype=np.float32)[:, None] * np.array(REFRESH_ANISOTROPY, np.float32)
    radii = [0x3f8ccccd, 0x400ccccd, 0x40533334, 0x400ccccd, 0x3f8ccccd]
    for fix_branching in (True, False):
        r = oracle.skeletonize(np.asfortranarray(refresh_sample()), dict(PARAMS, soma_detection_threshold=5),
                               REFRESH_ANISOTROPY, 0, fix_branching=fix_branching, fix_avocados=True)
        assert r["labels"].tolist() == [1]
        assert np.array_equal(r["vertices"], vertices)
        assert r["edges"].tolist() == [[0, 1], [1, 2], [2, 3], [3, 4]]
        assert r["radii"].view(np.uint32).tolist() == radii


@pytest.fixture(scope="module")
def brook():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import brook_amd
    return brook_amd


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
def test_gpu_avocado_survivors_and_order(brook, fix_branching):
    options = dict(dust_threshold=0, fix_branching=fix_branching, fix_avocados=True, progress=False)
    for sample, keys in zip(rule_samples(), KEYS):
        for volume in (np.asfortranarray(sample), np.ascontiguousarray(sample), sample.astype(np.uint8)):
            assert list(brook.skeletonize(volume, PARAMS, **options)) == keys
    batch = brook.skeletonize_batch(rule_samples(), teasar_params=PARAMS, output="skeletons", **options)
    assert [list(result) for result in batch] == KEYS


@pytest.mark.gpu
def test_gpu_avocado_refreshed_field(brook):
    """The merged tie sample's DBF is the EDT of the merged labels; the others keep theirs."""
    for i, sample in enumerate(rule_samples()):
        got = brook.prepare_fields(sample, teasar_params=PARAMS, dust_threshold=0, fix_borders=True, fix_avocados=True)
        want_labels = (sample > 0).astype(np.uint32) if i == 0 else sample
        cc, _, _ = oracle.ccl(np.asfortranarray(want_labels))
        assert np.array_equal(got["dbf"], oracle.edt(cc, (1, 1, 1), False)), i


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
def test_gpu_avocado_refresh_skeleton(brook, fix_branching):
    vertices = np.arange(1, 6, dtype=np.float32)[:, None] * np.array(REFRESH_ANISOTROPY, np.float32)
    radii = [0x3f8ccccd, 0x400ccccd, 0x40533334, 0x400ccccd, 0x3f8ccccd]
    params = dict(PARAMS, soma_detection_threshold=5)
    for volume in (np.ascontiguousarray(refresh_sample()), np.asfortranarray(refresh_sample())):
        result = brook.skeletonize(volume, params, REFRESH_ANISOTROPY, dust_threshold=0, fix_branching=fix_branching,
                                   fix_avocados=True, progress=False)
        assert list(result) == [1]
        np.testing.assert_array_equal(result[1].vertices, vertices)
        np.testing.assert_array_equal(result[1].edges, [[0, 1], [1, 2], [2, 3], [3, 4]])
        assert np.asarray(result[1].radii, np.float32).view(np.uint32).tolist() == radii


def avocado_cells(seed=0, size=64):
    """Fruits (shells) with pits (balls of another label inside, one breaking out through a gap),
    a fruit with two pits, tubes: several merges and survivors per pass."""
    rng = np.random.default_rng(seed)
    out = np.zeros((size, size, size), np.uint32)
    grid = np.indices(out.shape)

    def ball(c, r):
        return ((grid - np.reshape(c, (3, 1, 1, 1))) ** 2).sum(axis=0) <= r * r

    out[ball((16, 16, 16), 10)] = 4
    out[ball((16, 16, 16), 6)] = 9                      # pit fully inside
    out[ball((44, 20, 40), 11)] = 2
    out[ball((44, 20, 40), 7)] = 6
    out[ball((53, 20, 40), 4)] = 6                      # the pit breaks out on +x
    out[ball((20, 46, 44), 12)] = 3
    out[ball((14, 46, 44), 4)] = 8                      # two pits in one fruit
    out[ball((26, 46, 44), 4)] = 11
    for k in range(2):
        x, z = rng.integers(4, size - 4, size=2)
        sel = (slice(x - 1, x + 2), slice(2, size - 2), slice(z - 1, z + 2))
        out[sel][out[sel] == 0] = 20 + k
    return np.asfortranarray(out)


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
@pytest.mark.parametrize("seed", [0, 1])
def test_gpu_avocados_equal_oracle(brook, fix_branching, seed):
    labels = avocado_cells(seed)
    params = dict(PARAMS, soma_detection_threshold=12)
    got = brook.skeletonize(labels, params, (1, 1, 1), dust_threshold=0, fix_branching=fix_branching,
                            fix_avocados=True, output="packed", progress=False).as_dict()
    want = oracle.skeletonize(labels, params, (1, 1, 1), 0, fix_branching=fix_branching, fix_avocados=True)
    assert got["labels"].tolist() == want["labels"].tolist()
    for name in ("v_off", "e_off", "vertices", "edges", "radii"):
        assert np.array_equal(np.asarray(got[name]), np.asarray(want[name])), name


def test_oracle_avocado_cells_merge():
    labels = avocado_cells()
    params = dict(PARAMS, soma_detection_threshold=12)
    plain = oracle.skeletonize(labels, params, (1, 1, 1), 0)["labels"].tolist()
    fixed = oracle.skeletonize(labels, params, (1, 1, 1), 0, fix_avocados=True)["labels"].tolist()
    assert len(fixed) < len(plain)  # some pit merged
