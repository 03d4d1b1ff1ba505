"""Manual targets: extra_targets_before (traced ahead of the DAF order, last given first) and
extra_targets_after (traced once nothing of the component is left alive, last given first;
trace.cu:206-245). The samples restate the reference suite's `_targeted_samples`
(tests/python/test_package.py:513-543): a stub that the bar's invalidation ball swallows reached
as a `before` and as an `after` target, a sample with no targets, points on background and one
written from the far edge (a negative coordinate, filed under the wrapped point's component and
traced at the coordinates as written, skeletonize.cu:59-65), and a soma-sized body whose root goes
ahead of the manual targets.

CPU tests pin the oracle: the targets must change the skeletons where the reference's test says
they do and nowhere else, the target voxel becomes a vertex, `after` targets count toward
max_paths (a budget they fill leaves the label untraced), an `after` target adds one path that
walks from it to the skeleton the DAF order laid, and points whose written coordinates flatten off
their component are rejected. GPU tests compare brook_amd with the oracle bit for bit, single calls
and batches.
"""
import numpy as np
import pytest

import oracle

PARAMS = dict(scale=1.5, const=2, pdrf_scale=100000, pdrf_exponent=4, soma_detection_threshold=6,
              soma_acceptance_threshold=8, soma_invalidation_scale=1., soma_invalidation_const=0)
CANVAS = (28, 40, 40)
STUB_TIP = (12, 19, 22)


def stubbed_bar():
    volume = np.zeros(CANVAS, np.uint16)
    volume[10:14, 10:14, 4:36] = 1     # a bar across the volume
    volume[10:14, 14:20, 20:24] = 1    # a stub the bar's invalidation swallows
    return volume


def targeted_samples():
    samples, before, after = [], [], []
    samples.append(stubbed_bar()); before.append([STUB_TIP]); after.append([])
    samples.append(stubbed_bar()); before.append([]); after.append([STUB_TIP])
    plain = np.zeros(CANVAS, np.uint16)
    plain[10:14, 10:14, 4:36] = 3
    plain[20:26, 24:34, 24:34] = 4
    samples.append(plain); before.append([]); after.append([])
    mixed = stubbed_bar()
    mixed[2:6, 30:36, 30:36] = 2
    samples.append(mixed); before.append([(1, 1, 1), (12, -21, 22)]); after.append([(2, 2, 2)])
    z, y, x = np.meshgrid(np.arange(28), np.arange(40), np.arange(40), indexing="ij")
    body = np.zeros(CANVAS, np.uint16)
    body[(x - 14) ** 2 + (y - 14) ** 2 + (z - 14) ** 2 <= 100] = 5
    body[12:16, 12:16, 14:38] = 5
    samples.append(body); before.append([(14, 14, 37)]); after.append([])
    return samples, before, after


def by_label(r):
    out = {}
    for k, label in enumerate(r["labels"].tolist()):
        v0, v1, e0, e1 = r["v_off"][k], r["v_off"][k + 1], r["e_off"][k], r["e_off"][k + 1]
        out[label] = (np.asarray(r["vertices"][v0:v1]), np.asarray(r["edges"][e0:e1]), np.asarray(r["radii"][v0:v1]))
    return out


def oracle_run(sample, before=(), after=(), fix_branching=True, params=PARAMS):
    return oracle.skeletonize(np.asfortranarray(sample), params, (1, 1, 1), 10, fix_branching=fix_branching,
                              extra_targets_before=before, extra_targets_after=after)


@pytest.mark.parametrize("fix_branching", [True, False])
def test_oracle_targets_matter_where_the_reference_says(fix_branching):
    """The reference's own check (test_package.py:559-562): samples 0, 1, 3, 4 change, 2 does not;
    the stub tip is a vertex wherever it is a target."""
    samples, before, after = targeted_samples()
    for i, (sample, b, a) in enumerate(zip(samples, before, after)):
        got = by_label(oracle_run(sample, b, a, fix_branching))
        plain = by_label(oracle_run(sample, fix_branching=fix_branching))
        assert list(got) == list(plain)
        changed = any(not np.array_equal(got[k][0], plain[k][0]) for k in plain)
        assert changed == (i != 2), i
        if i in (0, 1):
            assert (got[1][0] == np.float32(STUB_TIP)).all(1).any()
            assert not (plain[1][0] == np.float32(STUB_TIP)).all(1).any()


def test_oracle_after_target_is_traced_last():
    """An `after` target adds exactly one path, from the target to the rail the DAF-order paths
    laid: the skeleton with it holds the plain one's vertices, plus the stub's path."""
    sample = stubbed_bar()
    plain = by_label(oracle_run(sample))[1]
    got = oracle_run(sample, after=[STUB_TIP])
    assert got["stats"]["paths"] == oracle_run(sample)["stats"]["paths"] + 1
    verts = {tuple(v) for v in by_label(got)[1][0].tolist()}
    assert {tuple(v) for v in plain[0].tolist()} <= verts
    stub = {v for v in verts if v[1] >= 14}
    assert STUB_TIP in stub and len(stub) >= 19 - 14 + 1   # a path walks the stub down to the bar


def test_oracle_targets_count_toward_max_paths():
    """A budget the manual targets fill leaves the label untraced (trace.cu:206); with no `before`
    target the DAF target counts as one (trace.cu:161)."""
    sample = stubbed_bar()
    two, three = dict(PARAMS, max_paths=2), dict(PARAMS, max_paths=3)
    assert 1 in by_label(oracle_run(sample, params=two))
    assert 1 not in by_label(oracle_run(sample, after=[STUB_TIP], params=two))
    got = by_label(oracle_run(sample, after=[STUB_TIP], params=three))
    assert (got[1][0] == np.float32(STUB_TIP)).all(1).any()
    assert 1 not in by_label(oracle_run(sample, before=[STUB_TIP], after=[(12, 12, 30), (12, 12, 6)], params=three))
    assert 1 in by_label(oracle_run(sample, before=[STUB_TIP], after=[(12, 12, 30)], params=three))


def test_oracle_rejects_targets_off_their_component():
    sample = stubbed_bar()
    with pytest.raises(RuntimeError, match="outside"):
        oracle_run(sample, after=[(28, 0, 0)])
    with pytest.raises(RuntimeError, match="flatten"):
        oracle_run(sample, after=[(-16, 10, 22)])  # wraps onto the bar; as written it flattens onto background


# ---- GPU ----

@pytest.fixture(scope="module")
def brook():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import brook_amd
    return brook_amd


def assert_equal_packed(got, want):
    assert got["labels"].tolist() == want["labels"].tolist()
    for name in ("v_off", "e_off", "vertices", "edges", "radii"):
        assert np.array_equal(np.asarray(got[name]), np.asarray(want[name])), name


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
def test_gpu_targets_equal_oracle(brook, fix_branching):
    samples, before, after = targeted_samples()
    options = dict(dust_threshold=10, fix_branching=fix_branching, progress=False)
    for i, (sample, b, a) in enumerate(zip(samples, before, after)):
        got = brook.skeletonize(np.asfortranarray(sample), PARAMS, extra_targets_before=b, extra_targets_after=a,
                                output="packed", **options).as_dict()
        assert_equal_packed(got, oracle_run(sample, b, a, fix_branching))
    batch = brook.skeletonize_batch(samples, teasar_params=PARAMS, extra_targets_before=before,
                                    extra_targets_after=after, output="packed", **options)
    wants = [oracle_run(sample, b, a, fix_branching) for sample, b, a in zip(samples, before, after)]
    for i in range(len(samples)):
        assert_equal_packed(batch[i].as_dict(), wants[i])
    reps = -(-8 // len(samples))  # at least eight samples: the per-sample stages run on worker threads
    big = brook.skeletonize_batch(samples * reps, teasar_params=PARAMS, extra_targets_before=before * reps,
                                  extra_targets_after=after * reps, output="packed", **options)
    for i in range(len(big)):
        assert_equal_packed(big[i].as_dict(), wants[i % len(samples)])


@pytest.mark.gpu
def test_gpu_batch_error_on_a_worker_thread(brook):
    """In a batch of nine (per-sample stages on the engine's worker threads), a manual target off
    the volume in one sample raises ValueError from that worker; the context then runs the same
    batch with valid targets, every sample equal to the oracle."""
    samples, before, after = targeted_samples()
    reps = -(-9 // len(samples))
    s9, b9, a9 = (samples * reps)[:9], (before * reps)[:9], (after * reps)[:9]
    bad = [list(b) for b in b9]
    bad[5] = [(1000, 0, 0)]
    options = dict(dust_threshold=10, progress=False)
    with pytest.raises(ValueError):
        brook.skeletonize_batch(s9, teasar_params=PARAMS, extra_targets_before=bad, extra_targets_after=a9,
                                output="packed", **options)
    ok = brook.skeletonize_batch(s9, teasar_params=PARAMS, extra_targets_before=b9, extra_targets_after=a9,
                                 output="packed", **options)
    for i in range(9):
        assert_equal_packed(ok[i].as_dict(), oracle_run(s9[i], b9[i], a9[i]))


@pytest.mark.gpu
def test_gpu_after_targets_on_tubes(brook):
    """Several `after` targets per label, on labels many paths cover, and a budget they fill."""
    from brook_amd import synth
    labels = synth.tubes_128(7)
    params = dict(scale=1.5, const=300, pdrf_scale=100000, pdrf_exponent=4)
    rng = np.random.default_rng(3)
    fg = np.argwhere(labels > 0)
    after = [tuple(int(c) for c in fg[i]) for i in rng.choice(len(fg), 12, replace=False)]
    before = [tuple(int(c) for c in fg[i]) for i in rng.choice(len(fg), 4, replace=False)]
    for fix_branching in (True, False):
        for p in (params, dict(params, max_paths=3)):
            got = brook.skeletonize(labels, p, (16, 16, 40), extra_targets_before=before, extra_targets_after=after,
                                    fix_branching=fix_branching, output="packed", progress=False).as_dict()
            want = oracle.skeletonize(labels, p, (16, 16, 40), 1000, fix_branching=fix_branching,
                                      extra_targets_before=before, extra_targets_after=after, threads=8)
            assert_equal_packed(got, want)


@pytest.mark.gpu
def test_gpu_rejects_targets_off_their_component(brook):
    with pytest.raises(ValueError, match="flatten"):
        brook.skeletonize(np.asfortranarray(stubbed_bar()), PARAMS, dust_threshold=10, extra_targets_after=[(-16, 10, 22)],
                          progress=False)


@pytest.mark.gpu
@pytest.mark.parametrize("fix_branching", [True, False])
def test_gpu_after_targets_on_a_soma(brook, fix_branching):
    """A soma-sized body with `after` targets (the process tip and a point inside the soma ball):
    the ball dies first, the targets come once nothing is alive, and their paths are trimmed to
    the soma radius like the others."""
    body = targeted_samples()[0][4]
    for before, after in (([], [(14, 14, 37)]), ([(14, 14, 37)], [(14, 14, 14), (13, 15, 30)])):
        got = brook.skeletonize(np.asfortranarray(body), PARAMS, dust_threshold=10, fix_branching=fix_branching,
                                extra_targets_before=before, extra_targets_after=after, output="packed",
                                progress=False).as_dict()
        assert_equal_packed(got, oracle_run(body, before, after, fix_branching))


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_gpu_random_options_equal_oracle(brook, seed):
    """Tube volumes under random anisotropy, TEASAR constants, budgets and manual targets, in both
    branching modes, against the oracle."""
    from brook_amd import synth
    rng = np.random.default_rng(seed)
    labels = synth.volume("tubes", (96, 80, 64), int(rng.integers(4, 12)), seed)
    aniso = tuple(float(a) for a in rng.choice([1.0, 2.0, 4.0, 8.0, 16.0, 40.0], 3))
    params = dict(scale=float(rng.choice([1.0, 1.5, 3.0])), const=float(rng.choice([2.0, 50.0, 300.0])),
                  pdrf_scale=100000, pdrf_exponent=int(rng.choice([2, 4, 8])))
    if rng.random() < 0.5:
        params["max_paths"] = int(rng.integers(4, 16))
    fg = np.argwhere(labels > 0)
    before = [tuple(int(c) for c in fg[i]) for i in rng.choice(len(fg), int(rng.integers(0, 4)), replace=False)]
    after = [tuple(int(c) for c in fg[i]) for i in rng.choice(len(fg), int(rng.integers(0, 6)), replace=False)]
    for fix_branching in (True, False):
        got = brook.skeletonize(labels, params, aniso, dust_threshold=100, fix_branching=fix_branching,
                                extra_targets_before=before, extra_targets_after=after, output="packed",
                                progress=False).as_dict()
        want = oracle.skeletonize(labels, params, aniso, 100, fix_branching=fix_branching,
                                  extra_targets_before=before, extra_targets_after=after, threads=8)
        assert_equal_packed(got, want)
