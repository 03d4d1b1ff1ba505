"""Pins of the oracle's target order, invalidation and analysis that do not restate their code:

* target order (lockstep_order.cu:21-43): the rule of the reference's component_order_smoke.cpp:9-21
  (descending DAF with +-inf and +-0 as 0, ties by descending flat index) evaluated by numpy's
  lexsort, on the reference test's grid of lengths, densities and tie rates, with +inf every 19th
  and -0 every 23rd voxel, over binary masks and masks of several ids;
* invalidation (invalidation_kernels.cuh:105-299): with one source the claimed set is the
  26-connected component of {same label, alive, distance to the source < ball} that holds the
  source, and every claim key is the exact minimax bottleneck (a Dijkstra over max); with many
  sources every claim holds (its distance to its source is inside the source's ball, its bottleneck
  bounds that distance, it touches a claimed voxel or is a source) and the claimed set is closed
  (no alive neighbour of a claimed voxel lies inside the ball of that voxel's final source);
* analysis (preamble.cu:10-29): numpy bincount and per-id coordinate minima and maxima.
"""
import heapq

import numpy as np
import pytest

import oracle

NX = [-1, 1, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, -1, -1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1]
NY = [0, 0, -1, 1, 0, 0, -1, 1, -1, 1, -1, -1, 1, 1, 0, 0, 0, 0, -1, -1, 1, 1, -1, -1, 1, 1]
NZ = [0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1]


def _expected_order(ids, daf):
    flat_ids, flat_daf = ids.ravel(order="F"), daf.ravel(order="F")
    key = np.where(np.isinf(flat_daf) | (flat_daf == 0), np.float32(0), flat_daf).astype(np.float64)
    out = []
    for label in np.unique(flat_ids[flat_ids > 0]):
        idx = np.flatnonzero(flat_ids == label)
        out.append(idx[np.lexsort((-idx, -key[idx]))])  # last key first: DAF descending, then index descending
    return np.concatenate(out) if out else np.zeros(0, np.int64)


@pytest.mark.parametrize("n", [1, 17, 1025, 100003])
def test_target_order_rule(n):
    rng = np.random.default_rng(482 + n)
    for density in (0, 3, 50, 100):
        for ties in (1, 7, 10000):
            for ids_max in (1, 7):  # a binary mask, and several ids (one order per id)
                on = rng.integers(0, 100, n) < density
                ids = np.where(on, rng.integers(1, ids_max + 1, n), 0).astype(np.uint32)
                daf = (rng.integers(0, ties, n).astype(np.float32) * np.float32(0.37)).astype(np.float32)
                daf[::19] = np.inf
                daf[::23] = -0.0
                ids3, daf3 = ids[:, None, None], daf[:, None, None]
                got = oracle.target_order(ids3, daf3)
                want = _expected_order(ids3, daf3)
                assert np.array_equal(got, want), (n, density, ties, ids_max)


def _f32_distance(a, b, aniso):
    """sqrtf(fl(dx^2) + fl(dy^2) + fl(dz^2)) with dx = fl(w_x (x_a - x_b)) (offer_distance)."""
    d = [np.float32(aniso[i]) * np.float32(a[i] - b[i]) for i in range(3)]
    return np.sqrt(np.float32(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])) + np.float32(d[2] * d[2]))


def _ball(dbf, s, scale, const):
    return np.float32(np.float32(np.float32(scale) * dbf[s]) + np.float32(const))


def _neighbours(p, shape):
    for k in range(26):
        q = (p[0] + NX[k], p[1] + NY[k], p[2] + NZ[k])
        if all(0 <= q[a] < shape[a] for a in range(3)):
            yield q


def _random_case(seed, shape=(18, 16, 12)):
    rng = np.random.default_rng(seed)
    labels = (rng.random(shape) < 0.8).astype(np.uint32) + (rng.random(shape) < 0.1)
    labels = np.asfortranarray(labels)
    dbf = np.asfortranarray(rng.choice([1.0, 1.5, 2.25, 3.0], shape).astype(np.float32))
    alive = np.asfortranarray((rng.random(shape) < 0.9).astype(np.uint8))
    return rng, labels, dbf, alive


@pytest.mark.parametrize("seed", range(6))
def test_invalidation_single_source_is_the_ball_component(seed):
    rng, labels, dbf, alive = _random_case(seed)
    shape = labels.shape
    aniso = [(1, 1, 1), (4, 4, 40), (16, 16, 40), (1.5, 2.25, 3.0)][seed % 4]
    scale, const = [(1.0, 2.0), (1.5, 10.0), (3.0, 5.0)][seed % 3]
    while True:
        s = tuple(int(rng.integers(0, shape[a])) for a in range(3))
        if labels[s] and alive[s]:
            break
    L = labels[s]
    r = _ball(dbf, s, scale, const)
    # expected: the claim keys are the minimax bottleneck over paths inside the ball (Dijkstra on max)
    best = {s: np.float32(0)}
    heap = [(0.0, s)]
    done = set()
    while heap:
        m, p = heapq.heappop(heap)
        if p in done:
            continue
        done.add(p)
        for q in _neighbours(p, shape):
            if labels[q] != L or not alive[q] or q in done:
                continue
            d = _f32_distance(q, s, aniso)
            if not d < r:
                continue
            cand = max(np.float32(best[p]), d)
            if q not in best or cand < best[q]:
                best[q] = cand
                heapq.heappush(heap, (float(cand), q))
    flat_s = s[0] + shape[0] * (s[1] + shape[1] * s[2])
    n, after, claims = oracle.invalidate(labels, dbf, alive, [flat_s], aniso, scale, const)
    claimed = {tuple(int(c) for c in np.unravel_index(i, shape, order="F"))
               for i in np.flatnonzero(claims.ravel(order="F") != np.uint64(0xFFFFFFFFFFFFFFFF))}
    assert claimed == done and n == len(done)
    for p in done:
        k = int(claims[p])
        assert k & 0xFFFFFFFF == flat_s and np.uint32(k >> 32).view(np.float32) == best[p], p
        assert after[p] == 0
    assert np.count_nonzero(alive) - np.count_nonzero(after) == len(done)


@pytest.mark.parametrize("seed", range(6))
def test_invalidation_claims_are_valid_and_closed(seed):
    rng, labels, dbf, alive = _random_case(100 + seed, shape=(22, 18, 14))
    shape = labels.shape
    aniso = [(1, 1, 1), (4, 4, 40), (1.5, 2.25, 3.0)][seed % 3]
    scale, const = (1.5, 3.0)
    # a path: a random walk through alive voxels of one label
    while True:
        p = tuple(int(rng.integers(0, shape[a])) for a in range(3))
        if labels[p] == 1 and alive[p]:
            break
    path = [p]
    for _ in range(40):
        options = [q for q in _neighbours(path[-1], shape) if labels[q] == 1]
        if not options:
            break
        path.append(options[int(rng.integers(0, len(options)))])
    flat = lambda q: q[0] + shape[0] * (q[1] + shape[1] * q[2])
    sources = {flat(q) for q in path if alive[q]}
    n, after, claims = oracle.invalidate(labels, dbf, alive, [flat(q) for q in path], aniso, scale, const)
    none = np.uint64(0xFFFFFFFFFFFFFFFF)
    claimed = {tuple(int(c) for c in np.unravel_index(i, shape, order="F"))
               for i in np.flatnonzero(claims.ravel(order="F") != none)}
    assert n == len(claimed) and sources <= {flat(q) for q in claimed}
    unflat = lambda i: tuple(int(c) for c in np.unravel_index(i, shape, order="F"))
    for q in claimed:
        k = int(claims[q])
        src, m = k & 0xFFFFFFFF, np.uint32(k >> 32).view(np.float32)
        assert labels[q] == 1 and alive[q] and after[q] == 0
        d = _f32_distance(q, unflat(src), aniso)
        if flat(q) == src:
            assert m == 0 and src in sources
            continue
        assert d < _ball(dbf, unflat(src), scale, const) and m >= d, q
        assert any(claims[u] != none for u in _neighbours(q, shape)), q
        # closure under the final key
        for w in _neighbours(q, shape):
            if labels[w] == 1 and alive[w] and claims[w] == none:
                assert not _f32_distance(w, unflat(src), aniso) < _ball(dbf, unflat(src), scale, const), (q, w)


@pytest.mark.parametrize("seed", range(4))
def test_analyze_matches_numpy(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(x) for x in rng.integers(5, 40, 3))
    count = int(rng.integers(1, 30))
    cc = np.asfortranarray(rng.integers(0, count + 1, shape).astype(np.uint32) * (rng.random(shape) < 0.3))
    lo, hi, counts = oracle.analyze(cc, count)
    assert np.array_equal(counts, np.bincount(cc.ravel(), minlength=count + 1))
    coords = np.indices(shape)
    for i in range(1, count + 1):
        m = cc == i
        if not m.any():
            assert list(lo[i]) == [2147483647] * 3 and list(hi[i]) == [-1] * 3
            continue
        assert [int(coords[a][m].min()) for a in range(3)] == list(lo[i])
        assert [int(coords[a][m].max()) for a in range(3)] == list(hi[i])


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 17, 1025, 100003])
def test_gpu_target_order_equals_oracle(n):
    """The engine's target-order sort (hipCUB selection and a stable pair sort), through the C ABI."""
    import ctypes
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from brook_amd import _native as native
    from brook_amd._intake import volume_struct
    L = native.lib()
    L.brook_target_order.restype = ctypes.c_int
    L.brook_target_order.argtypes = [ctypes.c_void_p, ctypes.POINTER(native.Volume), ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    rng = np.random.default_rng(482 + n)
    for density in (0, 3, 50, 100):
        for ties in (1, 7, 10000):
            for ids_max, dtype in ((1, np.uint8), (7, np.uint16), (300, np.uint32)):
                on = rng.integers(0, 100, n) < density
                ids = np.where(on, rng.integers(1, ids_max + 1, n), 0).astype(dtype)[:, None, None]
                daf = (rng.integers(0, ties, n).astype(np.float32) * np.float32(0.37)).astype(np.float32)
                daf[::19] = np.inf
                daf[::23] = -0.0
                keep = []
                vol, _ = volume_struct(np.asfortranarray(ids), keep)
                order = np.zeros(max(n, 1), np.int64)
                count = ctypes.c_size_t()
                native.check(L.brook_target_order(native.context().handle, ctypes.byref(vol), daf.ctypes.data,
                                                  order.ctypes.data, order.size, ctypes.byref(count)))
                want = oracle.target_order(ids.astype(np.uint32), daf[:, None, None])
                assert count.value == len(want) and np.array_equal(order[:count.value], want), (density, ties, ids_max)
