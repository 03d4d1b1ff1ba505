"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle (oracle/teasar_oracle.cpp,
oracle/post_oracle.cpp).

The oracle is the CPU restatement of the reference's TEASAR path. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module; the product
package (brook_amd) never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_D = ctypes.c_double

DEFAULT_PARAMS = dict(scale=10.0, const=10.0, pdrf_scale=5000.0, pdrf_exponent=16.0,
                      soma_detection_threshold=1100.0, soma_acceptance_threshold=4000.0,
                      soma_invalidation_scale=0.5, soma_invalidation_const=0.0)
_PARAM_ORDER = ["scale", "const", "pdrf_scale", "pdrf_exponent", "soma_detection_threshold",
                "soma_acceptance_threshold", "soma_invalidation_scale", "soma_invalidation_const"]


def build() -> str:
    """Compile the oracle library (make in oracle/); returns its path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_ccl.restype = _I64
        L.oracle_ccl.argtypes = [_P, _I64, _I64, _I64, _P, _P, _P]
        L.oracle_edt.restype = _I
        L.oracle_edt.argtypes = [_P, _I64, _I64, _I64, _P, _I, _I, _P]
        L.oracle_edt_f64.restype = _I
        L.oracle_edt_f64.argtypes = [_P, _I64, _I64, _I64, _P, _I, _I, _P]
        L.oracle_pointset_order.restype = _I
        L.oracle_pointset_order.argtypes = [_P, _I64, _P]
        L.oracle_border_targets.restype = _I64
        L.oracle_border_targets.argtypes = [_P, _I64, _I64, _I64, _I64, _P, _P, _I64, _P]
        L.oracle_geometric_flood.restype = _I
        L.oracle_geometric_flood.argtypes = [_P, _I64, _I64, _I64, _P, _P, _I64, _P]
        L.oracle_target_order.restype = _I64
        L.oracle_target_order.argtypes = [_P, _I64, _P, _P]
        L.oracle_invalidate.restype = _I64
        L.oracle_invalidate.argtypes = [_P, _I64, _I64, _I64, _P, _P, _P, _I64, _P, ctypes.c_float, ctypes.c_float, _P]
        L.oracle_analyze.restype = _I
        L.oracle_analyze.argtypes = [_P, _I64, _I64, _I64, _I64, _P, _P, _P]
        L.oracle_railroad.restype = _I64
        L.oracle_railroad.argtypes = [_P, _P, _I64, _I64, _I64, _I64, _P, _I64]
        L.oracle_prepare.restype = _I64
        L.oracle_prepare.argtypes = [_P, _I64, _I64, _I64, _P, _P, _D, _I, _I, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P]
        L.oracle_vw_parents.argtypes = [_P, _I64, _I64, _I64, _P, _P, _I64, _P, _P]
        L.oracle_skeletonize.restype = _P
        L.oracle_skeletonize.argtypes = [_P, _I64, _I64, _I64, _P, _I64, _P, _D, _I, _I, _P, _I64, _P, _I64, _I, _P, _I]
        L.oracle_graph_ccl.restype = _I64
        L.oracle_graph_ccl.argtypes = [_P, _I64, _I64, _I64, _P, _I, _P, _P, _P, _P, _P]
        L.oracle_graph_edt.restype = _I
        L.oracle_graph_edt.argtypes = [_P, _I64, _I64, _I64, _P, _P, _I, _P]
        L.oracle_result_sizes.argtypes = [_P, _P]
        L.oracle_result_copy.argtypes = [_P, _P, _P, _P, _P, _P, _P]
        L.oracle_result_free.argtypes = [_P]
        L.oracle_post.restype = _P
        L.oracle_post.argtypes = [_I, ctypes.c_size_t, _P, _P, _P, _P, _P, _P, _D, _D]
        L.oracle_post_error.restype = ctypes.c_char_p
        L.oracle_post_sizes.argtypes = [_P, _P]
        L.oracle_post_copy.argtypes = [_P, _P, _P, _P, _P, _P, _P]
        L.oracle_post_free.argtypes = [_P]
        L.oracle_post_components.restype = _I
        L.oracle_post_components.argtypes = [ctypes.c_size_t, _P, _P, _P, _P, _P, _P, _P]
        L.oracle_pairwise_sum.restype = ctypes.c_float
        L.oracle_pairwise_sum.argtypes = [_P, ctypes.c_size_t]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_P)


def _fail(what):
    raise RuntimeError(f"oracle {what}: {lib().oracle_last_error().decode()}")


def _fortran_int64(labels):
    a = np.asarray(labels)
    while a.ndim < 3:
        a = a[..., None]
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    return np.asfortranarray(a.astype(np.int64, copy=False)), a.shape


def params_vector(teasar_params=None):
    p = dict(DEFAULT_PARAMS)
    p.update({k: v for k, v in (teasar_params or {}).items() if k != "max_paths"})
    return np.array([float(p[k]) for k in _PARAM_ORDER], np.float64)


def ccl(labels):
    """(cc uint32 Fortran array, mapping int64[N+1], roots int64[N])."""
    a, shape = _fortran_int64(labels)
    n = a.size
    cc = np.zeros(shape, np.uint32, order="F")
    mapping = np.zeros(n + 1, np.int64)
    roots = np.zeros(max(n, 1), np.int64)
    count = lib().oracle_ccl(_ptr(a), *shape, _ptr(cc), _ptr(mapping), _ptr(roots))
    if count < 0:
        _fail("ccl")
    return cc, mapping[:count + 1].copy(), roots[:count].copy()


def _graph(graph, shape):
    """(uint32 Fortran graph, compact_2d): a uint8 graph (one XY plane) is widened value for value."""
    g = np.asarray(graph)
    while g.ndim < 3:
        g = g[..., None]
    if g.shape != tuple(shape):
        raise ValueError("voxel_graph must match the label shape")
    compact = g.dtype == np.uint8
    if compact and g.shape[2] != 1:
        raise ValueError("a uint8 voxel_graph must be one XY plane")
    if not compact and g.dtype != np.uint32:
        raise ValueError("voxel_graph must be uint32, or uint8 for an XY plane")
    return np.asfortranarray(g.astype(np.uint32)), compact


def graph_ccl(labels, graph):
    """Voxel-graph components: (cc uint32 Fortran array, mapping int64[N+1], roots int64[N],
    keys int64[N+1], input_uniform)."""
    a, shape = _fortran_int64(labels)
    g, compact = _graph(graph, shape)
    n = a.size
    cc = np.zeros(shape, np.uint32, order="F")
    mapping, keys = np.zeros(n + 1, np.int64), np.zeros(n + 1, np.int64)
    roots = np.zeros(max(n, 1), np.int64)
    uniform = np.zeros(1, np.int32)
    count = lib().oracle_graph_ccl(_ptr(a), *shape, _ptr(g), int(compact), _ptr(cc), _ptr(mapping), _ptr(roots),
                                   _ptr(keys), _ptr(uniform))
    if count < 0:
        _fail("graph_ccl")
    return cc, mapping[:count + 1].copy(), roots[:count].copy(), keys[:count + 1].copy(), bool(uniform[0])


def graph_edt(cc, graph, anisotropy=(1, 1, 1), black_border=False):
    """The graph-aware EDT of a label volume (float32, Fortran order)."""
    cc = np.asfortranarray(np.asarray(cc).astype(np.uint32, copy=False))
    while cc.ndim < 3:
        cc = cc[..., None]
    g, _ = _graph(graph, cc.shape)
    out = np.zeros(cc.shape, np.float32, order="F")
    an = np.asarray(anisotropy, np.float32)
    if lib().oracle_graph_edt(_ptr(cc), *cc.shape, _ptr(g), _ptr(an), int(black_border), _ptr(out)):
        _fail("graph_edt")
    return out


def edt(cc, anisotropy=(1, 1, 1), black_border=False, dims=3):
    cc = np.asfortranarray(np.asarray(cc).astype(np.uint32, copy=False))
    while cc.ndim < 3:
        cc = cc[..., None]
    out = np.zeros(cc.shape, np.float32, order="F")
    an = np.asarray(anisotropy, np.float32)
    if lib().oracle_edt(_ptr(cc), *cc.shape, _ptr(an), int(black_border), dims, _ptr(out)):
        _fail("edt")
    return out


def edt_f64(cc, anisotropy=(1, 1, 1), black_border=False, dims=3):
    cc = np.asfortranarray(np.asarray(cc).astype(np.uint32, copy=False))
    while cc.ndim < 3:
        cc = cc[..., None]
    out = np.zeros(cc.shape, np.float64, order="F")
    an = np.asarray([float(np.float32(a)) for a in anisotropy], np.float64)
    if lib().oracle_edt_f64(_ptr(cc), *cc.shape, _ptr(an), int(black_border), dims, _ptr(out)):
        _fail("edt_f64")
    return out


def pointset_order(points):
    pts = np.ascontiguousarray(np.asarray(points, np.int64).reshape(-1, 3))
    order = np.full(len(pts), -1, np.int64)
    if lib().oracle_pointset_order(_ptr(pts), len(pts), _ptr(order)):
        _fail("pointset")
    return order[order >= 0]


def border_targets(cc, count, anisotropy):
    cc = np.asfortranarray(np.asarray(cc).astype(np.uint32, copy=False))
    cap = 4 * (cc.shape[0] * cc.shape[1] + cc.shape[1] * cc.shape[2] + cc.shape[0] * cc.shape[2]) + 16
    pts = np.zeros(3 * cap, np.int64)
    offs = np.zeros(count + 2, np.int64)
    an = np.asarray(anisotropy, np.float32)
    total = lib().oracle_border_targets(_ptr(cc), *cc.shape, count, _ptr(an), _ptr(pts), cap, _ptr(offs))
    if total < 0:
        _fail("border_targets")
    pts = pts[:3 * total].reshape(-1, 3)
    return [[tuple(int(c) for c in p) for p in pts[offs[i]:offs[i + 1]]] for i in range(count + 1)]


def geometric_flood(cc, anisotropy, sources):
    cc = np.asfortranarray(np.asarray(cc).astype(np.uint32, copy=False))
    out = np.zeros(cc.shape, np.float32, order="F")
    src = np.ascontiguousarray(np.asarray(sources, np.int64))
    an = np.asarray(anisotropy, np.float32)
    if lib().oracle_geometric_flood(_ptr(cc), *cc.shape, _ptr(an), _ptr(src), len(src), _ptr(out)):
        _fail("geometric_flood")
    return out


def target_order(cc, daf):
    """Per-id target orders of a component volume (ids ascending, each the id's voxels by DAF
    descending with +-inf and +-0 as 0, ties by flat index descending), concatenated: flat indices."""
    cc = np.asfortranarray(np.asarray(cc).astype(np.uint32, copy=False))
    daf = np.asfortranarray(np.asarray(daf, np.float32))
    out = np.zeros(max(int(np.count_nonzero(cc)), 1), np.int64)
    n = lib().oracle_target_order(_ptr(cc), cc.size, _ptr(daf), _ptr(out))
    if n < 0:
        _fail("target_order")
    return out[:n]


def invalidate(cc, dbf, alive, path, anisotropy, scale, constant):
    """One rolling-ball invalidation of `path` (flat indices) in the label of path[0]:
    (claimed count, alive after, claim keys: uint64 (bottleneck bits << 32) | source, all ones where
    unclaimed)."""
    cc = np.asfortranarray(np.asarray(cc).astype(np.uint32, copy=False))
    dbf = np.asfortranarray(np.asarray(dbf, np.float32))
    alive = np.asfortranarray(np.asarray(alive).astype(np.uint8))
    p = np.ascontiguousarray(np.asarray(path, np.int64))
    an = np.asarray(anisotropy, np.float32)
    claims = np.zeros(cc.shape, np.uint64, order="F")
    n = lib().oracle_invalidate(_ptr(cc), *cc.shape, _ptr(dbf), _ptr(alive), _ptr(p), len(p), _ptr(an),
                                float(scale), float(constant), _ptr(claims))
    if n < 0:
        _fail("invalidate")
    return n, alive, claims


def analyze(cc, count):
    """Per-id boxes and counts: (lo (count+1, 3), hi (count+1, 3), counts (count+1))."""
    cc = np.asfortranarray(np.asarray(cc).astype(np.uint32, copy=False))
    lo = np.zeros((count + 1, 3), np.int64)
    hi = np.zeros((count + 1, 3), np.int64)
    counts = np.zeros(count + 1, np.int64)
    if lib().oracle_analyze(_ptr(cc), *cc.shape, count, _ptr(lo), _ptr(hi), _ptr(counts)):
        _fail("analyze")
    return lo, hi, counts


def railroad(field, target, mask=None):
    field = np.asfortranarray(np.asarray(field, np.float32))
    cc = np.asfortranarray((np.ones(field.shape, np.uint32) if mask is None else np.asarray(mask) != 0).astype(np.uint32))
    path = np.zeros(field.size + 1, np.int64)
    n = lib().oracle_railroad(_ptr(field), _ptr(cc), *field.shape, int(target), _ptr(path), len(path))
    if n < 0:
        _fail("railroad")
    return path[:n].tolist()


def prepare(labels, teasar_params=None, anisotropy=(1, 1, 1), dust_threshold=1000, fix_borders=True, threads=1,
            fill_holes=False, fix_avocados=False):
    a, shape = _fortran_int64(labels)
    n = a.size
    cc = np.zeros(shape, np.uint32, order="F")
    dbf, daf, pdrf = (np.zeros(shape, np.float32, order="F") for _ in range(3))
    cap = n + 2
    active = np.zeros(cap, np.uint8)
    dbf_max, max_daf = np.zeros(cap, np.float32), np.zeros(cap, np.float32)
    root, target = np.zeros(cap, np.int64), np.zeros(cap, np.int64)
    soma = np.zeros(cap, np.uint8)
    pv = params_vector(teasar_params)
    an = np.asarray(anisotropy, np.float32)
    flags = int(bool(fix_borders)) | (2 if fill_holes else 0) | (4 if fix_avocados else 0)
    count = lib().oracle_prepare(_ptr(a), *shape, _ptr(pv), _ptr(an), float(dust_threshold), flags, threads,
                                 _ptr(cc), _ptr(dbf), _ptr(daf), _ptr(pdrf), cap, _ptr(active), _ptr(dbf_max),
                                 _ptr(max_daf), _ptr(root), _ptr(target), _ptr(soma))
    if count < 0:
        _fail("prepare")
    k = count + 1
    return dict(cc=cc, dbf=dbf, daf=daf, pdrf=pdrf, count=count, active=active[:k].copy(), dbf_max=dbf_max[:k].copy(),
                max_daf=max_daf[:k].copy(), root=root[:k].copy(), target=target[:k].copy(), soma=soma[:k].copy())


def vw_parents(labels, vw, sources):
    """Vertex-weighted flood over the float32 weights `vw` from the flat `sources` (label-gated by the
    26-connected components of `labels`) and the parent field: (dist f32, parent int64), Fortran
    order."""
    a, shape = _fortran_int64(labels)
    w = np.asfortranarray(np.asarray(vw, np.float32))
    src = np.ascontiguousarray(np.asarray(sources, np.int64))
    dist = np.zeros(shape, np.float32, order="F")
    parent = np.zeros(shape, np.int64, order="F")
    if lib().oracle_vw_parents(_ptr(a), *shape, _ptr(w), _ptr(src), len(src), _ptr(dist), _ptr(parent)) != 0:
        _fail("vw_parents")
    return dist, parent


def skeletonize(labels, teasar_params=None, anisotropy=(1, 1, 1), dust_threshold=1000, fix_branching=True,
                fix_borders=True, extra_targets_before=(), threads=1, stop_emitted=True, fill_holes=False,
                fix_avocados=False, extra_targets_after=(), voxel_graph=None):
    """Packed result dict: labels, v_off, e_off, vertices (V,3) f32, edges (E,2) u32, radii f32, stats.
    fix_branching=False walks the parent field; stop_emitted=False makes every walk run to the root
    (Kimimaro's paths) instead of ending on the first emitted vertex."""
    a, shape = _fortran_int64(labels)
    g, compact = _graph(voxel_graph, shape) if voxel_graph is not None else (None, False)
    pv = params_vector(teasar_params)
    mp = (teasar_params or {}).get("max_paths")
    an = np.asarray(anisotropy, np.float32)
    tb = np.ascontiguousarray(np.asarray(list(extra_targets_before), np.int64).reshape(-1, 3))
    ta = np.ascontiguousarray(np.asarray(list(extra_targets_after), np.int64).reshape(-1, 3))
    L = lib()
    h = L.oracle_skeletonize(_ptr(a), *shape, _ptr(pv), -1 if mp is None else int(mp), _ptr(an), float(dust_threshold),
                             1 if fix_branching else (0 if stop_emitted else -1),
                             int(bool(fix_borders)) | (2 if fill_holes else 0) | (4 if fix_avocados else 0), _ptr(tb),
                             len(tb), _ptr(ta), len(ta), threads, None if g is None else _ptr(g), int(compact))
    if not h:
        _fail("skeletonize")
    try:
        sizes = np.zeros(6, np.int64)
        L.oracle_result_sizes(h, _ptr(sizes))
        ns, nv, ne = (int(x) for x in sizes[:3])
        out = dict(labels=np.zeros(ns, np.int64), v_off=np.zeros(ns + 1, np.int64), e_off=np.zeros(ns + 1, np.int64),
                   vertices=np.zeros((nv, 3), np.float32), edges=np.zeros((ne, 2), np.uint32),
                   radii=np.zeros(nv, np.float32))
        L.oracle_result_copy(h, _ptr(out["labels"]), _ptr(out["v_off"]), _ptr(out["e_off"]), _ptr(out["vertices"]),
                             _ptr(out["edges"]), _ptr(out["radii"]))
        out["stats"] = dict(paths=int(sizes[3]), components=int(sizes[4]), invalidated=int(sizes[5]))
        return out
    finally:
        L.oracle_result_free(h)

from .merge import merge_fragments  # noqa: E402,F401

_POST_STEPS = {"postprocess": 0, "remove_dust": 1, "remove_loops": 2, "join_components": 3, "remove_ticks": 4,
               "prune_unused": 5}


def _packed_arrays(p):
    labels = np.ascontiguousarray(np.asarray(p["labels"], np.int64))
    voff = np.ascontiguousarray(np.asarray(p["v_off"], np.int64))
    eoff = np.ascontiguousarray(np.asarray(p["e_off"], np.int64))
    verts = np.ascontiguousarray(np.asarray(p["vertices"], np.float32).reshape(-1, 3))
    edges = np.ascontiguousarray(np.asarray(p["edges"], np.uint32).reshape(-1, 2))
    radii = np.ascontiguousarray(np.asarray(p["radii"], np.float32))
    return labels, voff, eoff, verts, edges, radii


def post(step, packed, a=0.0, b=0.0):
    """The canonical postprocess of a packed result (dict of labels, v_off, e_off, vertices,
    edges, radii), one step or all of them: postprocess(a = dust, b = ticks), remove_dust(a),
    remove_loops, join_components, remove_ticks(a), prune_unused. Returns a packed dict; raises
    ValueError on an invalid packed graph."""
    arrays = _packed_arrays(packed)
    L = lib()
    h = L.oracle_post(_POST_STEPS[step], len(arrays[0]), *(_ptr(x) for x in arrays), float(a), float(b))
    if not h:
        raise ValueError(L.oracle_post_error().decode())
    try:
        sizes = np.zeros(3, np.int64)
        L.oracle_post_sizes(h, _ptr(sizes))
        k, nv, ne = (int(x) for x in sizes)
        out = dict(labels=np.zeros(k, np.int64), v_off=np.zeros(k + 1, np.int64), e_off=np.zeros(k + 1, np.int64),
                   vertices=np.zeros((nv, 3), np.float32), edges=np.zeros((ne, 2), np.uint32),
                   radii=np.zeros(nv, np.float32))
        L.oracle_post_copy(h, _ptr(out["labels"]), _ptr(out["v_off"]), _ptr(out["e_off"]), _ptr(out["vertices"]),
                           _ptr(out["edges"]), _ptr(out["radii"]))
        return out
    finally:
        L.oracle_post_free(h)


def post_components(packed):
    """Per vertex of a packed result: the least vertex index of its connected component."""
    arrays = _packed_arrays(packed)
    out = np.zeros(len(arrays[5]), np.int64)
    if lib().oracle_post_components(len(arrays[0]), *(_ptr(x) for x in arrays), _ptr(out)) != 0:
        raise ValueError(lib().oracle_post_error().decode())
    return out


def pairwise_sum(values):
    """NumPy's pairwise float32 sum, as the dust rule uses it."""
    v = np.ascontiguousarray(np.asarray(values, np.float32))
    return np.float32(lib().oracle_pairwise_sum(_ptr(v), len(v)))
