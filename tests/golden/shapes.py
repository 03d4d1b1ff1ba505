"""Input volumes of the reference's golden skeleton tests, regenerated here.

The reference's test suite (tests/python/test_package.py:2019-2050) defines its golden inputs
by short numpy expressions; the expected digests it holds (test_package.py:2079-2106, Kimimaro
5.8.1 + dijkstra3d 1.15.2) are kept in golden_digests.json beside this file. The expressions below
produce the same arrays (the digests would not match otherwise).
"""
from __future__ import annotations

import hashlib

import numpy as np


def sine_band(amplitude, period, width, phase=0.0, axis=0):
    """One sine band |y - (32 + A sin(2 pi x / P + phase))| < width in a 64 x 160 plane, one voxel
    thick along `axis`."""
    rows, cols = np.mgrid[:64, :160]
    inside = np.abs(rows - (32 + amplitude * np.sin(2 * np.pi * cols / period + phase))) < width
    return np.moveaxis(inside[None].astype(np.uint32), 0, axis)


def sine_bands():
    """Labels 1..3: three bands of one plane, each painted only where earlier ones left 0."""
    rows, cols = np.mgrid[:64, :160]
    plane = np.zeros((64, 160), np.uint32)
    specs = [(12, 4, 50, 1.7), (30, 6, 70, 2.2), (48, 5, 45, 1.5)]
    for label, (centre, amplitude, period, width) in enumerate(specs, 1):
        band = np.abs(rows - (centre + amplitude * np.sin(2 * np.pi * cols / period))) < width
        plane[band & (plane == 0)] = label
    return plane[None]


def spiral_band():
    """A one-voxel-thick spiral band of a 96 x 96 plane, turns 11 voxels apart."""
    rows, cols = np.mgrid[:96, :96]
    radius = np.hypot(rows - 47.3, cols - 48.6)
    angle = np.mod(np.arctan2(rows - 47.3, cols - 48.6), 2 * np.pi)
    turn = (radius - 6 - 5.5 * angle / np.pi) / 11.0
    band = (np.abs(turn - np.round(turn)) * 11.0 < 1.8) & (radius > 5) & (radius < 44)
    return band[None].astype(np.uint32)


def helix_tube(radius, turn_radius, pitch, length):
    """A tube of `radius` around a helix running along x."""
    n = int(2 * turn_radius + 2 * radius + 6)
    x = np.arange(length, dtype=np.float64)[:, None, None]
    y = np.arange(n, dtype=np.float64)[None, :, None]
    z = np.arange(n, dtype=np.float64)[None, None, :]
    cy = n / 2 + turn_radius * np.cos(2 * np.pi * x / pitch)
    cz = n / 2 + turn_radius * np.sin(2 * np.pi * x / pitch)
    return ((y - cy) ** 2 + (z - cz) ** 2 <= radius * radius).astype(np.uint32)


# Options of the golden cases (test_package.py:2052-2053).
GOLDEN_OPTIONS = dict(teasar_params=dict(scale=1.0, const=2.0), anisotropy=(4, 4, 4), dust_threshold=0,
                      fix_branching=True, fix_borders=True)


def golden_cases():
    """[(name, labels, options override)] of every golden digest case."""
    return [
        ("single_line_axis0", sine_band(6.0, 60.0, 1.6, axis=0), {}),
        ("single_line_axis1", sine_band(6.0, 60.0, 1.6, axis=1), {}),
        ("single_line_axis2", sine_band(6.0, 60.0, 1.6, axis=2), {}),
        ("bands", sine_bands(), {}),
        ("wide", sine_band(9.0, 80.0, 2.1, 0.7), {}),
        ("spiral", spiral_band(), {"anisotropy": (1, 1, 1)}),
        ("helix", helix_tube(1.9, 8.0, 28.0, 28), {}),
    ]


def digest(vertices, edges, radii):
    """Order-independent exact digest of one skeleton: its (vertex, radius) rows and its edges
    (the summary rule of test_package.py:1365-1373)."""
    rows = np.column_stack([np.asarray(vertices, np.float32), np.asarray(radii, np.float32)]).astype(np.float32)
    order = np.lexsort(rows.T[::-1])
    rank = np.empty(len(order), np.int64)
    rank[order] = np.arange(len(order))
    e = np.sort(rank[np.asarray(edges, np.int64)], axis=1).reshape(-1, 2)
    e = e[np.lexsort(e.T[::-1])]
    return hashlib.sha256(rows[order].tobytes() + e.tobytes()).hexdigest()[:16]


def summary_from_packed(packed):
    """{label: [vertices, edges, max radius, digest]} of a packed result dict."""
    out = {}
    for k, label in enumerate(np.asarray(packed["labels"]).tolist()):
        v0, v1 = int(packed["v_off"][k]), int(packed["v_off"][k + 1])
        e0, e1 = int(packed["e_off"][k]), int(packed["e_off"][k + 1])
        verts, edges, radii = packed["vertices"][v0:v1], packed["edges"][e0:e1], packed["radii"][v0:v1]
        out[int(label)] = [v1 - v0, e1 - e0, float(np.max(radii)), digest(verts, edges, radii)]
    return out


# ---- somata (the reference's test_package.py:347-367 input and variants) ----

SOMA_PARAMS = dict(scale=1.5, const=2, pdrf_scale=100000, pdrf_exponent=4, soma_detection_threshold=6,
                   soma_acceptance_threshold=9, soma_invalidation_scale=1.0, soma_invalidation_const=2)


def soma_ball(size=72, centre=36, radius=12):
    """A ball with six processes: three 3x3 tubes through the centre along x, y and z, from 2 to
    size - 2 (test_package.py:351-357)."""
    grid = np.indices((size,) * 3) - centre
    labels = ((grid ** 2).sum(axis=0) <= radius ** 2).astype(np.uint8)
    for axis in range(3):
        tube = [slice(centre - 1, centre + 2)] * 3
        tube[axis] = slice(2, size - 2)
        labels[tuple(tube)] = 1
    return np.asfortranarray(labels)


def soma_cells(seed=7, shape=(96, 80, 64)):
    """Several labels, some with an off-centre ball on a bent tube, some plain tubes, one ball
    touching a face: somata beside ordinary labels in one volume."""
    rng = np.random.default_rng(seed)
    out = np.zeros(shape, np.uint32)
    zz, yy, xx = None, None, None
    grid = np.indices(shape)
    for label in range(1, 7):
        c = np.array([rng.integers(12, s - 12) for s in shape])
        tube_axis = int(rng.integers(0, 3))
        r_tube = 1 + int(rng.integers(0, 2))
        d = np.abs(grid - c.reshape(3, 1, 1, 1))
        others = [a for a in range(3) if a != tube_axis]
        tube = (d[others[0]] <= r_tube) & (d[others[1]] <= r_tube)
        bend = grid[others[0]] - c[others[0]] - np.round(6 * np.sin(grid[tube_axis] / 9.0)).astype(int)
        tube = (np.abs(bend) <= r_tube) & (d[others[1]] <= r_tube)
        body = tube
        if label % 2:
            r = 10 + int(rng.integers(0, 4))
            body = body | ((((grid - c.reshape(3, 1, 1, 1)) ** 2).sum(axis=0)) <= r * r)
        out[body & (out == 0)] = label
    # a ball cut by the x = 0 face
    ball = (((grid - np.array([2, shape[1] // 2, shape[2] // 2]).reshape(3, 1, 1, 1)) ** 2).sum(axis=0)) <= 49
    out[ball & (out == 0)] = 9
    return np.asfortranarray(out)


def soma_with_void(size=40):
    """A ball with a one-voxel cavity at its centre and a tube: a soma label with an enclosed void."""
    grid = np.indices((size,) * 3) - size // 2
    labels = ((grid ** 2).sum(axis=0) <= 10 ** 2).astype(np.uint8)
    labels[size // 2 - 1:size // 2 + 2, size // 2 - 1:size // 2 + 2, 2:size - 2] = 1
    labels[size // 2, size // 2, size // 2 - 5] = 0
    return np.asfortranarray(labels)


def soma_cavity(size=48, nucleus=False, staircase=False):
    """A ball (radius 12) with a 5x5x5 cavity beside its centre and 3x3 processes along x, y and z
    through the whole volume, so the label's box is the volume. `nucleus`: a label 2 inside the
    cavity. `staircase`: a label 3 running diagonally from the cavity through the wall to the
    outside (26-connected, so part of it lies in the voids and part outside)."""
    c = size // 2
    grid = np.indices((size,) * 3) - c
    labels = ((grid ** 2).sum(axis=0) <= 12 ** 2).astype(np.uint8)
    for axis in range(3):
        tube = [slice(c - 1, c + 2)] * 3
        tube[axis] = slice(0, size)
        labels[tuple(tube)] = 1
    cav = (slice(c + 3, c + 8), slice(c - 2, c + 3), slice(c - 2, c + 3))
    labels[cav] = 0
    if nucleus:
        labels[c + 4:c + 7, c - 1:c + 2, c - 1:c + 2] = 2
    if staircase:
        for i in range(12):
            labels[c + 5 + i, c + i, c] = 3
    return np.asfortranarray(labels)


# ---- fill_holes (the reference's test_package.py:1422-1435 input, 1457-1458 options) ----

HOLE_PARAMS = dict(scale=1.5, const=2, pdrf_scale=100000, pdrf_exponent=4, soma_detection_threshold=100,
                   soma_acceptance_threshold=120, soma_invalidation_scale=1.0, soma_invalidation_const=0)


def painted_root():
    """Label 1's lowest voxels lie in label 2's cavity and it leaves through a diagonal step in the
    wall, so filling label 2's holes paints label 1's first voxel; a dust speck (label 3) lies
    between label 1's first voxel before and after the fill."""
    labels = np.zeros((22, 22, 26), np.uint32)
    labels[3:17, 3:17, 3:17] = 2
    labels[5:15, 5:15, 5:15] = 0
    labels[6:13, 6:13, 6:12] = 1
    labels[9, 9, 12:16] = 1
    labels[10, 10, 16:18] = 1
    labels[6:16, 6:16, 18:24] = 1
    labels[19, 19, 13] = 3
    return np.asfortranarray(labels)


# Kimimaro 5.8.1 summaries of painted_root() with HOLE_PARAMS, dust 10, fix_borders=False,
# fill_holes=True (test_package.py:1810-1811): key order [2, 1]; with branching correction the
# reference compares label 2 by its first three fields only. 1001 voxels are filled.
PAINTED_ROOT_EXPECTED = {
    False: {2: [38, 37, 7.0, "8850fde826c023e3"], 1: [24, 23, 3.1622776985168457, "bd25239c8e51a899"]},
    True: {2: [35, 34, 7.0], 1: [20, 19, 3.1622776985168457, "13904653524930cd"]},
}
PAINTED_ROOT_FILLED = 1001
