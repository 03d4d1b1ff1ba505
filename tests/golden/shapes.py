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
