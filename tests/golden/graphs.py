"""Voxel-graph inputs of the reference's test suite, regenerated here, with the answers its tests hold.

The reference's tests (tests/python/test_package.py:1246-1453) build their voxel-graph cases from
numpy expressions and compare Brook with Kimimaro 5.8.1 results (test_package.py:1599-1830); the
expressions below produce the same arrays, and the expected summaries are data. A graph word has a
bit per 26-neighbour step in cc3d's convention: the step (dx, dy, dz) from a voxel is permitted
when bit STEP_BITS[(dx, dy, dz)] of its word is set.
"""
from __future__ import annotations

import numpy as np

# cc3d's voxel-connectivity-graph bit of each step (dx, dy, dz) along the array's own axes.
STEP_BITS = {(1, 0, 0): 0, (-1, 0, 0): 1, (0, 1, 0): 2, (0, -1, 0): 3, (0, 0, 1): 4, (0, 0, -1): 5,
             (1, 1, 0): 6, (-1, 1, 0): 7, (1, -1, 0): 8, (-1, -1, 0): 9,
             (1, 0, 1): 10, (-1, 0, 1): 11, (0, 1, 1): 12, (0, -1, 1): 13,
             (1, 0, -1): 14, (-1, 0, -1): 15, (0, 1, -1): 16, (0, -1, -1): 17,
             (1, 1, 1): 18, (-1, 1, 1): 19, (1, -1, 1): 20, (-1, -1, 1): 21,
             (1, 1, -1): 22, (-1, 1, -1): 23, (1, -1, -1): 24, (-1, -1, -1): 25}

# teasar_params of the reference's Kimimaro comparisons (test_package.py:1457-1458).
HOLE_PARAMS = dict(scale=1.5, const=2, pdrf_scale=100000, pdrf_exponent=4, soma_detection_threshold=100,
                   soma_acceptance_threshold=120, soma_invalidation_scale=1., soma_invalidation_const=0)
AVOCADO_JITTER = (0.15, 0.48, 0.33, -0.33, -0.24, 0.45, -0.59, 0.39)


def _slices(d, shape):
    """(source, destination) slices of the voxel pairs one step d apart."""
    src = tuple(slice(max(0, -s), shape[a] - max(0, s)) for a, s in enumerate(d))
    dst = tuple(slice(max(0, s), shape[a] - max(0, -s)) for a, s in enumerate(d))
    return src, dst


def voxel_graph(labels, side, region):
    """Steps between equal nonzero labels, except those crossing from `side` to its complement with
    both ends in `region` (test_package.py:1255-1267)."""
    graph = np.zeros(labels.shape, np.uint32)
    fg = labels != 0
    for d, bit in STEP_BITS.items():
        src, dst = _slices(d, labels.shape)
        cut = (side[src] != side[dst]) & region[src] & region[dst]
        ok = fg[src] & (labels[src] == labels[dst]) & ~cut
        graph[src] |= ok.astype(np.uint32) << np.uint32(bit)
    return graph


def open_graph(side):
    """Every step permitted, background and the steps out of the volume included, except across the
    cut (test_package.py:1339-1347)."""
    graph = np.full(side.shape, (1 << 26) - 1, np.uint32)
    for d, bit in STEP_BITS.items():
        src, dst = _slices(d, side.shape)
        graph[src] &= ~((side[src] != side[dst]).astype(np.uint32) << np.uint32(bit))
    return graph


def label_graph(labels, side):
    """cc3d.voxel_connectivity_graph(labels, 26): steps between equal labels, background included,
    the steps out of the volume set, without the steps across the cut (test_package.py:1350-1362)."""
    graph = np.zeros(labels.shape, np.uint32)
    for d, bit in STEP_BITS.items():
        src, dst = _slices(d, labels.shape)
        inside = np.zeros(labels.shape, bool)
        inside[src] = True
        graph[~inside] |= np.uint32(1 << bit)
        graph[src] |= ((labels[src] == labels[dst]) & (side[src] == side[dst])).astype(np.uint32) << np.uint32(bit)
    return graph


def graph_cases():
    """Self-touching volumes and the cut that separates the contact: [(labels, side, region)]
    (test_package.py:1270-1290)."""
    bars = np.zeros((40, 40, 24), np.uint32)
    bars[4:36, 8:16, 8:16] = 1
    bars[4:36, 16:24, 8:16] = 1
    side = np.zeros(bars.shape, bool)
    side[:, 16:, :] = True
    everywhere = np.ones(bars.shape, bool)
    partial = np.zeros(bars.shape, bool)
    partial[:30] = True
    ball = np.zeros((40, 40, 40), np.uint32)
    x, y, z = np.indices(ball.shape)
    ball[(x - 14) ** 2 + (y - 14) ** 2 + (z - 14) ** 2 <= 100] = 1
    ball[12:17, 12:17, 24:38] = 1
    ball[28:32, 4:36, 28:32] = 2
    process = np.zeros(ball.shape, bool)
    process[12:17, 12:17, 24:] = True
    return [(bars, side, everywhere), (bars.astype(np.uint16), side, partial), (ball, process, np.ones(ball.shape, bool))]


# Options of the graph cases (test_package.py:1290-1292).
GRAPH_OPTIONS = dict(teasar_params=dict(scale=1.5, const=2, pdrf_scale=100000, pdrf_exponent=4, soma_detection_threshold=6,
                                        soma_acceptance_threshold=8, soma_invalidation_scale=1.,
                                        soma_invalidation_const=0),
                     dust_threshold=10, fix_borders=True)


def cut_tube(phases=(0.538, 1.488, 5.035, 3.658, 0.591, 2.721)):
    """One label: a wavy tube along x cut in two at z = 21 by the graph, a wavy arm along y on each
    side, an enclosed cavity with a loose blob in the lower part and a cavity straddling the cut
    (test_package.py:1316-1336). Returns (labels, side of the cut)."""
    p = phases
    x, y, z = np.indices((40, 36, 44), dtype=np.float64)
    labels = np.zeros(x.shape, np.uint32)
    yc, zc = 17.3 + 2.9 * np.sin(0.17 * x + p[0]), 20.6 + 1.3 * np.cos(0.21 * x + p[1])
    labels[((y - yc) / 6.6) ** 2 + ((z - zc) / 9.5) ** 2 <= 1] = 5
    xa, za = 21.6 + 2.3 * np.sin(0.19 * y + p[2]), 13.2 + 1.4 * np.sin(0.27 * y + p[3])
    labels[((x - xa) / 3.7) ** 2 + ((z - za) / 3.3) ** 2 <= 1] = 5
    xb, zb = 16.4 + 2.1 * np.sin(0.16 * y + p[4]), 29.1 + 1.6 * np.cos(0.23 * y + p[5])
    labels[((x - xb) / 3.4) ** 2 + ((z - zb) / 3.9) ** 2 <= 1] = 5
    cx, cy = 8.6, 17.3 + 2.9 * np.sin(0.17 * 8.6 + p[0])
    cavity = ((x - cx) / 4.1) ** 2 + ((y - cy) / 3.4) ** 2 + ((z - 16.4) / 2.8) ** 2 <= 1
    blob = ((x - cx) / 2.0) ** 2 + ((y - cy) / 1.6) ** 2 + ((z - 16.4) / 1.1) ** 2 <= 1
    sx, sy = 30.2, 17.3 + 2.9 * np.sin(0.17 * 30.2 + p[0])
    labels[cavity & ~blob] = 0
    labels[((x - sx) / 3.1) ** 2 + ((y - sy) / 2.6) ** 2 + ((z - 20.7) / 2.5) ** 2 <= 1] = 0
    return np.asfortranarray(labels), z >= 21


def cut_tube_graphs():
    """The cut tube and its open, cut and labels graphs (test_package.py:1437-1443)."""
    labels, side = cut_tube()
    return labels, {"open": np.asfortranarray(open_graph(side)),
                    "cut": np.asfortranarray(voxel_graph(labels, side, np.ones(side.shape, bool))),
                    "labels": np.asfortranarray(label_graph(labels, side))}


def avocado(jitter=(0.0,) * 8):
    """A soma (label 3) clipped by the four side faces, a nucleus (label 8) poking out below it and a
    process leaving upwards (test_package.py:1381-1392)."""
    j = jitter
    x, y, z = np.indices((30, 28, 44), dtype=np.float64)
    labels = np.zeros(x.shape, np.uint32)
    labels[((x - 15.2 - j[0]) / 16.1) ** 2 + ((y - 13.9 - j[1]) / 15.3) ** 2 + ((z - 22.4 - j[2]) / 12.7) ** 2 <= 1] = 3
    labels[((x - 14.3 - j[3]) / 6.2) ** 2 + ((y - 13.6 - j[4]) / 5.8) ** 2 + ((z - 10.5 - j[5]) / 5.1) ** 2 <= 1] = 8
    px, py = 12.1 + j[6] + 1.7 * np.sin(0.21 * z), 15.4 + j[7] + 1.3 * np.cos(0.17 * z)
    labels[(((x - px) / 2.6) ** 2 + ((y - py) / 2.3) ** 2 <= 1) & (z >= 30)] = 3
    return np.asfortranarray(labels)


def avocado_order():
    """A soma around a nucleus, a void the graph joins to the soma and five dust specks first in
    Fortran order (test_package.py:1395-1409). Returns (labels, graph)."""
    x, y, z = np.indices((40, 36, 48), dtype=np.float64)
    labels = np.zeros(x.shape, np.uint32)
    for i in range(5):
        labels[3 + 7 * i:5 + 7 * i, 2:4, 1] = 20 + i
    labels[((x - 19.6) / 22) ** 2 + ((y - 17.8) / 20) ** 2 + ((z - 27.2) / 13.1) ** 2 <= 1] = 3
    labels[((x - 19.3) / 6.1) ** 2 + ((y - 17.6) / 5.7) ** 2 + ((z - 26.8) / 5.4) ** 2 <= 1] = 8
    graph = label_graph(labels, np.zeros(labels.shape, bool))
    labels[((x - 7.8) / 2.6) ** 2 + ((y - 17.9) / 2.4) ** 2 + ((z - 27.3) / 2.2) ** 2 <= 1] = 0
    return np.asfortranarray(labels), np.asfortranarray(graph)


def soma_voids():
    """A soma-sized component clipped by the side faces with an enclosed cavity, and its labels'
    graph (test_package.py:1412-1419). Returns (labels, graph)."""
    x, y, z = np.indices((26, 24, 30), dtype=np.float64)
    ball = ((x - 12.6) / 14.2) ** 2 + ((y - 11.7) / 13.1) ** 2 + ((z - 14.8) / 11.3) ** 2 <= 1
    cavity = ((x - 12.3) / 4.2) ** 2 + ((y - 11.4) / 3.6) ** 2 + ((z - 14.6) / 3.1) ** 2 <= 1
    labels = np.asfortranarray((ball & ~cavity).astype(np.uint32))
    return labels, np.asfortranarray(label_graph(labels, np.zeros(labels.shape, bool)))


def spanning_labels():
    """Labels 5 and 9 that the all-steps graph joins into one component whose last run starts in
    label 5 (test_package.py:1446-1453). Returns (labels, graph)."""
    labels = np.zeros((12, 6, 10), np.uint32)
    labels[1:6, 1:5, 5:9] = 5
    labels[6:11, 1:5, 1:9] = 9
    return np.asfortranarray(labels), np.asfortranarray(open_graph(np.zeros(labels.shape, bool)))


# Kimimaro 5.8.1 answers the reference's tests hold: {label: (vertices, edges, largest radius, digest)}.
SPANNING = {5: (12, 11, 2.0, "e4fa75811ac3b077")}                                   # test_package.py:1775
CUT_TUBE_FILLED = {"open": 146, "cut": 163, "labels": 163}                           # :1609
CUT_TUBE_RADIUS = {"open": (5.220153331756592, 6.164999961853027),                   # :1610-1611
                   "cut": (4.924428939819336, 6.112677097320557),
                   "labels": (4.924428939819336, 6.112677097320557)}
CUT_TUBE_DIGESTS = {("open", (1., 1., 1.)): (177, 175, 5.220153331756592, "79d57c10fa68bdb2"),  # :1612-1615
                    ("open", (1., 1.13, 1.37)): (219, 217, 6.164999961853027, "7282c1e0c830083c"),
                    ("cut", (1., 1.13, 1.37)): (376, 374, 6.112677097320557, "efa10b05120eddc4"),
                    ("labels", (1., 1.13, 1.37)): (315, 313, 6.112677097320557, "f09248b04e476d32")}
AVOCADO_PLAIN = {8: (9, 8, 4.609772205352783, "6048950761de63d6"), 3: (316, 315, 9.5, "922c9a29f61e3a8e")}  # :1651-1653
AVOCADO_ORDER_NUCLEUS = (11, 10, 5.464176177978516, "9693fc7b4f82a49c")                # :1686
AVOCADO_ORDER_SOMA = (580, 579, 9.179778099060059)
AVOCADO_ORDER_CORRECTED = {3: (573, 572, 9.179778099060059, "74eed02464c34728")}       # :1688
AVOCADO_ORDER_BRANCHED_RADIUS = 8.973200798034668                                       # :1694
AVOCADO_ORDER_KEYS = [0, 2, 3, 4, 5, 6, 7, 8]                                           # :1701
SOMA_VOID_RADIUS = {(1., 1., 1.): 5.7662811279296875, (1., 1.13, 1.37): 6.881685733795166}  # :1721
PAINTED_ROOT_GRAPH_RADIUS = {2: 1.5, 1: 2.5495097637176514}                             # :1821
