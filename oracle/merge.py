"""TEST INFRASTRUCTURE ONLY: numpy restatement of the reference's packed merge.

merge_fragments(parts, origins) follows packed.cpp:66-86 (each fragment's skeletons become parts
offset by the fragment's origin) and packed_merge.cu:88-166 (merge_parts with sorted labels and
empty labels kept):
  - labels of the skeletons with vertices, ascending (packed_merge.cu:91-98);
  - vertices shifted by float32 addition of the origin, ordered by (label, x, y, z) with input
    order breaking ties, -0 equal to +0 and NaN above every number (packed_merge.cu:142-146, the
    float keys of :33-37); one merged vertex per run of equal (label, x, y, z), a NaN coordinate
    equal to nothing (:42-49), carrying the run's first vertex's position and radius;
  - edges as (min id, max id) of merged ids, self-loops dropped, duplicates removed, sorted
    (:57-64, :151-154); out-of-range local endpoints raise ValueError (:60, :155);
  - vertices used by some edge kept in id order (:65-74, :157-165); edge endpoints local to
    their label (:82-86).
Parts are dicts (or objects) with labels, v_off, e_off, vertices (n x 3 float32), edges (m x 2
uint32), radii (n float32) in host memory.
"""
from __future__ import annotations

import numpy as np


def _get(part, name):
    return part[name] if isinstance(part, dict) else getattr(part, name)


def merge_fragments(parts, origins):
    parts = list(parts)
    if len(parts) != len(origins):
        raise ValueError("fragment/origin count mismatch")
    label_set = set()
    rows = []  # (label, vertices, radii, edges) per skeleton with vertices, in input order
    for part, origin in zip(parts, origins):
        labels = np.asarray(_get(part, "labels"), np.int64)
        v_off = np.asarray(_get(part, "v_off"), np.int64)
        e_off = np.asarray(_get(part, "e_off"), np.int64)
        verts = np.asarray(_get(part, "vertices"), np.float32).reshape(-1, 3)
        edges = np.asarray(_get(part, "edges"), np.uint32).reshape(-1, 2)
        radii = np.asarray(_get(part, "radii"), np.float32).reshape(-1)
        if v_off[0] != 0 or e_off[0] != 0:
            raise ValueError("invalid packed offsets")
        o = np.asarray(origin, np.float32).reshape(3)
        for k, label in enumerate(labels.tolist()):
            v0, v1, e0, e1 = int(v_off[k]), int(v_off[k + 1]), int(e_off[k]), int(e_off[k + 1])
            if v1 < v0 or e1 < e0:
                raise ValueError("invalid packed offsets")
            if v1 == v0:
                if e1 > e0:
                    raise ValueError("edges without vertices")
                continue
            label_set.add(label)
            rows.append((label, verts[v0:v1] + o, radii[v0:v1], edges[e0:e1]))
    out_labels = np.array(sorted(label_set), np.int64)
    empty = dict(labels=out_labels, v_off=np.zeros(len(out_labels) + 1, np.int64),
                 e_off=np.zeros(len(out_labels) + 1, np.int64), vertices=np.zeros((0, 3), np.float32),
                 edges=np.zeros((0, 2), np.uint32), radii=np.zeros(0, np.float32))
    if not rows:
        return empty
    rank_of_label = {int(l): i for i, l in enumerate(out_labels.tolist())}
    rank = np.concatenate([np.full(len(r[1]), rank_of_label[r[0]], np.int64) for r in rows])
    pos = np.concatenate([r[1] for r in rows]).astype(np.float32)
    rad = np.concatenate([r[2] for r in rows]).astype(np.float32)
    base = np.cumsum([0] + [len(r[1]) for r in rows])
    global_edges = []
    for r, b in zip(rows, base[:-1]):
        e = r[3].astype(np.int64)
        if e.size and e.max() >= len(r[1]):
            raise ValueError("packed edge is out of bounds")
        global_edges.append(e + b)
    ge = np.concatenate(global_edges) if global_edges else np.zeros((0, 2), np.int64)
    n = len(pos)
    # stable order by (rank, x, y, z); numpy sorts -0 with +0 and NaN last
    order = np.lexsort((np.arange(n), pos[:, 2], pos[:, 1], pos[:, 0], rank))
    sp, sr = pos[order], rank[order]
    fresh = np.ones(n, bool)
    fresh[1:] = (sr[1:] != sr[:-1]) | np.any(sp[1:] != sp[:-1], axis=1)  # NaN != NaN
    ids_sorted = np.cumsum(fresh) - 1
    merged = np.empty(n, np.int64)
    merged[order] = ids_sorted
    first = order[fresh]  # input vertex of each merged id
    a, b = merged[ge[:, 0]], merged[ge[:, 1]]
    keep = a != b
    lo, hi = np.minimum(a, b)[keep], np.maximum(a, b)[keep]
    keys = np.unique((lo.astype(np.uint64) << np.uint64(32)) | hi.astype(np.uint64))
    ea, eb = (keys >> np.uint64(32)).astype(np.int64), (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    used = np.zeros(len(first), bool)
    used[ea] = True
    used[eb] = True
    slot = np.cumsum(used) - 1
    kept_first = first[used]
    kept_rank = rank[kept_first]
    ngroups = len(out_labels)
    vcount = np.bincount(kept_rank, minlength=ngroups)
    vbase = np.concatenate([[0], np.cumsum(vcount)])
    erank = rank[first[ea]]
    ecount = np.bincount(erank, minlength=ngroups)
    edges_out = np.stack([slot[ea] - vbase[erank], slot[eb] - vbase[erank]], axis=1).astype(np.uint32)
    return dict(labels=out_labels, v_off=vbase.astype(np.int64), e_off=np.concatenate([[0], np.cumsum(ecount)]).astype(np.int64),
                vertices=pos[kept_first].reshape(-1, 3), edges=edges_out.reshape(-1, 2), radii=rad[kept_first])
