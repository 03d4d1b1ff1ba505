"""Packed skeleton containers (the reference's _packed.py:13-95 and batch.py:15-99 layouts)."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np


class Skeleton:
    """A skeleton in physical space: the fields of osteoid.Skeleton that the path produces
    (used when osteoid is not installed)."""

    def __init__(self, vertices, edges, radii, segid, transform):
        self.vertices = vertices
        self.edges = edges
        self.radii = radii
        self.id = segid
        self.space = "physical"
        self.transform = transform

    def __repr__(self):
        return f"Skeleton(id={self.id}, vertices={len(self.vertices)}, edges={len(self.edges)})"


def make_skeleton(vertices, edges, radii, label, anisotropy):
    a = anisotropy
    transform = np.array([[a[0], 0, 0, 0], [0, a[1], 0, 0], [0, 0, a[2], 0]], np.float32)
    try:
        from osteoid import Skeleton as Osteoid
    except ImportError:
        return Skeleton(vertices, edges, radii, int(label), transform)
    return Osteoid(vertices, edges, radii=radii, segid=int(label), space="physical", transform=transform.copy())


@dataclass
class PackedSkeletons:
    """Skeletons in flat arrays: skeleton k has label labels[k], vertices[v_off[k]:v_off[k+1]]
    (float32, physical units), the same range of radii, and edges[e_off[k]:e_off[k+1]] (uint32,
    local to the skeleton). The arrays are torch tensors on the GPU when `on_device`."""

    labels: np.ndarray
    v_off: np.ndarray
    e_off: np.ndarray
    vertices: object
    edges: object
    radii: object
    anisotropy: tuple

    def __len__(self):
        return int(self.labels.size)

    @property
    def on_device(self):
        return not isinstance(self.vertices, np.ndarray)

    def to_host(self):
        if not self.on_device:
            return self
        return PackedSkeletons(self.labels, self.v_off, self.e_off, self.vertices.cpu().numpy(),
                               self.edges.cpu().numpy().astype(np.uint32), self.radii.cpu().numpy(), self.anisotropy)

    def to_device(self, device="cuda"):
        if self.on_device:
            return self
        import torch
        return PackedSkeletons(self.labels, self.v_off, self.e_off, torch.from_numpy(self.vertices).to(device),
                               torch.from_numpy(self.edges.astype(np.int64)).to(device),
                               torch.from_numpy(self.radii).to(device), self.anisotropy)

    def as_dict(self):
        h = self.to_host()
        return dict(labels=h.labels, v_off=h.v_off, e_off=h.e_off, vertices=h.vertices, edges=h.edges, radii=h.radii)

    def to_skeletons(self):
        """{label: skeleton} in physical space (osteoid.Skeleton when osteoid is installed)."""
        h = self.to_host()
        out = {}
        for k, label in enumerate(h.labels.tolist()):
            v0, v1, e0, e1 = int(h.v_off[k]), int(h.v_off[k + 1]), int(h.e_off[k]), int(h.e_off[k + 1])
            out[label] = make_skeleton(h.vertices[v0:v1], h.edges[e0:e1], h.radii[v0:v1], label, h.anisotropy)
        return out

    @classmethod
    def from_skeletons(cls, skeletons, anisotropy):
        labels = np.fromiter(skeletons, dtype=np.int64, count=len(skeletons))
        verts, edges, radii, voff, eoff = [], [], [], [0], [0]
        for s in skeletons.values():
            verts.append(np.asarray(s.vertices, np.float32))
            edges.append(np.asarray(s.edges, np.uint32))
            radii.append(np.asarray(s.radii, np.float32))
            voff.append(voff[-1] + len(s.vertices))
            eoff.append(eoff[-1] + len(s.edges))
        cat = lambda parts, shape, dt: np.concatenate(parts).astype(dt, copy=False) if parts else np.empty(shape, dt)
        return cls(labels, np.asarray(voff, np.int64), np.asarray(eoff, np.int64), cat(verts, (0, 3), np.float32),
                   cat(edges, (0, 2), np.uint32), cat(radii, (0,), np.float32), tuple(float(a) for a in anisotropy))

    @classmethod
    def from_view(cls, view):
        """Copy a brook_skeletons_view into numpy arrays."""
        ns, nv, ne = int(view.skeleton_count), int(view.vertex_count), int(view.edge_count)

        def arr(ptr, count, dtype, shape):
            out = np.empty(shape, dtype)
            if count:
                ctypes.memmove(out.ctypes.data, ptr, count * np.dtype(dtype).itemsize)
            return out

        return cls(arr(view.labels, ns, np.int64, (ns,)), arr(view.vertex_offsets, ns + 1, np.int64, (ns + 1,)),
                   arr(view.edge_offsets, ns + 1, np.int64, (ns + 1,)), arr(view.vertices, 3 * nv, np.float32, (nv, 3)),
                   arr(view.edges, 2 * ne, np.uint32, (ne, 2)), arr(view.radii, nv, np.float32, (nv,)),
                   tuple(float(a) for a in view.anisotropy))


@dataclass
class BatchedSkeletons:
    """Every skeleton of every sample, packed in sample order: labels[k] is skeleton k's label in
    its own sample, sample[k] its sample index and s_off[i]:s_off[i+1] the skeleton range of
    sample i. Vertices are in each sample's own frame."""

    packed: PackedSkeletons
    labels: np.ndarray
    sample: np.ndarray
    s_off: np.ndarray
    parts: list = field(default_factory=list, repr=False)

    def __len__(self):
        return int(self.s_off.size) - 1

    vertices = property(lambda self: self.packed.vertices)
    edges = property(lambda self: self.packed.edges)
    radii = property(lambda self: self.packed.radii)
    v_off = property(lambda self: self.packed.v_off)
    e_off = property(lambda self: self.packed.e_off)
    on_device = property(lambda self: self.packed.on_device)

    def __getitem__(self, i):
        i = int(i)
        if i < 0:
            i += len(self)
        if not 0 <= i < len(self):
            raise IndexError(i)
        return self.parts[i]

    def to_host(self):
        if not self.on_device:
            return self
        return BatchedSkeletons.from_parts([p.to_host() for p in self.parts])

    def to_skeletons(self):
        return [p.to_skeletons() for p in self.parts]

    @classmethod
    def from_parts(cls, parts):
        parts = list(parts)
        aniso = parts[0].anisotropy if parts else (1.0, 1.0, 1.0)
        host = [p.to_host() for p in parts]
        labels = np.concatenate([p.labels for p in host]) if host else np.empty(0, np.int64)
        sample = np.concatenate([np.full(len(p), i, np.int64) for i, p in enumerate(host)]) if host else np.empty(0, np.int64)
        s_off = np.concatenate([[0], np.cumsum([len(p) for p in host])]).astype(np.int64)
        v_off, e_off, verts, edges, radii = [0], [0], [], [], []
        for p in host:
            v_off.extend((p.v_off[1:] + v_off[-1]).tolist())
            e_off.extend((p.e_off[1:] + e_off[-1]).tolist())
            verts.append(p.vertices)
            edges.append(p.edges)
            radii.append(p.radii)
        cat = lambda xs, shape, dt: np.concatenate(xs).astype(dt, copy=False) if xs else np.empty(shape, dt)
        packed = PackedSkeletons(labels, np.asarray(v_off, np.int64), np.asarray(e_off, np.int64),
                                 cat(verts, (0, 3), np.float32), cat(edges, (0, 2), np.uint32),
                                 cat(radii, (0,), np.float32), aniso)
        return cls(packed, labels, sample, s_off, parts)
