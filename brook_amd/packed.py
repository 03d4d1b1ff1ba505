"""Packed results and their merge on the GPU (the reference's packed.py:1-6, _packed.py:98-112)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as native
from ._intake import _packed_from_handle
from ._packed import PackedSkeletons

__all__ = ["PackedSkeletons", "merge_fragments"]


def _view(part, keep):
    """A brook_skeletons_view of `part`; the arrays it points into are appended to `keep`."""
    if part.on_device:
        import torch
        vertices = part.vertices.to(torch.float32).contiguous()
        edges = part.edges.to(torch.int32).contiguous()  # local indices < 2**31: the uint32 bits
        radii = part.radii.to(torch.float32).contiguous()
        torch.cuda.current_stream(vertices.device).synchronize()  # device arrays must be ready
        pointers = [vertices.data_ptr(), edges.data_ptr(), radii.data_ptr()]
        nv, ne = int(vertices.shape[0]), int(edges.shape[0])
    else:
        vertices = np.ascontiguousarray(part.vertices, np.float32).reshape(-1, 3)
        edges = np.ascontiguousarray(part.edges, np.uint32).reshape(-1, 2)
        radii = np.ascontiguousarray(part.radii, np.float32).reshape(-1)
        pointers = [vertices.ctypes.data, edges.ctypes.data, radii.ctypes.data]
        nv, ne = len(vertices), len(edges)
    labels = np.ascontiguousarray(part.labels, np.int64)
    v_off = np.ascontiguousarray(part.v_off, np.int64)
    e_off = np.ascontiguousarray(part.e_off, np.int64)
    if len(v_off) != len(labels) + 1 or len(e_off) != len(labels) + 1 or len(radii) != nv:
        raise ValueError("invalid packed skeleton arrays")
    keep.extend([vertices, edges, radii, labels, v_off, e_off])
    v = native.SkeletonsView()
    v.skeleton_count, v.vertex_count, v.edge_count = len(labels), nv, ne
    v.labels, v.vertex_offsets, v.edge_offsets = labels.ctypes.data, v_off.ctypes.data, e_off.ctypes.data
    v.vertices, v.edges, v.radii = pointers
    for a in range(3):
        v.anisotropy[a] = float(part.anisotropy[a])
    return v


def merge_fragments(parts, origins, on_device=False):
    """Merge `PackedSkeletons` chunk results placed at `origins` (one (x, y, z) per part, physical
    units) on the GPU: one vertex per (label, position) keeping the first radius, edges remapped
    without self-loops or duplicates, vertices on no edge dropped, labels ascending. Parts may be
    on the host or on the GPU; the result stays on the GPU (torch tensors) when `on_device`."""
    parts = list(parts)
    origins = list(origins)
    if len(parts) != len(origins):
        raise ValueError("fragment/origin count mismatch")
    aniso = parts[0].anisotropy if parts else (1.0, 1.0, 1.0)
    if not parts or not any(len(p) for p in parts):
        return PackedSkeletons.from_skeletons({}, aniso)
    at = np.ascontiguousarray(np.asarray(origins, np.float32).reshape(len(parts), 3))
    device = next((p.vertices.device for p in parts if p.on_device), None)
    ctx = native.context(None if device is None else device.index)
    keep = []
    views = (native.SkeletonsView * len(parts))(*[_view(p, keep) for p in parts])
    handle = ctypes.c_void_p()
    native.check(native.lib().brook_merge_views(ctx.handle, views, at.ctypes.data, len(parts), ctypes.byref(handle)))
    result = _packed_from_handle(handle)
    return result.to_device(device or "cuda") if on_device else result
