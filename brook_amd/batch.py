"""`skeletonize_batch`: independent samples in one call (the reference's batch.py:193-229);
result[i] equals skeletonize(samples[i], ...)."""
from __future__ import annotations

import ctypes

from . import _native as native
from ._intake import (DEFAULT_TEASAR_PARAMS, _packed_from_handle, fence, graph_struct, options_struct, stream_handle,
                      volume_struct)
from ._packed import BatchedSkeletons, PackedSkeletons


def _split(samples):
    try:
        import torch
        if isinstance(samples, torch.Tensor):
            return [samples[i] for i in range(samples.shape[0])]
    except ImportError:
        pass
    if hasattr(samples, "ndim") and getattr(samples, "ndim", 0) >= 2 and not isinstance(samples, (list, tuple)):
        return [samples[i] for i in range(samples.shape[0])]
    return list(samples)


def _per_sample(entries, count, name):
    if entries is None:
        return [[] for _ in range(count)]
    entries = list(entries)
    if len(entries) != count:
        raise ValueError(f"brook: {name} has {len(entries)} entries for {count} samples")
    return [[] if e is None else [tuple(int(c) for c in p) for p in e] for e in entries]


def batch_parts(items, teasar_params, anisotropy, dust_threshold, fix_branching=True, fix_borders=True,
                fill_holes=False, fix_avocados=False, before=None, after=None, stream=None, graphs=None):
    """One PackedSkeletons per sample (host arrays): samples without per-sample targets share one
    engine batch (one flood and one trace launch per group); otherwise one engine call each.
    Device samples are fenced on `stream` (see _intake.stream_handle). graphs: one voxel graph or
    None per sample (None: none at all)."""
    before = before or [[] for _ in items]
    after = after or [[] for _ in items]
    keep, devices = [], []
    vols = [volume_struct(s, keep, devices)[0] for s in items]
    gvols = [None if g is None else graph_struct(g, s, keep, devices) for s, g in zip(items, graphs or [None] * len(items))]
    ctx = native.context(devices[0].device.index if devices else None)
    for handle in sorted({stream_handle(stream, t) for t in devices}):
        fence(ctx, handle)
    if all(not b for b in before) and all(not a for a in after):
        if not items:
            return []
        opts = options_struct(teasar_params, anisotropy, dust_threshold, fix_branching, fix_borders, fill_holes,
                              fix_avocados, None, (), (), keep)
        arr = (ctypes.POINTER(native.Volume) * len(vols))(*[ctypes.pointer(v) for v in vols])
        handles = (ctypes.c_void_p * len(vols))()
        if any(g is not None for g in gvols):
            garr = (ctypes.POINTER(native.Volume) * len(vols))(
                *[ctypes.POINTER(native.Volume)() if g is None else ctypes.pointer(g) for g in gvols])
            native.check(native.lib().brook_skeletonize_batch_graphs(ctx.handle, arr, garr, len(vols),
                                                                     ctypes.byref(opts), handles))
        else:
            native.check(native.lib().brook_skeletonize_batch(ctx.handle, arr, len(vols), ctypes.byref(opts), handles))
        return [_packed_from_handle(ctypes.c_void_p(h)) for h in handles]
    parts = []
    for vol, g, b, a in zip(vols, gvols, before, after):
        opts = options_struct(teasar_params, anisotropy, dust_threshold, fix_branching, fix_borders, fill_holes,
                              fix_avocados, None, b, a, keep)
        if g is not None:
            opts.voxel_graph = ctypes.addressof(g)
        h = ctypes.c_void_p()
        native.check(native.lib().brook_skeletonize(ctx.handle, ctypes.byref(vol), ctypes.byref(opts), ctypes.byref(h)))
        parts.append(_packed_from_handle(h))
    return parts


def skeletonize_batch(samples, teasar_params=DEFAULT_TEASAR_PARAMS, anisotropy=(1, 1, 1), dust_threshold=1000,
                      fix_branching=True, fix_borders=True, fill_holes=False, fix_avocados=False,
                      extra_targets_before=None, extra_targets_after=None, progress=False, output="packed_device",
                      stream=None, borrow=True, voxel_graph=None):
    """Skeletonize independent label volumes (shapes may differ) in one call. Returns
    BatchedSkeletons (device tensors for "packed_device", host arrays for "packed") or, for
    output="skeletons", a list of `{label: skeleton}` dicts."""
    if output not in ("packed_device", "packed", "skeletons"):
        raise ValueError("brook: output must be 'packed_device', 'packed' or 'skeletons'")
    items = _split(samples)
    graphs = None
    if voxel_graph is not None:
        graphs = _split(voxel_graph)
        if len(graphs) != len(items):
            raise ValueError(f"brook: voxel_graph has {len(graphs)} entries for {len(items)} samples")
    before = _per_sample(extra_targets_before, len(items), "extra_targets_before")
    after = _per_sample(extra_targets_after, len(items), "extra_targets_after")
    anisotropy = tuple(float(a) for a in anisotropy)
    parts = batch_parts(items, teasar_params, anisotropy, dust_threshold, fix_branching, fix_borders, fill_holes,
                        fix_avocados, before, after, stream, graphs)
    parts = [p if len(p) else PackedSkeletons.from_skeletons({}, anisotropy) for p in parts]
    if output == "skeletons":
        return [p.to_skeletons() for p in parts]
    result = BatchedSkeletons.from_parts(parts)
    if output == "packed_device":
        return BatchedSkeletons.from_parts([p.to_device() if len(p) else p for p in parts])
    return result
