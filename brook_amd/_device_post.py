"""GPU postprocessing of `PackedSkeletons` (the reference's _device_post.py:1-49): dust, loops,
joins and ticks with the reference's tie rules, computed on the device (post.cu).

Results come back as `PackedSkeletons`, on the input's device when the input is on the device.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as native
from ._intake import _packed_from_handle
from .packed import _view

__all__ = ["components", "remove_dust", "remove_loops", "join_close_components", "remove_ticks", "postprocess"]

_STEPS = {"postprocess": 0, "remove_dust": 1, "remove_loops": 2, "join_components": 3, "remove_ticks": 4,
          "prune_unused": 5}


def _device(p):
    return p.vertices.device if p.on_device else None


def _context(p):
    device = _device(p)
    return native.context(None if device is None else device.index)


def _call(p, step, a=0.0, b=0.0):
    ctx = _context(p)
    keep = []
    view = _view(p, keep)
    handle = ctypes.c_void_p()
    native.check(native.lib().brook_postprocess_views(ctx.handle, ctypes.byref(view), _STEPS[step], float(a), float(b),
                                                      ctypes.byref(handle)))
    result = _packed_from_handle(handle)
    device = _device(p)
    return result.to_device(device) if device is not None else result


def components(p):
    """A connected-component id for every packed vertex: the least vertex index of its component
    (int64; a tensor on the input's device for device input)."""
    ctx = _context(p)
    keep = []
    view = _view(p, keep)
    out = np.zeros(int(view.vertex_count), np.int64)
    native.check(native.lib().brook_skeleton_components(ctx.handle, ctypes.byref(view), out.ctypes.data, len(out)))
    device = _device(p)
    if device is None:
        return out
    import torch
    return torch.from_numpy(out).to(device)


def remove_dust(p, dust_threshold):
    """Remove connected components whose cable length is below `dust_threshold` (physical units),
    on the GPU."""
    if dust_threshold == 0 or int(p.v_off[-1]) == 0:
        return p
    return _call(p, "remove_dust", dust_threshold)


def remove_loops(p):
    """Break the cycles of each skeleton on the GPU."""
    return _call(p, "remove_loops")


def join_close_components(p):
    """Join the connected components of each skeleton into one tree on the GPU."""
    return _call(p, "join_components")


def remove_ticks(p, tick_threshold):
    """Remove terminal branches shorter than `tick_threshold` (physical units), on the GPU."""
    return _call(p, "remove_ticks", tick_threshold)


def prune_unused(p):
    """Drop the vertices that are on no edge (the postprocess's last step)."""
    return _call(p, "prune_unused")


def postprocess(p, dust_threshold=1500, tick_threshold=3500):
    """GPU postprocessing of packed skeletons: dust, loops, joins and ticks, with the reference's
    tie rules. Returns new `PackedSkeletons`."""
    return _call(p, "postprocess", dust_threshold, tick_threshold)
