"""Device results of the C ABI as torch tensors, zero copy through DLPack (brook_array_to_dlpack):
the components and the EDT of a label volume stay in HBM. The tensor owns the export; the
library's memory is released when the tensor is (the export's deleter runs exactly once)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as native
from ._intake import fence, graph_struct, stream_handle, volume_struct

_capsule_new = ctypes.pythonapi.PyCapsule_New
_capsule_new.restype = ctypes.py_object
_capsule_new.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]


def array_to_torch(handle):
    """A torch tensor viewing the brook_array `handle` (which may be destroyed afterwards)."""
    import torch.utils.dlpack
    tensor = ctypes.c_void_p()
    native.check(native.lib().brook_array_to_dlpack(handle, ctypes.byref(tensor)))
    # torch takes the DLManagedTensor from the capsule and calls its deleter when it is done
    return torch.utils.dlpack.from_dlpack(_capsule_new(tensor, b"dltensor", None))


def _call(labels, stream, fn):
    keep, devices = [], []
    vol, _ = volume_struct(labels, keep, devices)
    ctx = native.context(devices[0].device.index if devices else None)
    if devices:
        fence(ctx, stream_handle(stream, devices[0]))
    return fn(ctx, vol)


def _components(lib, h):
    """(ids tensor, mapping, roots, keys) of a brook_components handle (destroyed here)."""
    try:
        mapping, roots, count = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_size_t()
        native.check(lib.brook_components_view(h, None, ctypes.byref(mapping), ctypes.byref(roots), ctypes.byref(count)))
        n = count.value
        m = np.ctypeslib.as_array(ctypes.cast(mapping, ctypes.POINTER(ctypes.c_int64)), (n + 1,)).copy()
        r = np.ctypeslib.as_array(ctypes.cast(roots, ctypes.POINTER(ctypes.c_int64)), (max(n, 1),))[:n].copy()
        keys, nk = ctypes.c_void_p(), ctypes.c_size_t()
        native.check(lib.brook_components_keys(h, ctypes.byref(keys), ctypes.byref(nk)))
        k = (np.ctypeslib.as_array(ctypes.cast(keys, ctypes.POINTER(ctypes.c_int64)), (nk.value,)).copy()
             if nk.value else np.zeros(0, np.int64))
        arr = ctypes.c_void_p()
        native.check(lib.brook_components_labels(h, ctypes.byref(arr)))
        try:
            ids = array_to_torch(arr)
        finally:
            lib.brook_array_destroy(arr)
    finally:
        lib.brook_components_destroy(h)
    return ids, m, r, k


def _call_graph(labels, graph, stream, fn):
    keep, devices = [], []
    vol, _ = volume_struct(labels, keep, devices)
    gvol = graph_struct(graph, labels, keep, devices)
    ctx = native.context(devices[0].device.index if devices else None)
    for handle in sorted({stream_handle(stream, t) for t in devices}):
        fence(ctx, handle)
    return fn(ctx, vol, gvol)


def graph_components(labels, voxel_graph, stream=None):
    """Components of a voxel graph (cc3d's color_connectivity_graph, as the reference's
    graph_components): (ids, mapping, roots, keys); keys[id] is the 1-based rank of the id's graph
    component among all of them, background ones included (Kimimaro's ids)."""
    def run(ctx, vol, gvol):
        lib = native.lib()
        h = ctypes.c_void_p()
        native.check(lib.brook_graph_components(ctx.handle, ctypes.byref(vol), ctypes.byref(gvol), ctypes.byref(h)))
        return _components(lib, h)
    return _call_graph(labels, voxel_graph, stream, run)


def graph_edt(labels, voxel_graph, anisotropy=(1, 1, 1), black_border=False, stream=None):
    """The graph-aware EDT of unsigned labels (float32 device tensor, Fortran strides)."""
    def run(ctx, vol, gvol):
        lib = native.lib()
        an = np.asarray(anisotropy, np.float32)
        h = ctypes.c_void_p()
        native.check(lib.brook_graph_edt(ctx.handle, ctypes.byref(vol), ctypes.byref(gvol), an.ctypes.data,
                                         int(bool(black_border)), ctypes.byref(h)))
        try:
            return array_to_torch(h)
        finally:
            lib.brook_array_destroy(h)
    return _call_graph(labels, voxel_graph, stream, run)


def connected_components(labels, stream=None):
    """26-connected components of a label volume (host array or device tensor): (ids, mapping,
    roots) with ids a device tensor of shape labels.shape in Fortran strides (the smallest unsigned
    dtype that holds the count), mapping[id] = input value, roots[id - 1] = first voxel (flat
    Fortran index)."""
    def run(ctx, vol):
        lib = native.lib()
        h = ctypes.c_void_p()
        native.check(lib.brook_connected_components(ctx.handle, ctypes.byref(vol), ctypes.byref(h)))
        try:
            mapping, roots, count = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_size_t()
            native.check(lib.brook_components_view(h, None, ctypes.byref(mapping), ctypes.byref(roots),
                                                   ctypes.byref(count)))
            n = count.value
            m = np.ctypeslib.as_array(ctypes.cast(mapping, ctypes.POINTER(ctypes.c_int64)), (n + 1,)).copy()
            r = np.ctypeslib.as_array(ctypes.cast(roots, ctypes.POINTER(ctypes.c_int64)), (max(n, 1),))[:n].copy()
            arr = ctypes.c_void_p()
            native.check(lib.brook_components_labels(h, ctypes.byref(arr)))
            try:
                ids = array_to_torch(arr)
            finally:
                lib.brook_array_destroy(arr)
        finally:
            lib.brook_components_destroy(h)
        return ids, m, r
    return _call(labels, stream, run)


def edt(labels, anisotropy=(1, 1, 1), black_border=False, stream=None):
    """Multi-label anisotropic EDT (float32 device tensor, Fortran strides, background 0)."""
    def run(ctx, vol):
        lib = native.lib()
        an = np.asarray(anisotropy, np.float32)
        h = ctypes.c_void_p()
        native.check(lib.brook_edt(ctx.handle, ctypes.byref(vol), an.ctypes.data, int(bool(black_border)),
                                   ctypes.byref(h)))
        try:
            return array_to_torch(h)
        finally:
            lib.brook_array_destroy(h)
    return _call(labels, stream, run)
