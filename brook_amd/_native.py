"""ctypes binding of the C ABI in include/brook.h (libbrook_amd.so, built in-tree).

The library is required: every call fails loudly when it is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
# BROOK_AMD_LIBRARY names another build of the same sources (brook_amd.tune times its variants so)
LIBRARY = os.environ.get("BROOK_AMD_LIBRARY") or os.path.join(_HERE, "libbrook_amd.so")

ABI_VERSION = 1
SUCCESS, INVALID_ARGUMENT, OUT_OF_MEMORY, DEVICE_ERROR, INTERNAL_ERROR = range(5)
U8, U16, U32, U64, I8, I16, I32, I64, F32, F64 = range(10)
HOST, DEVICE = 0, 1
FIX_BRANCHING, FIX_BORDERS, FILTER_OBJECTS, DEVICE_OUTPUT, FILL_HOLES, FIX_AVOCADOS = 1, 2, 4, 8, 16, 32

_c = ctypes
_P = _c.c_void_p


class Volume(_c.Structure):
    _fields_ = [("struct_size", _c.c_size_t), ("abi_version", _c.c_uint32), ("data", _P),
                ("shape", _c.c_int64 * 3), ("strides", _c.c_int64 * 3), ("dtype", _c.c_int), ("memory", _c.c_int)]


class TeasarParameters(_c.Structure):
    _fields_ = [("scale", _c.c_double), ("constant", _c.c_double), ("pdrf_scale", _c.c_double),
                ("pdrf_exponent", _c.c_double), ("soma_detection_threshold", _c.c_double),
                ("soma_acceptance_threshold", _c.c_double), ("soma_invalidation_scale", _c.c_double),
                ("soma_invalidation_const", _c.c_double), ("max_paths", _c.c_int64), ("has_max_paths", _c.c_uint32)]


class SkeletonizeOptions(_c.Structure):
    _fields_ = [("struct_size", _c.c_size_t), ("abi_version", _c.c_uint32), ("teasar", TeasarParameters),
                ("anisotropy", _c.c_float * 3), ("dust_threshold", _c.c_double), ("flags", _c.c_uint32),
                ("object_ids", _P), ("object_ids_count", _c.c_size_t),
                ("targets_before", _P), ("targets_before_count", _c.c_size_t),
                ("targets_after", _P), ("targets_after_count", _c.c_size_t), ("voxel_graph", _P)]


class SkeletonsView(_c.Structure):
    _fields_ = [("skeleton_count", _c.c_size_t), ("vertex_count", _c.c_size_t), ("edge_count", _c.c_size_t),
                ("labels", _P), ("vertex_offsets", _P), ("edge_offsets", _P), ("vertices", _P), ("edges", _P),
                ("radii", _P), ("anisotropy", _c.c_float * 3)]


class SkeletonsDeviceView(_c.Structure):
    _fields_ = [("skeleton_count", _c.c_size_t), ("vertex_count", _c.c_size_t), ("edge_count", _c.c_size_t),
                ("labels", _P), ("vertex_offsets", _P), ("edge_offsets", _P), ("vertices", Volume), ("edges", Volume),
                ("radii", Volume), ("anisotropy", _c.c_float * 3)]


VERTICES, EDGES, RADII = range(3)


class DeviceError(RuntimeError):
    """An error reported by the device engine (a HIP error or an engine failure)."""


_lib = None
_lock = threading.Lock()


def lib():
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIBRARY):
                raise RuntimeError(f"brook_amd: the native library {LIBRARY} is not built "
                                   "(run __graft_entry__.build()); there is no CPU fallback")
            L = _c.CDLL(LIBRARY)
            S = _c.c_int
            sig = {
                "brook_abi_version": (_c.c_uint32, []),
                "brook_version": (_c.c_char_p, []),
                "brook_last_error": (_c.c_char_p, []),
                "brook_context_create": (S, [_c.c_int, _c.POINTER(_P)]),
                "brook_context_destroy": (None, [_P]),
                "brook_context_synchronize": (S, [_P]),
                "brook_array_upload": (S, [_P, _c.POINTER(Volume), _c.POINTER(_P)]),
                "brook_array_destroy": (None, [_P]),
                "brook_array_view": (S, [_P, _c.POINTER(Volume)]),
                "brook_array_copy_to_host": (S, [_P, _P, _c.c_size_t]),
                "brook_connected_components": (S, [_P, _c.POINTER(Volume), _c.POINTER(_P)]),
                "brook_components_destroy": (None, [_P]),
                "brook_components_view": (S, [_P, _c.POINTER(Volume), _c.POINTER(_P), _c.POINTER(_P),
                                              _c.POINTER(_c.c_size_t)]),
                "brook_components_copy_to_host": (S, [_P, _P, _c.c_size_t]),
                "brook_components_labels": (S, [_P, _c.POINTER(_P)]),
                "brook_edt": (S, [_P, _c.POINTER(Volume), _P, _c.c_int, _c.POINTER(_P)]),
                "brook_graph_components": (S, [_P, _c.POINTER(Volume), _c.POINTER(Volume), _c.POINTER(_P)]),
                "brook_graph_edt": (S, [_P, _c.POINTER(Volume), _c.POINTER(Volume), _P, _c.c_int, _c.POINTER(_P)]),
                "brook_components_keys": (S, [_P, _c.POINTER(_P), _c.POINTER(_c.c_size_t)]),
                "brook_skeletonize_options_init": (None, [_c.POINTER(SkeletonizeOptions)]),
                "brook_skeletonize": (S, [_P, _c.POINTER(Volume), _c.POINTER(SkeletonizeOptions), _c.POINTER(_P)]),
                "brook_skeletonize_batch": (S, [_P, _c.POINTER(_c.POINTER(Volume)), _c.c_size_t,
                                                _c.POINTER(SkeletonizeOptions), _c.POINTER(_P)]),
                "brook_skeletonize_batch_graphs": (S, [_P, _c.POINTER(_c.POINTER(Volume)), _c.POINTER(_c.POINTER(Volume)),
                                                       _c.c_size_t, _c.POINTER(SkeletonizeOptions), _c.POINTER(_P)]),
                "brook_skeletons_get_view": (S, [_P, _c.POINTER(SkeletonsView)]),
                "brook_skeletons_destroy": (None, [_P]),
                "brook_skeletons_get_device_view": (S, [_P, _c.POINTER(SkeletonsDeviceView)]),
                "brook_skeletons_get_array": (S, [_P, _c.c_int, _c.POINTER(_P)]),
                "brook_merge_skeletons": (S, [_P, _c.POINTER(_P), _P, _c.c_size_t, _c.POINTER(_P)]),
                "brook_merge_views": (S, [_P, _c.POINTER(SkeletonsView), _P, _c.c_size_t, _c.POINTER(_P)]),
                "brook_postprocess_canonical": (S, [_P, _P, _c.c_double, _c.c_double, _c.POINTER(_P)]),
                "brook_postprocess_views": (S, [_P, _c.POINTER(SkeletonsView), _c.c_int, _c.c_double, _c.c_double,
                                                _c.POINTER(_P)]),
                "brook_skeleton_components": (S, [_P, _c.POINTER(SkeletonsView), _P, _c.c_size_t]),
                "brook_prepare_fields": (S, [_P, _c.POINTER(Volume), _c.POINTER(SkeletonizeOptions), _P, _P, _P,
                                             _c.c_size_t, _P, _P, _P, _P, _P, _P]),
                "brook_context_stat": (S, [_P, _c.c_char_p, _c.POINTER(_c.c_double)]),
                "brook_context_wait_stream": (S, [_P, _P]),
                "brook_array_to_dlpack": (S, [_P, _c.POINTER(_P)]),
                "brook_array_to_dlpack_versioned": (S, [_P, _c.c_uint32, _c.POINTER(_P)]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            if L.brook_abi_version() != ABI_VERSION:
                raise RuntimeError("brook_amd: native library ABI version mismatch")
            _lib = L
    return _lib


def check(status):
    if status == SUCCESS:
        return
    text = lib().brook_last_error().decode(errors="replace")
    if status == INVALID_ARGUMENT:
        if "not supported by this build" in text:
            raise NotImplementedError(text)
        raise ValueError(text)
    if status == OUT_OF_MEMORY:
        raise MemoryError(text)
    raise DeviceError(text)


class Context:
    """One device context (stream, cached device memory) of the engine."""

    def __init__(self, device=0):
        h = _P()
        check(lib().brook_context_create(int(device), _c.byref(h)))
        self.handle = h
        self.device = int(device)

    def close(self):
        if getattr(self, "handle", None):
            lib().brook_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stat(self, name):
        v = _c.c_double()
        check(lib().brook_context_stat(self.handle, name.encode(), _c.byref(v)))
        return v.value


_contexts = {}


def context(device=None):
    """The engine context of this process for `device` (default: torch's current device)."""
    if device is None:
        import torch
        device = torch.cuda.current_device() if torch.cuda.is_available() else 0
    key = (os.getpid(), threading.get_ident(), int(device))
    ctx = _contexts.get(key)
    if ctx is None:
        ctx = Context(device)
        _contexts[key] = ctx
    return ctx
