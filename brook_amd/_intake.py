"""`skeletonize`: Kimimaro's arguments in, skeletons out (the reference's _intake.py:99-205),
over the C ABI. Host arrays (numpy) and device tensors (torch on ROCm, or any DLPack producer on a
ROCm device) are accepted; device input is read in place when it is Fortran-contiguous."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as native
from ._packed import PackedSkeletons

# The reference's DEFAULT_TEASAR_PARAMS (src/brook/__init__.py:8-17).
DEFAULT_TEASAR_PARAMS = {
    "scale": 1.5,
    "const": 300,
    "pdrf_scale": 100000,
    "pdrf_exponent": 4,
    "soma_acceptance_threshold": 3500,
    "soma_detection_threshold": 750,
    "soma_invalidation_const": 300,
    "soma_invalidation_scale": 2,
}

# Keys a caller leaves out take kimimaro.trace's defaults (skeleton.hpp:24-25), not the above.
_TRACE_DEFAULTS = dict(scale=10.0, const=10.0, pdrf_scale=5000.0, pdrf_exponent=16.0, soma_detection_threshold=1100.0,
                       soma_acceptance_threshold=4000.0, soma_invalidation_scale=0.5, soma_invalidation_const=0.0)
_FIELD = dict(scale="scale", const="constant", pdrf_scale="pdrf_scale", pdrf_exponent="pdrf_exponent",
              soma_detection_threshold="soma_detection_threshold", soma_acceptance_threshold="soma_acceptance_threshold",
              soma_invalidation_scale="soma_invalidation_scale", soma_invalidation_const="soma_invalidation_const")

_DTYPES = {np.dtype(np.uint8): native.U8, np.dtype(np.uint16): native.U16, np.dtype(np.uint32): native.U32,
           np.dtype(np.uint64): native.U64, np.dtype(np.int8): native.I8, np.dtype(np.int16): native.I16,
           np.dtype(np.int32): native.I32, np.dtype(np.int64): native.I64}


class DimensionError(Exception):
    """The input volume has unsupported dimensions."""


def format_labels(labels):
    """Host labels as a 3-D numpy array (bool as uint8, trailing unit axes dropped, padded to 3-D)
    without rearranging contiguous bytes (_intake.py:23-36)."""
    labels = np.asarray(labels)
    if labels.dtype == bool:
        labels = labels.view(np.uint8)
    while labels.ndim > 3 and labels.shape[-1] == 1:
        labels = labels[..., 0]
    if labels.ndim > 3:
        raise DimensionError(f"{labels.ndim} dimensions is not supported (max 3).")
    while labels.ndim < 3:
        labels = labels[..., np.newaxis]
    return labels


def teasar_struct(params):
    t = native.TeasarParameters()
    values = dict(_TRACE_DEFAULTS)
    max_paths = None
    for key, value in dict(params or {}).items():
        if key == "max_paths":
            max_paths = value
            continue
        if key not in _FIELD:
            raise TypeError(f"unexpected TEASAR parameter: {key}")
        values[key] = float(value)
    for key, field_name in _FIELD.items():
        setattr(t, field_name, values[key])
    if max_paths is not None:
        t.max_paths = int(max_paths)
        t.has_max_paths = 1
    return t


def options_struct(teasar_params, anisotropy, dust_threshold, fix_branching, fix_borders, fill_holes, fix_avocados,
                   object_ids=None, extra_targets_before=(), extra_targets_after=(), keep=None):
    """brook_skeletonize_options; arrays it points to are appended to `keep`."""
    o = native.SkeletonizeOptions()
    native.lib().brook_skeletonize_options_init(ctypes.byref(o))
    o.teasar = teasar_struct(teasar_params)
    for i, a in enumerate(anisotropy):
        o.anisotropy[i] = float(a)
    o.dust_threshold = float(dust_threshold)
    flags = 0
    flags |= native.FIX_BRANCHING if fix_branching else 0
    flags |= native.FIX_BORDERS if fix_borders else 0
    flags |= native.FILL_HOLES if fill_holes else 0
    flags |= native.FIX_AVOCADOS if fix_avocados else 0
    keep = keep if keep is not None else []
    if object_ids is not None:
        ids = np.ascontiguousarray(np.asarray(list(object_ids), dtype=np.int64).reshape(-1))
        keep.append(ids)
        o.object_ids = ids.ctypes.data
        o.object_ids_count = ids.size
        flags |= native.FILTER_OBJECTS
    for name, pts in (("targets_before", extra_targets_before), ("targets_after", extra_targets_after)):
        arr = np.ascontiguousarray(np.asarray([tuple(int(c) for c in p) for p in (pts or ())], np.int64).reshape(-1, 3))
        keep.append(arr)
        setattr(o, name, arr.ctypes.data if arr.size else None)
        setattr(o, name + "_count", int(arr.shape[0]))
    o.flags = flags
    return o


def _is_device(obj):
    try:
        import torch
        if isinstance(obj, torch.Tensor):
            return obj.is_cuda
    except ImportError:
        pass
    dev = getattr(obj, "__dlpack_device__", None)
    return dev is not None and int(dev()[0]) in (2, 10)  # kDLCUDA / kDLROCM


def _device_tensor(obj):
    import torch
    if not isinstance(obj, torch.Tensor):
        obj = torch.from_dlpack(obj)
    if obj.dtype == torch.bool:
        obj = obj.to(torch.uint8)
    while obj.dim() > 3 and obj.shape[-1] == 1:
        obj = obj[..., 0]
    if obj.dim() > 3:
        raise DimensionError(f"{obj.dim()} dimensions is not supported (max 3).")
    while obj.dim() < 3:
        obj = obj.unsqueeze(-1)
    return obj


_TORCH_DTYPES = None


def _torch_dtype_code(t):
    global _TORCH_DTYPES
    import torch
    if _TORCH_DTYPES is None:
        _TORCH_DTYPES = {torch.uint8: native.U8, torch.int8: native.I8, torch.int16: native.I16, torch.int32: native.I32,
                         torch.int64: native.I64}
        for name, code in (("uint16", native.U16), ("uint32", native.U32), ("uint64", native.U64)):
            if hasattr(torch, name):
                _TORCH_DTYPES[getattr(torch, name)] = code
    if t.dtype not in _TORCH_DTYPES:
        raise TypeError(f"device labels must be integer or boolean, not {t.dtype}")
    return _TORCH_DTYPES[t.dtype]


def stream_handle(stream, tensor=None):
    """The producer stream device input waits for (the reference's stream argument,
    bindings.cpp:105-120): a torch.cuda.Stream, an integer handle (0 or 1: the null stream, 2: the
    per-thread default stream, else a hipStream_t), or None: torch's current stream on the
    tensor's device."""
    if stream is None:
        import torch
        device = tensor.device if tensor is not None else None
        return int(torch.cuda.current_stream(device).cuda_stream)
    handle = getattr(stream, "cuda_stream", stream)
    handle = int(handle)
    return 0 if handle == 1 else handle


def fence(ctx, handle):
    """The context's later work waits for the work queued so far on the producer stream."""
    native.check(native.lib().brook_context_wait_stream(ctx.handle, ctypes.c_void_p(handle or None)))


def volume_struct(labels, keep, streams=None):
    """brook_volume describing host or device labels (the object itself is kept alive in `keep`).
    For device labels the producer stream to wait for is appended to `streams` (see
    stream_handle); the caller fences the engine's context on it before the call."""
    v = native.Volume()
    v.struct_size = ctypes.sizeof(native.Volume)
    v.abi_version = native.ABI_VERSION
    if _is_device(labels):
        t = _device_tensor(labels)
        code = _torch_dtype_code(t)
        if streams is None:
            import torch
            torch.cuda.current_stream(t.device).synchronize()  # no fence requested: input made ready here
        else:
            streams.append(t)
        keep.append(t)
        v.data = t.data_ptr() if t.numel() else None
        item = t.element_size()
        for a in range(3):
            v.shape[a] = int(t.shape[a])
            v.strides[a] = int(t.stride(a)) * item
        v.dtype = code
        v.memory = native.DEVICE
        return v, t.numel()
    a = format_labels(labels)
    if a.dtype.kind == "f":
        raise NotImplementedError("floating-point labels are not supported by this build; pass integer labels")
    if a.dtype not in _DTYPES:
        a = a.astype(np.int64)
    keep.append(a)
    v.data = a.ctypes.data if a.size else None
    for i in range(3):
        v.shape[i] = int(a.shape[i])
        v.strides[i] = int(a.strides[i])
    v.dtype = _DTYPES[a.dtype]
    v.memory = native.HOST
    return v, a.size


def graph_struct(voxel_graph, labels, keep, streams=None):
    """brook_volume of a voxel graph (uint32 words, or uint8 for one XY plane; shape and layout
    checked by the engine). A device graph needs device labels, as the reference's intake requires
    (_intake.py:233-239)."""
    if _is_device(voxel_graph) and not _is_device(labels):
        raise TypeError("brook: a voxel_graph on the device needs device labels; pass the labels on the device too, "
                        "or the graph as a host array")
    v, _ = volume_struct(voxel_graph, keep, streams)
    keep.append(v)
    return v


def _packed_from_handle(handle):
    view = native.SkeletonsView()
    try:
        native.check(native.lib().brook_skeletons_get_view(handle, ctypes.byref(view)))
        return PackedSkeletons.from_view(view)
    finally:
        native.lib().brook_skeletons_destroy(handle)


def _empty(output, anisotropy):
    packed = PackedSkeletons.from_skeletons({}, anisotropy)
    return {} if output == "skeletons" else packed


def skeletonize(all_labels, teasar_params=DEFAULT_TEASAR_PARAMS, anisotropy=(1, 1, 1), object_ids=None,
                dust_threshold=1000, progress=True, fix_branching=True, in_place=False, fix_borders=True, parallel=1,
                parallel_chunk_size=100, extra_targets_before=(), extra_targets_after=(), fill_holes=False,
                fix_avocados=False, voxel_graph=None, streaming=None, output="skeletons", stream=None, borrow=True):
    """Skeletonize a 3-D label volume on one MI355X; a drop-in for `brook.skeletonize`.

    Returns `{label: skeleton}` (osteoid.Skeleton when osteoid is installed) by default,
    `PackedSkeletons` for output="packed" (host arrays) or "packed_device" (torch tensors on the GPU).
    Device input produced on another stream is fenced, not host-synchronized: the engine waits for
    `stream` (default: torch's current stream on the tensor's device; see stream_handle).
    `parallel`, `parallel_chunk_size`, `in_place`, `borrow` and `progress` are accepted for
    compatibility. Keys left out of `teasar_params` take kimimaro.trace's defaults. `voxel_graph`
    (uint32 words in cc3d's convention, or uint8 for one XY plane; host or device, device only with
    device labels) replaces the labels' own connectivity: components are the graph's, and every
    flood, path and invalidation steps only where the graph permits."""
    if output not in ("skeletons", "packed", "packed_device"):
        raise ValueError("brook: output must be 'skeletons', 'packed' or 'packed_device'")
    if streaming:
        raise NotImplementedError("streaming is not supported by this build (in-core on 288 GB of HBM)")
    anisotropy = tuple(float(np.float32(a)) for a in anisotropy)
    keep, devices = [], []
    vol, size = volume_struct(all_labels, keep, devices)
    if size <= dust_threshold:
        return _empty(output, anisotropy)
    opts = options_struct(teasar_params, anisotropy, dust_threshold, fix_branching, fix_borders, fill_holes,
                          fix_avocados, object_ids, extra_targets_before, extra_targets_after, keep)
    if voxel_graph is not None:
        opts.voxel_graph = ctypes.addressof(graph_struct(voxel_graph, all_labels, keep, devices))
    ctx = native.context(devices[0].device.index if devices else None)
    for handle in sorted({stream_handle(stream, t) for t in devices}):
        fence(ctx, handle)
    h = ctypes.c_void_p()
    native.check(native.lib().brook_skeletonize(ctx.handle, ctypes.byref(vol), ctypes.byref(opts), ctypes.byref(h)))
    packed = _packed_from_handle(h)
    if output == "skeletons":
        return packed.to_skeletons()
    if output == "packed_device" and len(packed):
        return packed.to_device()
    return packed


def prepare_fields(labels, teasar_params=DEFAULT_TEASAR_PARAMS, anisotropy=(1, 1, 1), dust_threshold=1000,
                   fix_borders=True, extra_targets_before=(), fill_holes=False, fix_avocados=False):
    """The preamble's fields on the host (diagnostics): dbf, daf, pdrf (Fortran-ordered float32
    volumes, pdrf before the roots become rails) and per-component tables."""
    keep = []
    vol, size = volume_struct(labels, keep)
    shape = tuple(int(vol.shape[a]) for a in range(3))
    opts = options_struct(teasar_params, anisotropy, dust_threshold, True, fix_borders, fill_holes, fix_avocados, None,
                          extra_targets_before, (), keep)
    dbf, daf, pdrf = (np.zeros(shape, np.float32, order="F") for _ in range(3))
    cap = size + 2
    traced = np.zeros(cap, np.uint8)
    root, target = np.zeros(cap, np.int64), np.zeros(cap, np.int64)
    dbf_max, max_daf = np.zeros(cap, np.float32), np.zeros(cap, np.float32)
    count = ctypes.c_int64()
    p = lambda a: a.ctypes.data
    native.check(native.lib().brook_prepare_fields(native.context().handle, ctypes.byref(vol), ctypes.byref(opts),
                                                   p(dbf), p(daf), p(pdrf), cap, p(traced), p(root), p(target),
                                                   p(dbf_max), p(max_daf), ctypes.byref(count)))
    k = count.value + 1
    return dict(dbf=dbf, daf=daf, pdrf=pdrf, count=count.value, active=traced[:k], root=root[:k], target=target[:k],
                dbf_max=dbf_max[:k], max_daf=max_daf[:k])
