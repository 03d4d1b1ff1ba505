"""brook_connected_components_noalloc (no device allocation: its working memory is a static device
array) against brook_connected_components, through the C ABI: ids, mapping, roots and count equal
on two calls in a row, the second on a smaller volume (the array keeps the first call's contents);
a table capacity below the count and a volume beyond the array's size are clean errors."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _device_volume(torch, labels):
    flat = torch.from_numpy(np.ravel(labels, order="F").copy()).cuda()
    sx, sy, sz = labels.shape
    return flat.view(sz, sy, sx).permute(2, 1, 0)


def _volume(t):
    from brook_amd import _native as native
    from brook_amd._intake import _torch_dtype_code
    v = native.Volume()
    v.struct_size = ctypes.sizeof(native.Volume)
    v.abi_version = native.ABI_VERSION
    v.data = t.data_ptr()
    for a in range(3):
        v.shape[a] = int(t.shape[a])
        v.strides[a] = int(t.stride(a)) * t.element_size()
    v.dtype = _torch_dtype_code(t)
    v.memory = native.DEVICE
    return v


def _reference(labels):
    from brook_amd import device
    ids, mapping, roots = device.connected_components(labels)
    return ids.cpu().numpy().astype(np.uint32), mapping, roots


def _noalloc(torch, dev, capacity):
    from brook_amd import _native as native
    L, ctx = native.lib(), native.context()
    vol = _volume(dev)
    ids = torch.empty(dev.numel(), dtype=torch.int32, device="cuda")
    mapping = np.zeros(capacity + 1, np.int64)
    roots = np.zeros(max(capacity, 1), np.int64)
    count = ctypes.c_size_t()
    status = L.brook_connected_components_noalloc(ctx.handle, ctypes.byref(vol), ctypes.c_void_p(ids.data_ptr()),
                                                 ctypes.c_size_t(ids.numel() * 4), mapping.ctypes.data,
                                                 roots.ctypes.data, ctypes.c_size_t(capacity), ctypes.byref(count))
    torch.cuda.synchronize()
    got = ids.cpu().numpy().view(np.uint32).reshape(dev.shape[::-1]).transpose(2, 1, 0)
    return status, count.value, got, mapping, roots


def _bind():
    from brook_amd import _native as native
    L = native.lib()
    L.brook_connected_components_noalloc.restype = ctypes.c_int
    L.brook_connected_components_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(native.Volume), ctypes.c_void_p,
                                                     ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    return native


def test_noalloc_equals_components_on_two_calls(torch):
    native = _bind()
    from brook_amd import synth
    rng = np.random.default_rng(7)
    big = synth.volume("tubes", (96, 80, 64), 12, 5)
    small = np.asfortranarray((rng.integers(0, 4, (40, 30, 20)) * (rng.random((40, 30, 20)) < 0.6)).astype(np.uint16))
    for labels in (big, small):
        ids, mapping, roots = _reference(labels)
        dev = _device_volume(torch, labels)
        status, count, got, gm, gr = _noalloc(torch, dev, len(roots) + 5)
        assert status == native.SUCCESS, native.lib().brook_last_error()
        assert count == len(roots) and count > 0
        assert np.array_equal(got, ids)
        assert np.array_equal(gm[:count + 1], mapping) and np.array_equal(gr[:count], roots)


def test_noalloc_rejects_small_tables_and_large_volumes(torch):
    native = _bind()
    from brook_amd import synth
    labels = synth.volume("tubes", (64, 48, 40), 6, 9)
    ids, mapping, roots = _reference(labels)
    status, count, *_ = _noalloc(torch, _device_volume(torch, labels), len(roots) - 1)
    assert status == native.INVALID_ARGUMENT and count == len(roots)
    assert b"capacity" in native.lib().brook_last_error()
    big = torch.zeros(160 * 512 * 512, dtype=torch.uint8, device="cuda").view(160, 512, 512).permute(2, 1, 0)
    status, *_ = _noalloc(torch, big, 4)
    assert status == native.INVALID_ARGUMENT and b"no-allocation scratch" in native.lib().brook_last_error()
