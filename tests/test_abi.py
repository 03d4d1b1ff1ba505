"""CPU checks of the C ABI: the library loads without a GPU and exports every symbol that
include/brook.h declares; options/defaults handling of the Python layer."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "brook.h")).read()
    return sorted(set(re.findall(r"\b(brook_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from brook_amd import _native
    lib = ctypes.CDLL(_native.LIBRARY)
    missing = [name for name in _declared() if not hasattr(lib, name)]
    assert not missing, missing
    assert len(_declared()) >= 20


def test_abi_version_and_options_defaults():
    from brook_amd import _native
    L = _native.lib()
    assert L.brook_abi_version() == 1
    o = _native.SkeletonizeOptions()
    L.brook_skeletonize_options_init(ctypes.byref(o))
    # the Python skeletonize() defaults (c_api.cpp:218-226 of the reference)
    assert (o.teasar.scale, o.teasar.constant, o.teasar.pdrf_scale, o.teasar.pdrf_exponent) == (1.5, 300, 100000, 4)
    assert o.dust_threshold == 1000 and o.flags == _native.FIX_BRANCHING | _native.FIX_BORDERS
    assert ctypes.sizeof(_native.Volume) == 80


def test_context_without_device_is_an_error():
    from brook_amd import _native
    h = ctypes.c_void_p()
    status = _native.lib().brook_context_create(-1, ctypes.byref(h))
    assert status == _native.INVALID_ARGUMENT
    assert b"negative" in _native.lib().brook_last_error()


def test_teasar_defaults_and_unknown_keys():
    from brook_amd._intake import teasar_struct
    t = teasar_struct({"scale": 2})
    assert t.scale == 2 and t.constant == 10 and t.pdrf_exponent == 16  # kimimaro.trace defaults
    with pytest.raises(TypeError):
        teasar_struct({"bogus": 1})


def test_format_labels():
    from brook_amd import DimensionError, format_labels
    assert format_labels(np.zeros((4, 5), bool)).shape == (4, 5, 1)
    assert format_labels(np.zeros((4, 5, 6, 1), np.uint8)).shape == (4, 5, 6)
    with pytest.raises(DimensionError):
        format_labels(np.zeros((2, 3, 4, 5)))


def test_synthetic_volumes_are_deterministic():
    from brook_amd import synth
    a = synth.volume("tubes", (48, 40, 32), 5, 7)
    b = synth.volume("tubes", (48, 40, 32), 5, 7)
    assert np.array_equal(a, b) and a.max() == 5 and a.flags.f_contiguous


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors have the sizes and field offsets a C compiler gives the header's structs."""
    import subprocess
    from brook_amd import _native
    structs = {"brook_volume": _native.Volume, "brook_skeletonize_options": _native.SkeletonizeOptions,
               "brook_skeletons_view": _native.SkeletonsView,
               "brook_skeletons_device_view": _native.SkeletonsDeviceView}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "brook.h"', 'int main(void) {']
    for name, mirror in structs.items():
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for field, _ in mirror._fields_:
            lines.append(f'  printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    include = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c11", "-I", include, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(line.split()[:2]): int(line.split()[2]) for line in out if line}
    for name, mirror in structs.items():
        assert got[(name, "size")] == ctypes.sizeof(mirror), name
        for field, _ in mirror._fields_:
            assert got[(name, field)] == getattr(mirror, field).offset, (name, field)
