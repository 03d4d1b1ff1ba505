"""brook_amd: TEASAR skeletonization of 3-D label volumes on one AMD Instinct MI355X.

The interface of giorgioangel/brook's hot path: `skeletonize` and `skeletonize_batch` take
Kimimaro's arguments and return skeletons; the C ABI is include/brook.h.
"""
from __future__ import annotations

__version__ = "0.1.0+mi355x"

from ._intake import DEFAULT_TEASAR_PARAMS, DimensionError, format_labels, prepare_fields, skeletonize  # noqa: E402
from ._native import DeviceError  # noqa: E402
from ._packed import BatchedSkeletons, PackedSkeletons, Skeleton  # noqa: E402
from .batch import skeletonize_batch  # noqa: E402
from .serve import Server  # noqa: E402
from . import _device_post as device_post  # noqa: E402

__all__ = ["skeletonize", "skeletonize_batch", "BatchedSkeletons", "PackedSkeletons", "Skeleton",
           "DEFAULT_TEASAR_PARAMS", "DimensionError", "DeviceError", "format_labels", "prepare_fields", "Server", "device_post",
           "__version__"]
