"""Serving: a request loop with continuous batching on one GPU.

`Server.submit(labels)` returns a future at once. One worker thread owns the GPU: whenever it is
free it admits every waiting request that fits the device-memory budget (first come, first
served; a request larger than the budget runs alone) and runs them as one engine batch, so the
requests' labels share the flood and trace launches and their work overlaps on the device.
Requests that arrive while a batch runs are admitted with the next one. Each result equals the
result of a plain `skeletonize` call on the same volume (the engine's per-sample stages are the
plain call's stages), which the parity tests check.

The reference has no serving loop (SURVEY of brook: `GpuPool.imap` in-flight bounding,
`_pool.py:198-266`, is the nearest); this is the request loop of the serving form this build is
measured on.
"""
from __future__ import annotations

import collections
import threading
import time
from concurrent.futures import Future

from ._intake import DEFAULT_TEASAR_PARAMS
from ._packed import PackedSkeletons

# Device bytes a request holds while the engine runs it: 40 per voxel, 56 per foreground voxel at
# up to 40% occupancy, plus slack (the engine's own group sizing uses the same figures).
_BYTES_PER_VOXEL = 40 + 0.4 * 56
_SLACK = 64 << 20


def request_bytes(labels) -> int:
    """Device-memory budget a request is admitted against."""
    n = 1
    for s in labels.shape:
        n *= int(s)
    return int(n * (_BYTES_PER_VOXEL + labels.dtype.itemsize)) + _SLACK


class Request:
    __slots__ = ("labels", "future", "arrival", "admitted", "done", "batch_size")

    def __init__(self, labels):
        self.labels = labels
        self.future = Future()
        self.arrival = time.perf_counter()
        self.admitted = self.done = None
        self.batch_size = 0


class Server:
    """Continuous-batching skeletonization server (one worker thread, one GPU).

    hbm_budget: device bytes the requests of one batch may take (default: half the free device
    memory when the server starts); max_batch: requests per batch (1 = serial plain calls).
    runner(list_of_labels) -> list of results replaces the engine (used by the host tests).
    """

    def __init__(self, teasar_params=DEFAULT_TEASAR_PARAMS, anisotropy=(1, 1, 1), dust_threshold=1000,
                 fix_branching=True, fix_borders=True, hbm_budget=None, max_batch=64, runner=None):
        if max_batch < 1:
            raise ValueError("brook: max_batch must be at least 1")
        self._params = dict(teasar_params)
        self._aniso = tuple(float(a) for a in anisotropy)
        self._dust = dust_threshold
        self._fix_branching, self._fix_borders = fix_branching, fix_borders
        self._max_batch = int(max_batch)
        self._runner = runner or self._engine
        if hbm_budget is None:
            hbm_budget = 8 << 30
            if runner is None:
                import torch
                free, _ = torch.cuda.mem_get_info()
                hbm_budget = free // 2
        self._budget = int(hbm_budget)
        self._pending = collections.deque()
        self._cv = threading.Condition()
        self._stop = False
        self.batches = []  # sizes of the batches run, in order
        self._thread = threading.Thread(target=self._loop, name="brook-serve", daemon=True)
        self._thread.start()

    def submit(self, labels) -> Future:
        """Queues a label volume (3-D numpy array or device tensor); the future's result is a
        PackedSkeletons (host arrays)."""
        req = Request(labels)
        with self._cv:
            if self._stop:
                raise RuntimeError("brook: server is closed")
            self._pending.append(req)
            self._cv.notify()
        req.future.request = req  # timing fields for callers that measure latency
        return req.future

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify()
        self._thread.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _admit(self):
        """FIFO requests that fit the budget (at least one)."""
        batch, used = [], 0
        while self._pending and len(batch) < self._max_batch:
            need = request_bytes(self._pending[0].labels)
            if batch and used + need > self._budget:
                break
            batch.append(self._pending.popleft())
            used += need
        return batch

    def _engine(self, volumes):
        from .batch import batch_parts
        parts = batch_parts(volumes, self._params, self._aniso, self._dust, self._fix_branching, self._fix_borders)
        return [p if len(p) else PackedSkeletons.from_skeletons({}, self._aniso) for p in parts]

    def _loop(self):
        while True:
            with self._cv:
                while not self._pending and not self._stop:
                    self._cv.wait()
                if not self._pending and self._stop:
                    return
                batch = self._admit()
            now = time.perf_counter()
            for r in batch:
                r.admitted, r.batch_size = now, len(batch)
            self.batches.append(len(batch))
            try:
                results = self._runner([r.labels for r in batch])
            except BaseException as e:  # every admitted request learns why
                for r in batch:
                    r.done = time.perf_counter()
                    r.future.set_exception(e)
                continue
            done = time.perf_counter()
            for r, res in zip(batch, results):
                r.done = done
                r.future.set_result(res)
