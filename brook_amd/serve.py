"""Serving: a request loop with continuous batching on one GPU.

`Server.submit(labels)` returns a future at once. The server runs `slots` engine workers, each a
thread with its own device context (its own HIP stream and device-buffer pool). Whenever a slot is
free it admits every waiting request that fits what is left of the device-memory budget beside
the requests already on the GPU (first come, first served; a request larger than the whole budget
runs alone once nothing else is in flight) and runs them as one engine batch, so those requests'
labels share the flood and trace launches. A request that arrives while another batch is on the
GPU starts at once on the next free slot: its CCL, EDT and floods run on its slot's stream beside
the other slot's trace kernel, whose persistent workgroups take only one compute unit per label.
Each result equals the result of a plain `skeletonize` call on the same volume (the engine's
per-sample stages are the plain call's stages), which the parity tests check alone and in the mix.

Device buffers stay in each slot's pool between requests (size classes of a quarter octave, so
volumes of other shapes reuse them), so a slot's label, EDT and frontier arrays are allocated once
and refilled per request.

The reference has no serving loop (`GpuPool.imap`'s in-flight bounding, `_pool.py:198-266`, is the
nearest); this is the request loop of the serving form this build is measured on.
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
    __slots__ = ("labels", "future", "arrival", "admitted", "done", "batch_size", "slot", "overlapped")

    def __init__(self, labels):
        self.labels = labels
        self.future = Future()
        self.arrival = time.perf_counter()
        self.admitted = self.done = None
        self.batch_size = 0
        self.slot = -1
        self.overlapped = False  # admitted while another slot's batch was on the GPU


class Server:
    """Continuous-batching skeletonization server on one GPU.

    slots: engine workers (device contexts) that run batches side by side; hbm_budget: device bytes
    the requests in flight may take together (default: half the free device memory when the server
    starts); max_batch: requests per batch. slots=1, max_batch=1 is serial plain calls.
    runner(list_of_labels) -> list of results replaces the engine (used by the host tests); it is
    called from the slot's worker thread.
    """

    def __init__(self, teasar_params=DEFAULT_TEASAR_PARAMS, anisotropy=(1, 1, 1), dust_threshold=1000,
                 fix_branching=True, fix_borders=True, hbm_budget=None, max_batch=64, runner=None, slots=2):
        if max_batch < 1:
            raise ValueError("brook: max_batch must be at least 1")
        if slots < 1:
            raise ValueError("brook: slots must be at least 1")
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
        self._in_flight = 0  # bytes admitted and not yet done, over all slots
        self._busy = 0  # slots running a batch
        self.batches = []  # sizes of the batches run, in order of admission
        self._threads = [threading.Thread(target=self._loop, args=(s,), name=f"brook-serve-{s}", daemon=True)
                         for s in range(int(slots))]
        for t in self._threads:
            t.start()

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
            self._cv.notify_all()
        for t in self._threads:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _admit(self):
        """FIFO requests that fit the budget beside what is in flight; an empty list when the head
        request must wait for memory (it runs alone, at the latest once nothing is in flight)."""
        batch, used = [], 0
        while self._pending and len(batch) < self._max_batch:
            need = request_bytes(self._pending[0].labels)
            if self._in_flight + used + need > self._budget and (batch or self._in_flight):
                break
            batch.append(self._pending.popleft())
            used += need
        return batch, used

    def _engine(self, volumes):
        from .batch import batch_parts
        parts = batch_parts(volumes, self._params, self._aniso, self._dust, self._fix_branching, self._fix_borders)
        return [p if len(p) else PackedSkeletons.from_skeletons({}, self._aniso) for p in parts]

    def _loop(self, slot):
        while True:
            with self._cv:
                while True:
                    if self._pending:
                        batch, used = self._admit()
                        if batch:
                            break
                    elif self._stop:
                        return
                    self._cv.wait()
                overlapped = self._busy > 0
                self._busy += 1
                self._in_flight += used
                now = time.perf_counter()
                for r in batch:
                    r.admitted, r.batch_size, r.slot, r.overlapped = now, len(batch), slot, overlapped
                self.batches.append(len(batch))
                if self._pending:
                    self._cv.notify()  # another free slot may take what this batch left
            try:
                results, error = self._runner([r.labels for r in batch]), None
            except BaseException as e:  # every admitted request learns why
                results, error = None, e
            done = time.perf_counter()
            with self._cv:
                self._busy -= 1
                self._in_flight -= used
                self._cv.notify_all()  # memory freed: waiting requests may fit now
            for i, r in enumerate(batch):
                r.done = done
                if error is not None:
                    r.future.set_exception(error)
                else:
                    r.future.set_result(results[i])
