"""Host tests of the serving loop's admission and bookkeeping (a stand-in runner replaces the
engine, so no GPU is needed)."""
import threading
import time

import numpy as np
import pytest

from brook_amd.serve import Server, request_bytes


def _vol(edge, dtype=np.uint32):
    return np.zeros((edge, edge, edge), dtype)


class Recorder:
    """Runner that records each batch and can be held back to let requests pile up."""

    def __init__(self):
        self.batches = []
        self.gate = threading.Event()
        self.gate.set()

    def __call__(self, volumes):
        self.gate.wait(10)
        self.batches.append([v.shape for v in volumes])
        return [v.shape for v in volumes]


def test_results_follow_their_requests():
    rec = Recorder()
    with Server(runner=rec, hbm_budget=1 << 40) as srv:
        futs = [srv.submit(_vol(e)) for e in (4, 5, 6, 7)]
        assert [f.result(5) for f in futs] == [(4, 4, 4), (5, 5, 5), (6, 6, 6), (7, 7, 7)]


def test_requests_arriving_during_a_batch_join_the_next_one():
    rec = Recorder()
    rec.gate.clear()
    with Server(runner=rec, hbm_budget=1 << 40, slots=1) as srv:
        first = srv.submit(_vol(3))
        time.sleep(0.2)  # the worker has admitted `first` and waits at the gate
        later = [srv.submit(_vol(4)) for _ in range(5)]
        rec.gate.set()
        first.result(5)
        for f in later:
            f.result(5)
    assert rec.batches[0] == [(3, 3, 3)]
    assert rec.batches[1] == [(4, 4, 4)] * 5


def test_budget_admission_is_fifo_and_never_starves_a_large_request():
    rec = Recorder()
    rec.gate.clear()
    small, big = _vol(8), _vol(32)
    budget = 2 * request_bytes(small) + 1
    with Server(runner=rec, hbm_budget=budget, slots=1) as srv:
        hold = srv.submit(_vol(2))
        time.sleep(0.2)
        futs = [srv.submit(v) for v in (small, small, small, big, small)]
        rec.gate.set()
        for f in [hold] + futs:
            f.result(5)
    # two smalls fit, the third waits; the big one does not fit beside it; then the last small
    assert rec.batches[1:] == [[(8, 8, 8)] * 2, [(8, 8, 8)], [(32, 32, 32)], [(8, 8, 8)]]


def test_max_batch_one_is_serial():
    rec = Recorder()
    rec.gate.clear()
    with Server(runner=rec, hbm_budget=1 << 40, max_batch=1, slots=1) as srv:
        futs = [srv.submit(_vol(3)) for _ in range(4)]
        rec.gate.set()
        for f in futs:
            f.result(5)
    assert all(len(b) == 1 for b in rec.batches)


def test_runner_errors_reach_every_request_of_the_batch():
    def boom(volumes):
        raise ValueError("bad batch")
    with Server(runner=boom, hbm_budget=1 << 40) as srv:
        futs = [srv.submit(_vol(3)) for _ in range(3)]
        for f in futs:
            with pytest.raises(ValueError, match="bad batch"):
                f.result(5)


def test_latency_fields_are_ordered():
    rec = Recorder()
    with Server(runner=rec, hbm_budget=1 << 40) as srv:
        f = srv.submit(_vol(3))
        f.result(5)
        r = f.request
        assert r.arrival <= r.admitted <= r.done and r.batch_size >= 1


def test_closed_server_refuses_requests():
    srv = Server(runner=Recorder(), hbm_budget=1 << 40)
    srv.close()
    with pytest.raises(RuntimeError):
        srv.submit(_vol(3))


class Gated:
    """Runner whose batches each wait for their own release, recording the order they start."""

    def __init__(self):
        self.started = []
        self.gates = {}
        self.lock = threading.Lock()

    def release(self, edge):
        with self.lock:
            self.gates.setdefault(edge, threading.Event()).set()

    def __call__(self, volumes):
        edge = volumes[0].shape[0]
        with self.lock:
            self.started.append([v.shape[0] for v in volumes])
            gate = self.gates.setdefault(edge, threading.Event())
        gate.wait(10)
        return [v.shape for v in volumes]


def _wait_for(cond, timeout=5.0):
    end = time.time() + timeout
    while not cond() and time.time() < end:
        time.sleep(0.01)
    assert cond()


def test_arrival_during_a_batch_starts_on_the_free_slot():
    run = Gated()
    with Server(runner=run, hbm_budget=1 << 40, slots=2) as srv:
        first = srv.submit(_vol(3))
        _wait_for(lambda: len(run.started) == 1)  # slot 0 holds the first batch
        later = srv.submit(_vol(5))
        _wait_for(lambda: len(run.started) == 2)  # admitted at once, beside the first
        run.release(5)
        assert later.result(5) == (5, 5, 5) and not first.done()
        run.release(3)
        assert first.result(5) == (3, 3, 3)
    assert not first.request.overlapped and later.request.overlapped
    assert first.request.slot != later.request.slot


def test_budget_counts_the_requests_in_flight_on_other_slots():
    run = Gated()
    small = _vol(8)
    budget = 2 * request_bytes(small) + 1
    with Server(runner=run, hbm_budget=budget, slots=3) as srv:
        a = srv.submit(_vol(8))
        _wait_for(lambda: len(run.started) == 1)
        b = srv.submit(_vol(8))
        _wait_for(lambda: len(run.started) == 2)
        c = srv.submit(_vol(8))  # a third slot is free, but the budget is not: c waits
        time.sleep(0.2)
        assert len(run.started) == 2 and not c.done()
        run.release(8)
        for f in (a, b, c):
            assert f.result(5) == (8, 8, 8)
    assert len(run.started) == 3


def test_a_request_above_the_budget_runs_alone_once_nothing_is_in_flight():
    run = Gated()
    small, big = _vol(8), _vol(32)
    budget = request_bytes(small) + request_bytes(small) // 2
    with Server(runner=run, hbm_budget=budget, slots=2) as srv:
        a = srv.submit(small)
        _wait_for(lambda: len(run.started) == 1)
        b = srv.submit(big)
        time.sleep(0.2)
        assert len(run.started) == 1  # the big one waits for the small one to finish
        run.release(8)
        run.release(32)
        assert a.result(5) == (8, 8, 8) and b.result(5) == (32, 32, 32)
    assert run.started == [[8], [32]]
