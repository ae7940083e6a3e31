"""Admission logic of the movie server (host code, no GPU): in-flight limit,
HBM budget with LRU eviction of idle plans, reuse of plans of the same key,
and refusal of a movie larger than the budget.  Plans are stand-ins that
record their life cycle; the GPU behaviour is in tests/test_gpu_serve.py."""
import threading
import time

import numpy as np
import pytest

from zorro_amd import serve as S


class FakeLib:
    def __init__(self, nbytes):
        self.nbytes = nbytes

    def zr_movie_device_bytes(self, mv, out):
        out._obj.value = self.nbytes
        return 0


class FakePlan:
    live = 0
    created = 0
    running = 0
    peak_running = 0
    lock = threading.Lock()
    bytes_per_pixel = 40

    def __init__(self, params, n, h, w, device=0):
        self.n, self.h, self.w = n, h, w
        self.params = params
        self.lib = FakeLib(n * h * w * self.bytes_per_pixel)
        self._mv = None
        with FakePlan.lock:
            FakePlan.live += 1
            FakePlan.created += 1
        self.closed = False

    def set_frames(self, frames):
        self.frames = frames

    def align(self):
        with FakePlan.lock:
            FakePlan.running += 1
            FakePlan.peak_running = max(FakePlan.peak_running, FakePlan.running)
        time.sleep(0.01)
        with FakePlan.lock:
            FakePlan.running -= 1

    def synchronize(self):
        pass

    def solution(self):
        return dict(translations=np.full((self.n, 2), float(self.frames[0, 0, 0])))

    def frc_curve(self):
        return np.zeros(4)

    def close(self):
        assert not self.closed
        self.closed = True
        with FakePlan.lock:
            FakePlan.live -= 1


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(S, "MoviePlan", FakePlan)
    FakePlan.live = FakePlan.created = FakePlan.running = FakePlan.peak_running = 0
    return FakePlan


KW = dict(triMode="tri", maxShift=10, Bmode="conv", Brad=150, fouCrop=(0, 0), filterMode=None, doLazyFRC=False)


def movie(n, h, w, tag):
    return np.full((n, h, w), float(tag), np.float32)


def test_results_in_flight_limit_and_plan_reuse(fake):
    srv = S.MovieServer(hbm_budget=1 << 40, max_inflight=2, return_sums="none")
    futs = [srv.submit(movie(4, 64, 64, i), shapePadded=(64, 64), **KW) for i in range(8)]
    res = [f.result(timeout=30) for f in futs]
    srv.close()
    for i, r in enumerate(res):
        assert np.all(r.translations == i) and r.done >= r.admitted >= r.arrival
    assert fake.peak_running <= 2
    assert fake.created <= 2            # one key: at most one plan per in-flight slot
    assert fake.live == 0               # close() frees every plan


def test_budget_evicts_idle_plans_of_other_shapes(fake):
    one = 4 * 64 * 64 * (FakePlan.bytes_per_pixel + 4)          # plan bytes + frames, as the server charges
    srv = S.MovieServer(hbm_budget=int(one * 1.5), max_inflight=2, return_sums="none")
    shapes = [(64, 64), (64, 64), (64, 64)]
    a = [srv.submit(movie(4, h, w, i), shapePadded=(h, w), **KW) for i, (h, w) in enumerate(shapes)]
    b = srv.submit(movie(4, 64, 64, 9), shapePadded=(128, 128), **KW)    # another key of the same size
    for f in a + [b]:
        f.result(timeout=30)
    srv.close()
    assert fake.peak_running == 1       # only one plan fits the budget
    assert srv.stats["plans_evicted"] >= 1
    assert srv.stats["peak_held"] <= srv.budget


def test_movie_larger_than_budget_is_refused(fake):
    srv = S.MovieServer(hbm_budget=1 << 20, max_inflight=2, return_sums="none")
    f = srv.submit(movie(4, 512, 512, 1), shapePadded=(512, 512), **KW)
    with pytest.raises(MemoryError):
        f.result(timeout=30)
    g = srv.submit(movie(2, 16, 16, 2), shapePadded=(16, 16), **KW)
    assert np.all(g.result(timeout=30).translations == 2)
    srv.close()


def test_estimate_covers_the_workspaces():
    params = S.make_params(triMode="tri", maxShift=40, fouCrop=(3072, 3072), shapePadded=(4096, 4096),
                           filterMode="dose", pixelsize=0.1)
    est = S.estimate_plan_bytes(40, 4096, 4096, params)
    spectra = 40 * 2049 * 4096 * 8                    # full-height half spectra kept for the dose filter
    assert est > spectra + 780 * 1537 * 80 * 8
    assert est < 16e9
