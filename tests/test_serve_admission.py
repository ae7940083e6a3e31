"""Admission logic of the native movie server (zr_server_*, host code, no
GPU): in-flight limit, HBM budget with LRU eviction of idle plans, reuse of
plans of the same key, refusal of a movie larger than the budget,
cancellation, and plan creation outside the scheduler's lock.  Plans are
Python stand-ins, driven by the native scheduler through zr_plan_ops, that
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
def fake():
    FakePlan.live = FakePlan.created = FakePlan.running = FakePlan.peak_running = 0
    return FakePlan


KW = dict(triMode="tri", maxShift=10, Bmode="conv", Brad=150, fouCrop=(0, 0), filterMode=None, doLazyFRC=False)


def movie(n, h, w, tag):
    return np.full((n, h, w), float(tag), np.float32)


def test_results_in_flight_limit_and_plan_reuse(fake):
    srv = S.MovieServer(hbm_budget=1 << 40, max_inflight=2, return_sums="none", plan_class=FakePlan)
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
    srv = S.MovieServer(hbm_budget=int(one * 1.5), max_inflight=2, return_sums="none", plan_class=FakePlan)
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
    srv = S.MovieServer(hbm_budget=1 << 20, max_inflight=2, return_sums="none", plan_class=FakePlan)
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


class SlowAlignPlan(FakePlan):
    def align(self):
        time.sleep(0.2)


def test_cancelled_request_is_dropped_and_released_once():
    FakePlan.live = FakePlan.created = FakePlan.running = FakePlan.peak_running = 0
    srv = S.MovieServer(hbm_budget=1 << 40, max_inflight=1, return_sums="none", plan_class=SlowAlignPlan)
    a = srv.submit(movie(4, 64, 64, 1), shapePadded=(64, 64), **KW)
    b = srv.submit(movie(4, 64, 64, 2), shapePadded=(64, 64), **KW)
    assert b.cancel()                           # still queued behind a: cancellable
    c = srv.submit(movie(4, 64, 64, 3), shapePadded=(64, 64), **KW)
    assert np.all(a.result(timeout=30).translations == 1)
    assert np.all(c.result(timeout=30).translations == 3)
    assert b.cancelled()
    assert not a.cancel()                        # a finished: too late to cancel
    st = srv.stats
    assert st["inflight"] == 0 and st["queued"] == 0 and st["admitted"] == 2
    srv.close()
    assert FakePlan.live == 0


class SlowCreatePlan(FakePlan):
    def __init__(self, params, n, h, w, device=0):
        if h == 128:
            time.sleep(0.6)                      # a multi-GB plan being allocated
        super().__init__(params, n, h, w, device)


def test_slow_plan_creation_does_not_hold_up_other_movies():
    FakePlan.live = FakePlan.created = FakePlan.running = FakePlan.peak_running = 0
    srv = S.MovieServer(hbm_budget=1 << 40, max_inflight=2, return_sums="none", plan_class=SlowCreatePlan)
    assert srv.warm((4, 64, 64), count=1, shapePadded=(64, 64), **KW) == 1
    t0 = time.perf_counter()
    big = srv.submit(movie(4, 128, 128, 5), shapePadded=(128, 128), **KW)    # admitted first, plan created slowly
    time.sleep(0.05)
    small = srv.submit(movie(4, 64, 64, 6), shapePadded=(64, 64), **KW)     # the warmed plan
    r = small.result(timeout=30)
    t_small = time.perf_counter() - t0
    assert np.all(r.translations == 6)
    assert t_small < 0.45, t_small               # done while the other plan is still being created
    assert not big.done()
    assert np.all(big.result(timeout=30).translations == 5)
    srv.close()
    assert FakePlan.live == 0


def test_smaller_movie_goes_ahead_of_a_head_that_does_not_fit():
    FakePlan.live = FakePlan.created = FakePlan.running = FakePlan.peak_running = 0
    one = 4 * 64 * 64 * (FakePlan.bytes_per_pixel + 4)
    srv = S.MovieServer(hbm_budget=int(one * 2.5), max_inflight=2, return_sums="none", plan_class=SlowAlignPlan)
    a = srv.submit(movie(4, 64, 64, 1), shapePadded=(64, 64), **KW)
    time.sleep(0.05)                               # a is running on the first plan
    # b: a new key, charged the server's upper estimate until its first plan exists,
    # which cannot be admitted while anything else is held
    b = srv.submit(movie(4, 64, 64, 2), shapePadded=(128, 128), **KW)
    c = srv.submit(movie(4, 64, 64, 3), shapePadded=(64, 64), **KW)    # a's key: a second plan fits
    rb, rc = b.result(timeout=30), c.result(timeout=30)
    assert rc.admitted < rb.admitted
    assert np.all(rb.translations == 2) and np.all(rc.translations == 3)
    a.result(timeout=30)
    srv.close()
    assert FakePlan.live == 0


def test_c_abi_poll_wait_result_release():
    import ctypes as C
    from zorro_amd import _lib as L
    lib = L.load()
    ops = S.PythonPlanOps(SlowAlignPlan)
    h = C.c_void_p()
    L.check(lib.zr_server_create_ops(C.byref(ops.ops), C.c_size_t(1 << 40), 1, C.byref(h)))
    try:
        frames = movie(3, 32, 32, 7)
        params = S.make_params(shapePadded=(32, 32), **KW)
        rid = C.c_longlong()
        L.check(lib.zr_server_submit(h, C.byref(params), 3, 32, 32, C.c_void_p(frames.ctypes.data), None, None,
                                     C.c_double(0.0), C.byref(rid)))
        st = C.c_int()
        L.check(lib.zr_server_poll(h, rid, C.byref(st)))
        assert st.value in (L.ZR_REQ_QUEUED, L.ZR_REQ_RUNNING)
        t = np.zeros((3, 2))
        assert lib.zr_server_result(h, rid, L.dptr(t), 3, None, 0, None, None, None, 0) == L.ZR_ERR_STATE
        assert lib.zr_server_poll(h, C.c_longlong(rid.value + 99), C.byref(st)) == L.ZR_ERR_ARG
        L.check(lib.zr_server_wait(h, rid, C.c_double(-1.0), C.byref(st)))
        assert st.value == L.ZR_REQ_DONE
        times = np.zeros(3)
        L.check(lib.zr_server_result(h, rid, L.dptr(t), 3, None, 0, None, L.dptr(times), None, 0))
        assert np.all(t == 7) and times[2] >= times[1] >= times[0] > 0
        assert lib.zr_server_cancel(h, rid) == L.ZR_ERR_STATE          # finished: too late
        L.check(lib.zr_server_release(h, rid))
        assert lib.zr_server_poll(h, rid, C.byref(st)) == L.ZR_ERR_ARG
        v = (C.c_longlong * 8)()
        L.check(lib.zr_server_stats(h, v))
        assert v[0] == 1 and v[1] == 1 and v[5] == 0 and v[6] == 0
    finally:
        lib.zr_server_destroy(h)
