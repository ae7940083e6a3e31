"""Continuous-batching movie server on one MI355X.

The reference registers one movie per process and scales by running several
processes (the Skulk manager, ``automator/zorroSkulkManager.py:917-1203``).
Here one process keeps a pool of movie plans resident in HBM and runs several
movies at once, each on its own HIP stream, so that the latency-bound stages
of one movie (the one-block shift solver, the logistic fit) overlap with the
bandwidth-bound pair transforms of another.

* Requests (frames + ``ImageRegistrator`` settings) queue in arrival order.
* A request is admitted when an in-flight slot is free and its plan fits the
  HBM budget: an idle plan of the same shape and settings is reused, otherwise
  a new one is created, evicting idle plans of other shapes (least recently
  used first) when the budget requires it.  A request that does not fit behind
  a larger head of the queue may go first ("admitted as slots and memory free
  up"); nothing is starved, because admission always retries the oldest
  request first.
* Each admitted movie runs ``zr_align`` on its plan's stream in a worker
  thread (the C calls release the GIL), copies its sums into caller-owned
  device tensors and its translations to the host, and frees the plan for the
  next request.

Results do not depend on what else is in flight: every kernel of a movie
reads and writes only that movie's buffers, and no reduction depends on
scheduling (tests/test_gpu_serve.py checks alone == in the mix).
"""
from __future__ import annotations

import collections
import ctypes as C
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor

import numpy as np

from . import _lib as L
from .engine import MoviePlan, make_params


def plan_key(n, h, w, params):
    return (int(n), int(h), int(w), bytes(memoryview(params)))


def estimate_plan_bytes(n, h, w, params) -> int:
    """Upper estimate of a plan's device memory before it exists (the exact
    figure, ``zr_movie_device_bytes``, replaces it once a plan of the key has
    been created)."""
    hp = params.shape_padded[0] or h
    wp = params.shape_padded[1] or w
    crop = params.fou_crop[0] != 0 or params.fou_crop[1] != 0
    hc = 2 * (params.fou_crop[0] // 2) if crop else hp
    wc = 2 * (params.fou_crop[1] // 2) if crop else wp
    m = params.max_shift
    masked = (hp != h or wp != w)
    npairs = n * (n - 1) // 2
    nspec = 2 * n + 1 if masked else n
    ptot = npairs + (4 * n + 1 if masked else 0)
    wk = wp // 2 + 1
    spectral = params.do_dose and not masked and h == w
    hs = hp if spectral else hc
    b = min(n, 8) * wk * hp * 8                       # row-pass scratch
    b += nspec * wk * hs * 8 + nspec * wk * 8          # spectra
    b += ptot * (wc // 2 + 1) * 2 * m * 8              # pair pass 1
    b += ptot * 4 * m * m * 4                          # lag windows
    b += 4 * h * w * 4 + n * h * w * 4                 # sums + an owned copy of the frames
    if any(params.filters[i] == L.ZR_FILTER_GAUSSLP for i in range(params.n_filters)):
        b += h * w * 4                                 # the Gaussian's first pass
    if params.do_dose:
        b += wk * hp * 8
        if not spectral:
            b += min(n, 8) * (wk + 1) * hp * 8         # a chunk of the direct dose path
        else:
            b += 3 * n * (wk + hp) * 8
    if params.do_frc:
        box = params.frc_box
        tiles = max(1, (h // (box // 2)) * (w // (box // 2)))
        b += tiles * 2 * (box // 2 + 1) * box * 8
    return int(b * 1.05) + (16 << 20)


class MovieResult:
    """Outputs of one served movie (the ImageRegistrator result attributes)."""

    def __init__(self, translations, imageSum, filtSum, FRC, arrival, admitted, done, key):
        self.translations = translations
        self.imageSum = imageSum
        self.filtSum = filtSum
        self.FRC = FRC
        self.arrival, self.admitted, self.done = arrival, admitted, done
        self.key = key

    @property
    def latency(self):
        return self.done - self.arrival


class _Request:
    __slots__ = ("frames", "n", "h", "w", "params", "key", "future", "arrival", "admitted")

    def __init__(self, frames, params, arrival, shape=None):
        self.frames = frames
        self.n, self.h, self.w = (int(v) for v in (shape or frames.shape))
        self.params = params
        self.key = plan_key(self.n, self.h, self.w, params)
        self.future = Future()
        self.arrival = arrival
        self.admitted = None


class MovieServer:
    """Serve registration requests of mixed shapes on one device.

    ``hbm_budget``: bytes of device memory the server may hold in plans
    (default: 85% of the free memory at start).  ``max_inflight``: movies
    running at once (each on its own stream)."""

    def __init__(self, device: int = 0, hbm_budget: int | None = None, max_inflight: int = 3,
                 return_sums: str = "device"):
        import torch
        self.torch = torch
        self.device = int(device)
        self.dev = torch.device("cuda", self.device)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.dev)
            hbm_budget = int(free * 0.85)
        self.budget = int(hbm_budget)
        self.max_inflight = int(max_inflight)
        self.return_sums = return_sums          # "device" (torch tensors), "host" (numpy) or "none"
        self._lock = threading.Condition()
        self._pending = collections.deque()
        self._idle = collections.OrderedDict()  # key -> [plans], LRU order of keys
        self._plan_bytes = {}                   # id(plan) -> bytes
        self._key_bytes = {}                    # key -> measured bytes of one plan
        self._held = 0                          # bytes in all plans
        self._inflight = 0
        self._creating = 0                      # plans being created outside the lock
        self._closed = False
        self._pool = ThreadPoolExecutor(max_workers=self.max_inflight, thread_name_prefix="zr-movie")
        self._sched = threading.Thread(target=self._schedule, name="zr-admit", daemon=True)
        self.stats = dict(admitted=0, plans_created=0, plans_evicted=0, peak_held=0)
        self._sched.start()

    # ------------------------------------------------------------ public
    def submit(self, frames, arrival: float | None = None, **settings) -> Future:
        """Queue one movie: frames f32 [N,H,W] (numpy or a CUDA tensor kept
        alive by the caller until the future completes) and ImageRegistrator
        settings as keywords (``engine.make_params``)."""
        params = make_params(**settings)
        req = _Request(frames, params, time.perf_counter() if arrival is None else arrival)
        with self._lock:
            if self._closed:
                raise RuntimeError("server is closed")
            self._pending.append(req)
            self._lock.notify_all()
        return req.future

    def close(self):
        with self._lock:
            self._closed = True
            self._lock.notify_all()
        self._sched.join()
        self._pool.shutdown(wait=True)
        with self._lock:
            for plans in self._idle.values():
                for p in plans:
                    p.close()
            self._idle.clear()
            self._held = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ admission
    def _bytes_for(self, req):
        return self._key_bytes.get(req.key) or estimate_plan_bytes(req.n, req.h, req.w, req.params)

    def _evict_for(self, need):
        """Free idle plans (least recently used keys first) until `need` bytes fit."""
        while self._held + need > self.budget and self._idle:
            key, plans = next(iter(self._idle.items()))
            p = plans.pop()
            if not plans:
                del self._idle[key]
            self._held -= self._plan_bytes.pop(id(p), 0)
            p.close()
            self.stats["plans_evicted"] += 1
        return self._held + need <= self.budget

    def _take_plan(self, req):
        """Called under the lock.  ('idle', plan) for an idle plan of the
        request's key, ('create', bytes) when a new plan may be created (its
        bytes are then reserved), or None if neither is possible now."""
        plans = self._idle.get(req.key)
        if plans:
            p = plans.pop()
            if not plans:
                del self._idle[req.key]
            return "idle", p
        known = req.key in self._key_bytes
        need = self._bytes_for(req)
        if not self._evict_for(need):
            # the first plan of a key is charged an upper estimate; with nothing
            # else held, create it and check the exact figure instead
            if known or self._held > 0:
                return None
        self._held += need
        return "create", need

    def _create_plan(self, req, reserved):
        """Create a plan outside the lock (multi-GB allocations and table
        uploads must not hold up completions), then settle its exact size."""
        p = MoviePlan(req.params, req.n, req.h, req.w, device=self.device)
        nb = C.c_size_t()
        L.check(p.lib.zr_movie_device_bytes(p._mv, C.byref(nb)))
        actual = int(nb.value) + req.n * req.h * req.w * 4
        with self._lock:
            self._key_bytes[req.key] = actual
            self._held += actual - reserved
            if self._held > self.budget:
                self._evict_for(0)
            if self._held > self.budget and self._held == actual:
                self._held -= actual
                p.close()
                raise MemoryError(f"movie {req.n}x{req.h}x{req.w} needs {actual / 1e9:.2f} GB, more than the "
                                  f"HBM budget {self.budget / 1e9:.2f} GB")
            self._plan_bytes[id(p)] = actual
            self.stats["plans_created"] += 1
            self.stats["peak_held"] = max(self.stats["peak_held"], self._held)
        return p

    def warm(self, frames_like, count: int = 1, **settings):
        """Create ``count`` idle plans for movies shaped like ``frames_like``
        ([N, H, W], or the shape itself) with these settings, ahead of the
        requests (serving warm-up: no request then waits for a plan)."""
        shape = tuple(int(v) for v in getattr(frames_like, "shape", frames_like))
        params = make_params(**settings)
        req = _Request(np.empty((0,), np.float32), params, 0.0, shape=shape)
        made = []
        for _ in range(int(count)):
            with self._lock:
                if not self._evict_for(self._bytes_for(req)) and self._held > 0:
                    break
                need = self._bytes_for(req)
                self._held += need
            made.append(self._create_plan(req, need))
        with self._lock:
            for p in made:
                self._idle.setdefault(req.key, []).append(p)
            if made:
                self._idle.move_to_end(req.key)
            self._lock.notify_all()
        return len(made)

    def _schedule(self):
        while True:
            with self._lock:
                while not self._closed and (not self._pending or self._inflight >= self.max_inflight):
                    self._lock.wait()
                if self._closed and not self._pending:
                    return
                if self._inflight >= self.max_inflight:
                    continue
                # requests cancelled while queued are dropped (Future.cancel succeeds
                # until the request is admitted)
                for req in [r for r in self._pending if r.future.cancelled()]:
                    self._pending.remove(req)
                chosen, got = None, None
                for req in list(self._pending):
                    got = self._take_plan(req)
                    if got is not None:
                        chosen = req
                        break
                if chosen is None:
                    if self._inflight == 0 and self._pending and self._creating == 0:
                        # nothing running and nothing fits even after eviction
                        req = self._pending.popleft()
                        if req.future.set_running_or_notify_cancel():
                            req.future.set_exception(MemoryError(
                                f"movie {req.n}x{req.h}x{req.w} needs ~{self._bytes_for(req) / 1e9:.1f} GB, "
                                f"more than the HBM budget {self.budget / 1e9:.1f} GB"))
                        continue
                    self._lock.wait()
                    continue
                self._pending.remove(chosen)
                if not chosen.future.set_running_or_notify_cancel():
                    # cancelled between the sweep and now: give the plan or the reservation back
                    if got[0] == "idle":
                        self._idle.setdefault(chosen.key, []).append(got[1])
                    else:
                        self._held -= got[1]
                    continue
                chosen.admitted = time.perf_counter()
                self._inflight += 1
                if got[0] == "create":
                    self._creating += 1
                self.stats["admitted"] += 1
            self._pool.submit(self._run, chosen, got)

    # ------------------------------------------------------------ execution
    def _run(self, req, got):
        """One admitted movie.  The plan is released exactly once, and the
        future (already marked running, so it can no longer be cancelled) gets
        its result or exception after the release."""
        torch = self.torch
        res = err = None
        plan = None
        try:
            if got[0] == "create":
                try:
                    plan = self._create_plan(req, got[1])
                except BaseException:
                    with self._lock:
                        self._held -= got[1] if req.key not in self._key_bytes else 0
                    raise
                finally:
                    with self._lock:
                        self._creating -= 1
            else:
                plan = got[1]
            plan.set_frames(req.frames)
            plan.align()
            h, w = req.h, req.w
            img = filt = None
            if self.return_sums != "none":
                img = torch.empty((h, w), dtype=torch.float32, device=self.dev)
                filt = torch.empty((h, w), dtype=torch.float32, device=self.dev) if req.params.n_filters else None
                L.check(plan.lib.zr_copy_sums_device(plan._mv, C.c_void_p(img.data_ptr()),
                                                     C.c_void_p(filt.data_ptr() if filt is not None else 0),
                                                     None, None))
            plan.synchronize()
            sol = plan.solution()
            frc = plan.frc_curve() if req.params.do_frc else None
            done = time.perf_counter()
            if self.return_sums == "host":
                img = img.cpu().numpy()
                filt = filt.cpu().numpy() if filt is not None else None
            res = MovieResult(sol["translations"], img, filt, frc, req.arrival, req.admitted, done, req.key)
            plan._frames_ref = None
        except BaseException as e:
            err = e
        finally:
            self._release(req, plan, broken=err is not None)
        if err is None:
            req.future.set_result(res)
        else:
            req.future.set_exception(err)

    def _release(self, req, plan, broken=False):
        with self._lock:
            self._inflight -= 1
            if plan is not None:
                if broken:
                    self._held -= self._plan_bytes.pop(id(plan), 0)
                    plan.close()
                else:
                    self._idle.setdefault(req.key, []).append(plan)
                    self._idle.move_to_end(req.key)
            self._lock.notify_all()


def latency_summary(results):
    lat = np.array([r.latency for r in results]) * 1e3
    return dict(p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), n=int(lat.size))
