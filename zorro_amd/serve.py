"""Continuous-batching movie server on one MI355X (Python face of zr_server).

The reference registers one movie per process and scales by running several
processes (the Skulk manager, ``automator/zorroSkulkManager.py:917-1203``).
Here one process keeps a pool of movie plans resident in HBM and runs several
movies at once, each on its own HIP stream, so that the latency-bound stages
of one movie (the one-block shift solver, the logistic fit) overlap with the
bandwidth-bound pair transforms of another.

The scheduler is native (``zr_server_*`` in ``include/zorro_hip.h``): requests
queue in arrival order and are admitted when an in-flight slot is free and
their plan fits the HBM budget (an idle plan of the same key is reused,
otherwise one is created after evicting idle plans of other keys, least
recently used first; a movie that fits may go ahead of a larger head of the
queue); plans are created and movies run in the server's worker threads.  This
module turns requests into futures: ``submit`` queues a movie and returns a
``concurrent.futures.Future`` of a ``MovieResult``; a completion thread waits
on the server (the ctypes call releases the GIL) and settles the futures.

Results do not depend on what else is in flight: every kernel of a movie reads
and writes only that movie's buffers, and no reduction depends on scheduling
(tests/test_gpu_serve.py checks alone == in the mix).  ``plan_class`` swaps
the plans for host-side stand-ins with MoviePlan's interface, which drive the
same native scheduler without a GPU (tests/test_serve_admission.py).
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from concurrent.futures import Future

import numpy as np

from . import _lib as L
from .engine import make_params


def plan_key(n, h, w, params):
    return (int(n), int(h), int(w), bytes(memoryview(params)))


def estimate_plan_bytes(n, h, w, params) -> int:
    """The server's upper estimate of a plan's device memory before it exists
    (``zr_server_estimate_bytes``; the exact figure, ``zr_movie_device_bytes``,
    replaces it once a plan of the key has been created)."""
    return int(L.load().zr_server_estimate_bytes(C.byref(params), int(n), int(h), int(w)))


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


class PythonPlanOps:
    """zr_plan_ops over a Python plan class with MoviePlan's interface
    (``cls(params, n, h, w, device=)``, ``set_frames``, ``align``,
    ``synchronize``, ``solution``, ``frc_curve``, ``close`` and
    ``lib.zr_movie_device_bytes``).  Frames reach the plan as a numpy view of
    the submitted host array."""

    def __init__(self, plan_class, device=0):
        self.cls = plan_class
        self.device = device
        self.plans = {}
        self.lock = threading.Lock()
        self.next = 1
        self.ops = L.ZrPlanOps(None, L.PLAN_ESTIMATE(self._estimate), L.PLAN_CREATE(self._create),
                               L.PLAN_RUN(self._run), L.PLAN_DESTROY(self._destroy))

    @staticmethod
    def _estimate(ctx, params, n, h, w):
        return estimate_plan_bytes(n, h, w, params.contents)

    def _create(self, ctx, params, n, h, w, plan_out, bytes_out):
        try:
            p = L.ZrParams()
            C.memmove(C.byref(p), params, C.sizeof(L.ZrParams))
            plan = self.cls(p, n, h, w, device=self.device)
            nb = C.c_size_t()
            L.check(plan.lib.zr_movie_device_bytes(plan._mv, C.byref(nb)))
            with self.lock:
                handle, self.next = self.next, self.next + 1
                self.plans[handle] = plan
            plan_out[0] = handle
            bytes_out[0] = int(nb.value) + n * h * w * 4
            return L.ZR_OK
        except MemoryError:
            return L.ZR_ERR_NOMEM
        except Exception:
            return L.ZR_ERR_HIP

    def _run(self, ctx, handle, job):
        try:
            j = job.contents
            plan = self.plans[handle]
            frames = np.ctypeslib.as_array(C.cast(j.frames, C.POINTER(C.c_float)), shape=(j.n, j.h, j.w))
            plan.set_frames(frames)
            plan.align()
            plan.synchronize()
            t = np.asarray(plan.solution()["translations"], np.float64).reshape(j.n, 2)
            np.ctypeslib.as_array(j.translations, shape=(j.n, 2))[:] = t
            if j.params.contents.do_frc and j.frc_n:
                f = np.asarray(plan.frc_curve(), np.float64)[: j.frc_cap]
                np.ctypeslib.as_array(j.frc, shape=(j.frc_cap,))[: f.size] = f
                j.frc_n[0] = f.size
            return L.ZR_OK
        except Exception:
            return L.ZR_ERR_HIP

    def _destroy(self, ctx, handle):
        with self.lock:
            plan = self.plans.pop(handle, None)
        if plan is not None:
            plan.close()


class _ServerFuture(Future):
    """A request's future: cancel() succeeds while the request is still queued
    in the native server."""

    def __init__(self, server, rid):
        super().__init__()
        self._server, self._rid = server, rid

    def cancel(self):
        if self.done():
            return super().cancel()
        if self._server._lib.zr_server_cancel(self._server._h, C.c_longlong(self._rid)) != L.ZR_OK:
            return False
        return super().cancel()


class MovieServer:
    """Serve registration requests of mixed shapes on one device.

    ``hbm_budget``: bytes of device memory the server may hold in plans
    (default: 85% of the free memory at start).  ``max_inflight``: movies
    running at once (each on its own stream).  ``return_sums``: "device"
    (torch tensors), "host" (numpy) or "none"."""

    FRC_CAP = 2048

    def __init__(self, device: int = 0, hbm_budget: int | None = None, max_inflight: int = 3,
                 return_sums: str = "device", plan_class=None):
        self._lib = L.load()
        self.device = int(device)
        self.max_inflight = int(max_inflight)
        self.return_sums = return_sums
        self._ops = PythonPlanOps(plan_class, self.device) if plan_class is not None else None
        if self._ops is None:
            import torch
            self.torch = torch
            self.dev = torch.device("cuda", self.device)
            if hbm_budget is None:
                free, _ = torch.cuda.mem_get_info(self.dev)
                hbm_budget = int(free * 0.85)
        elif hbm_budget is None:
            raise ValueError("a server over stand-in plans needs an HBM budget")
        self.budget = int(hbm_budget)
        h = C.c_void_p()
        if self._ops is None:
            L.check(self._lib.zr_server_create(self.device, C.c_size_t(self.budget), self.max_inflight, C.byref(h)))
        else:
            L.check(self._lib.zr_server_create_ops(C.byref(self._ops.ops), C.c_size_t(self.budget),
                                                   self.max_inflight, C.byref(h)))
        self._h = h
        self._lock = threading.Lock()
        self._outstanding = {}                 # id -> (future, request record)
        self._closing = False
        self._closed = False
        self._done_thread = threading.Thread(target=self._complete_loop, name="zr-complete", daemon=True)
        self._done_thread.start()

    # ------------------------------------------------------------ public
    def submit(self, frames, arrival: float | None = None, **settings) -> Future:
        """Queue one movie: frames f32 [N,H,W] (numpy or a CUDA tensor kept
        alive by the caller until the future completes) and ImageRegistrator
        settings as keywords (``engine.make_params``).  ``arrival`` is a
        ``time.perf_counter()`` value (the server's clock; default now)."""
        params = make_params(**settings)
        n, h, w = (int(v) for v in frames.shape)
        keep = [frames]
        if self._ops is None:
            torch = self.torch
            if not isinstance(frames, torch.Tensor):
                frames = torch.from_numpy(np.ascontiguousarray(frames, np.float32)).to(self.dev)
                keep.append(frames)
            if frames.dtype != torch.float32 or not frames.is_contiguous() or frames.device != self.dev:
                raise ValueError(f"frames must be a contiguous float32 tensor on {self.dev}")
            fptr = frames.data_ptr()
            img = filt = None
            if self.return_sums != "none":
                img = torch.empty((h, w), dtype=torch.float32, device=self.dev)
                if params.n_filters or params.do_dose:
                    filt = torch.empty((h, w), dtype=torch.float32, device=self.dev)
        else:
            a = np.ascontiguousarray(frames, np.float32)
            keep.append(a)
            fptr = a.ctypes.data
            img = filt = None
        rid = C.c_longlong()
        with self._lock:
            if self._closing:
                raise RuntimeError("server is closed")
            L.check(self._lib.zr_server_submit(
                self._h, C.byref(params), n, h, w, C.c_void_p(fptr),
                C.c_void_p(img.data_ptr() if img is not None else 0),
                C.c_void_p(filt.data_ptr() if filt is not None else 0),
                C.c_double(time.perf_counter() if arrival is None else float(arrival)), C.byref(rid)))
            fut = _ServerFuture(self, rid.value)
            self._outstanding[rid.value] = (fut, dict(keep=keep, img=img, filt=filt, n=n, params=params,
                                                      key=plan_key(n, h, w, params)))
        return fut

    def warm(self, frames_like, count: int = 1, **settings):
        """Create ``count`` idle plans for movies shaped like ``frames_like``
        ([N, H, W], or the shape itself) with these settings, ahead of the
        requests (serving warm-up: no request then waits for a plan)."""
        shape = tuple(int(v) for v in getattr(frames_like, "shape", frames_like))
        params = make_params(**settings)
        made = C.c_int()
        L.check(self._lib.zr_server_warm(self._h, C.byref(params), *shape, int(count), C.byref(made)))
        return made.value

    @property
    def stats(self):
        v = (C.c_longlong * 8)()
        L.check(self._lib.zr_server_stats(self._h, v))
        keys = ("admitted", "plans_created", "plans_evicted", "peak_held", "held", "inflight", "queued", "failed")
        return dict(zip(keys, (int(x) for x in v)))

    def close(self):
        """Finish every queued and running movie, then free the plans."""
        with self._lock:
            if self._closed:
                return
            self._closing = True
            futs = [f for f, _ in self._outstanding.values()]
        for f in futs:
            try:
                f.exception()
            except BaseException:
                pass
        self._done_thread.join()
        self._lib.zr_server_destroy(self._h)
        self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ completion
    def _complete_loop(self):
        rid, state = C.c_longlong(), C.c_int()
        while True:
            L.check(self._lib.zr_server_wait_any(self._h, C.c_double(0.05), C.byref(rid), C.byref(state)))
            if rid.value == 0:
                with self._lock:
                    if self._closing and not self._outstanding:
                        return
                continue
            self._settle(rid.value, state.value)

    def _settle(self, rid, state):
        with self._lock:
            fut, rec = self._outstanding.pop(rid)
        n = rec["n"]
        t = np.zeros((n, 2), np.float64)
        frc = np.zeros(self.FRC_CAP, np.float64)
        nfrc = C.c_int()
        times = np.zeros(3, np.float64)
        msg = C.create_string_buffer(512)
        rc = self._lib.zr_server_result(self._h, C.c_longlong(rid), L.dptr(t), n, L.dptr(frc), self.FRC_CAP,
                                        C.byref(nfrc), L.dptr(times), msg, 512)
        self._lib.zr_server_release(self._h, C.c_longlong(rid))
        if state == L.ZR_REQ_CANCELLED:
            super(_ServerFuture, fut).cancel()
            return
        if rc != L.ZR_OK:
            text = msg.value.decode(errors="replace")
            fut.set_exception(MemoryError(text) if rc == L.ZR_ERR_NOMEM else L.ZrError(rc, text))
            return
        img, filt = rec["img"], rec["filt"]
        if self.return_sums == "host":
            img = img.cpu().numpy() if img is not None else None
            filt = filt.cpu().numpy() if filt is not None else None
        elif self.return_sums == "none":
            img = filt = None
        res = MovieResult(t, img, filt, frc[: nfrc.value].copy() if rec["params"].do_frc else None,
                          times[0], times[1], times[2], rec["key"])
        fut.set_result(res)


def latency_summary(results):
    lat = np.array([r.latency for r in results]) * 1e3
    return dict(p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), n=int(lat.size))
