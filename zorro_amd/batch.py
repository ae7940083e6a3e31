"""Batches: several movies of one detector shape registered in single launches.

``BatchPlan`` wraps a ``zr_batch`` (include/zorro_hip.h): the forward spectra
of all frames, the pair transforms of every movie's pairs (a ragged pair list)
and the peak searches run as one launch each over the batch, and the movies'
solver, shift/filter and FRC stages run on streams of their own, joined back
into one captured HIP graph.  The reference runs one movie per process and
scales by processes (automator/zorroSkulkManager.py:1089-1113); a batch is
the GPU counterpart for movies that arrive together.  Each movie's results
equal those of a ``MoviePlan`` of its own (tests/test_gpu_batch.py).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .engine import MoviePlan


class BatchPlan:
    """Device plan for movies of ``n_frames[i]`` frames of ``h`` x ``w`` with
    one set of ``ImageRegistrator`` settings (``engine.make_params``)."""

    def __init__(self, params: L.ZrParams, n_frames, h: int, w: int, device: int = 0):
        self.lib = L.load()
        self.params = params
        self.n_frames = [int(v) for v in n_frames]
        self.h, self.w = int(h), int(w)
        arr = (C.c_int * len(self.n_frames))(*self.n_frames)
        self._b = C.c_void_p()
        L.check(self.lib.zr_batch_create(int(device), C.byref(params), len(self.n_frames), arr, self.h, self.w,
                                         C.byref(self._b)))
        self.movies = []
        for i, n in enumerate(self.n_frames):
            v = C.c_void_p()
            L.check(self.lib.zr_batch_movie(self._b, i, C.byref(v)))
            self.movies.append(MoviePlan(params, n, self.h, self.w, device=device, _handle=v))

    def __len__(self):
        return len(self.n_frames)

    def close(self):
        if self._b:
            for m in self.movies:
                m.close()
            self.lib.zr_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def frames_ptr(self, i):
        """Device address of movie i's frame slot (f32 [n_i, h, w])."""
        p = C.c_void_p()
        L.check(self.lib.zr_batch_frames(self._b, int(i), C.byref(p)))
        return p.value

    def set_frames(self, i, frames):
        """Copy movie i's frames in: NumPy float32 [n_i, h, w] or a float32 CUDA tensor."""
        want = (self.n_frames[i], self.h, self.w)
        if hasattr(frames, "data_ptr") and getattr(frames, "is_cuda", False):
            if str(frames.dtype) != "torch.float32" or not frames.is_contiguous() or tuple(frames.shape) != want:
                raise ValueError(f"movie {i}: frames must be a contiguous float32 CUDA tensor of shape {want}")
            import torch
            torch.cuda.current_stream(frames.device).synchronize()    # the producer's kernels have written it
            L.check(self.lib.zr_batch_set_frames(self._b, int(i), C.c_void_p(frames.data_ptr()), 1))
        else:
            a = np.ascontiguousarray(frames, dtype=np.float32)
            if a.shape != want:
                raise ValueError(f"movie {i}: frames shape {a.shape} != {want}")
            L.check(self.lib.zr_batch_set_frames(self._b, int(i), C.c_void_p(a.ctypes.data), 0))

    def use_graph(self, on: bool = True):
        L.check(self.lib.zr_batch_use_graph(self._b, 1 if on else 0))

    def align(self):
        L.check(self.lib.zr_batch_align(self._b))

    def synchronize(self):
        L.check(self.lib.zr_batch_synchronize(self._b))

    def device_bytes(self):
        nb = C.c_size_t()
        L.check(self.lib.zr_batch_device_bytes(self._b, C.byref(nb)))
        return int(nb.value)
