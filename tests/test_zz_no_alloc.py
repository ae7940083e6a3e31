"""zr_tiled_frc_noalloc (the allocation-free tiledFRC) against zr_tiled_frc,
called through the C ABI on device buffers.

Bar: bit-identical curves (the same kernels, tables and per-group summation
order; the no-allocation call transforms its tiles a chunk at a time), on two
calls in a row where the second is smaller than the first, and the size limit
(box 1024) refused with a message.
"""
import ctypes as C

import numpy as np
import pytest

from zorro_amd import _lib as L

pytestmark = pytest.mark.gpu


def _images(h, w, seed):
    rng = np.random.default_rng(seed)
    common = rng.standard_normal((h, w))
    a = (3.0 + common + 0.7 * rng.standard_normal((h, w))).astype(np.float32)
    b = (3.0 + common + 0.7 * rng.standard_normal((h, w))).astype(np.float32)
    return a, b


def _call(fn, a, b, trans, box, overlap):
    import torch
    dev = torch.device("cuda", 0)
    ta = torch.from_numpy(a).to(dev)
    tb = torch.from_numpy(b).to(dev)
    out = torch.full((2048,), -1.0, dtype=torch.float64, device=dev)
    t = None if trans is None else np.ascontiguousarray(trans, np.float64)
    n = C.c_int()
    stream = torch.cuda.current_stream(dev).cuda_stream
    L.check(fn(C.c_void_p(ta.data_ptr()), C.c_void_p(tb.data_ptr()), a.shape[0], a.shape[1],
               None if t is None else t.ctypes.data_as(C.POINTER(C.c_double)), 0 if t is None else t.shape[0],
               box, overlap, C.c_void_p(out.data_ptr()), C.byref(n), C.c_void_p(stream)))
    torch.cuda.synchronize(dev)
    res = out.cpu().numpy()
    assert np.all(res[n.value:] == -1.0)          # nothing written past the curve
    return res[:n.value]


def test_noalloc_tiled_frc_equals_allocating_call_twice():
    lib = L.load()
    cases = [
        # first call: 9 box-512 tiles (chunks of 4, 4, 1 without allocation)
        ("1536sq_box512", 1536, 1536, 512, 0.5, np.array([[-3.2, 7.5], [2.0, -0.4]]), 31),
        # second call, smaller: 48 box-256 tiles
        ("1100x1300_box256", 1100, 1300, 256, 0.5, np.array([[1.0, -2.0]]), 32),
    ]
    for name, h, w, box, ov, trans, seed in cases:
        print(f"[no-alloc] case={name}", flush=True)
        a, b = _images(h, w, seed)
        ref = _call(lib.zr_tiled_frc, a, b, trans, box, ov)
        got = _call(lib.zr_tiled_frc_noalloc, a, b, trans, box, ov)
        print(f"[no-alloc] {name}: {got.size} rings, max abs diff {np.max(np.abs(got - ref)):.1e}", flush=True)
        assert got.shape == ref.shape and np.array_equal(got, ref)


def test_noalloc_tiled_frc_refuses_box_beyond_its_workspace():
    lib = L.load()
    a, b = _images(2200, 2200, 33)
    with pytest.raises(L.ZrError, match="exceeds 512"):
        _call(lib.zr_tiled_frc_noalloc, a, b, None, 1024, 0.5)
