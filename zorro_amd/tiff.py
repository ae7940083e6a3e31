"""TIFF movie stacks for ``ImageRegistrator.loadData``.

The reference reads ``.tif``/``.tiff`` stacks with ``skimage.io.imread(path,
plugin='pil')`` and casts them to float32 (``zorro/zorro.py:3330-3344``); here
Pillow reads the pages directly.  Integer samples (8- and 16-bit counts) are
kept as integers so that the GPU input stage converts them (the same values:
every such integer is exact in float32); 32-bit integer and floating-point
samples become float32.
"""
from __future__ import annotations

import numpy as np


def read_tiff(path: str) -> np.ndarray:
    """All pages of a TIFF file as [N, H, W] (uint8, uint16 or float32)."""
    try:
        from PIL import Image, ImageSequence
    except ImportError as e:          # Pillow is the reference's TIFF reader too
        raise NotImplementedError("TIFF stacks need Pillow") from e
    with Image.open(path) as im:
        pages = [np.asarray(p) for p in ImageSequence.Iterator(im)]
    if not pages:
        raise ValueError(f"{path}: no pages")
    shapes = {p.shape for p in pages}
    if len(shapes) != 1 or len(next(iter(shapes))) != 2:
        raise ValueError(f"{path}: pages of one 2D shape expected, got {sorted(shapes)}")
    a = np.stack(pages)
    if a.dtype.kind == "u" and a.dtype.itemsize <= 2:
        return np.ascontiguousarray(a.astype(a.dtype.newbyteorder("=")))
    return np.ascontiguousarray(a, dtype=np.float32)


def write_tiff(path: str, stack) -> None:
    """A [N, H, W] uint8, uint16 or float32 stack as a multi-page TIFF."""
    from PIL import Image
    a = np.asarray(stack)
    if a.ndim == 2:
        a = a[None]
    if a.dtype not in (np.uint8, np.uint16, np.float32):
        raise ValueError(f"TIFF pages of {a.dtype} are not written (uint8, uint16, float32)")
    pages = [Image.fromarray(np.ascontiguousarray(f)) for f in a]
    pages[0].save(path, save_all=True, append_images=pages[1:])
