"""MRC / MRCS stack files: the format side of ``loadData``/``saveData``
(``zorro/zorro.py:3413-3426``, ``:3583-3642``), which the reference reads and
writes through the external ``mrcz`` package.

Only the header and the sample layout are handled here (host side); the
samples of integer stacks stay integers and are converted (times the gain
reference, binned) on the GPU by ``MoviePlan.set_frames_raw``.  Supported
modes (MRC2014, with the IMOD/SerialEM conventions for packed data):

* 0  8-bit (unsigned when the IMOD stamp's flag bit 0 is set, as SerialEM
  writes counting data; signed otherwise)
* 1  int16, 2 float32, 6 uint16, 12 float16
* 101, or 0 with IMOD flag bit 4 (16): 4-bit unsigned counts, two per byte,
  the first pixel in the low nibble, each row padded to whole bytes

The header's cell length over the sampling gives the pixel size in Angstrom;
``read_mrc`` reports it in nm as ``mrcz.readMRC(..., pixelunits='nm')`` does.
``mrcz``'s blosc-compressed ``.mrcz`` files are not readable here (no blosc).
"""
from __future__ import annotations

import os
import struct

import numpy as np

_IMOD_STAMP = 1146047817
_MODES = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}


class MRCHeader(dict):
    """Parsed header: nx, ny, nz, mode, pixelsize (nm, per axis z, y, x),
    voltage (kV, 0 if absent), packed4 (4-bit samples), unsigned8, ext_bytes."""


def _parse_header(buf: bytes) -> MRCHeader:
    if len(buf) < 1024:
        raise ValueError("MRC header shorter than 1024 bytes")
    endian = "<"
    nx, ny, nz, mode = struct.unpack(endian + "4i", buf[0:16])
    if not (0 <= mode < 1 << 16):
        endian = ">"
        nx, ny, nz, mode = struct.unpack(endian + "4i", buf[0:16])
    mx, my, mz = struct.unpack(endian + "3i", buf[28:40])
    cx, cy, cz = struct.unpack(endian + "3f", buf[40:52])
    nsymbt = struct.unpack(endian + "i", buf[92:96])[0]
    stamp, flags = struct.unpack(endian + "ii", buf[152:160])
    imod = stamp == _IMOD_STAMP
    h = MRCHeader(nx=nx, ny=ny, nz=nz, mode=mode, ext_bytes=max(nsymbt, 0), endian=endian)
    h["packed4"] = mode == 101 or (mode == 0 and imod and bool(flags & 16))
    h["unsigned8"] = mode == 0 and imod and bool(flags & 1)
    if h["packed4"] and mode == 0:
        h["nx"] = nx * 2        # SerialEM mode 0 + flag 16: the header counts bytes
    ps = [(c / m) if m else 0.0 for c, m in ((cz, mz), (cy, my), (cx, mx))]
    h["pixelsize"] = [p / 10.0 for p in ps]   # Angstrom -> nm
    h["voltage"] = 0.0
    return h


def read_mrc(path: str, unpack4: bool = True, memmap: bool = False):
    """Read an MRC/MRCS file: (data [nz, ny, nx], header).  4-bit files give
    uint8 counts (``unpack4``) or the packed bytes [nz, ny, ceil(nx/2)] for
    ``MoviePlan.set_frames_raw(..., packed4=True)``; 8-bit files give uint8
    (SerialEM's unsigned counts) or int8."""
    if os.path.splitext(path)[1].lower() in (".mrcz", ".mrczs"):
        raise NotImplementedError("blosc-compressed .mrcz files need the blosc package, which is absent")
    with open(path, "rb") as f:
        head = f.read(1024)
        h = _parse_header(head)
    nx, ny, nz, e = h["nx"], h["ny"], h["nz"], h["endian"]
    off = 1024 + h["ext_bytes"]
    if h["packed4"]:
        row = (nx + 1) // 2
        raw = np.memmap(path, np.uint8, "r", off, (nz, ny, row)) if memmap else \
            np.fromfile(path, np.uint8, nz * ny * row, offset=off).reshape(nz, ny, row)
        if not unpack4:
            return raw, h
        out = np.empty((nz, ny, 2 * row), np.uint8)
        out[..., 0::2] = raw & 15
        out[..., 1::2] = raw >> 4
        return out[..., :nx], h
    if h["mode"] not in _MODES:
        raise NotImplementedError(f"MRC mode {h['mode']} (complex or unknown) is not a movie stack")
    dt = np.dtype(_MODES[h["mode"]])
    if h["mode"] == 0 and h["unsigned8"]:
        dt = np.dtype(np.uint8)
    dt = dt.newbyteorder(e) if dt.itemsize > 1 else dt
    data = np.memmap(path, dt, "r", off, (nz, ny, nx)) if memmap else \
        np.fromfile(path, dt, nz * ny * nx, offset=off).reshape(nz, ny, nx)
    if dt.byteorder == ">":
        data = data.astype(dt.newbyteorder("<"))
    return data, h


def pack4(counts: np.ndarray) -> np.ndarray:
    """uint8 counts (< 16) [..., nx] -> packed bytes [..., ceil(nx/2)], the
    first pixel of each pair in the low nibble (IMOD mode 101)."""
    c = np.asarray(counts)
    if c.size and int(c.max()) > 15:
        raise ValueError("4-bit packing needs counts below 16")
    nx = c.shape[-1]
    if nx % 2:
        c = np.concatenate([c, np.zeros(c.shape[:-1] + (1,), c.dtype)], axis=-1)
    c = c.astype(np.uint8)
    return (c[..., 0::2] | (c[..., 1::2] << 4)).astype(np.uint8)


def write_mrc(path: str, data: np.ndarray, pixelsize_nm: float = 0.0, mode: int | None = None,
              voltage_kv: float = 0.0, label: str = "zorro_amd") -> None:
    """Write a stack or image as MRC2014 (little endian).  ``mode``: None picks
    2 (float32), 6 (uint16), 0 (uint8, IMOD-stamped unsigned) or 1 (int16)
    from the dtype; 101 packs uint8 counts into 4 bits."""
    a = np.asarray(data)
    if a.ndim == 2:
        a = a[None]
    nz, ny, nx = a.shape
    if mode is None:
        mode = {np.dtype(np.float32): 2, np.dtype(np.uint16): 6, np.dtype(np.uint8): 0,
                np.dtype(np.int16): 1, np.dtype(np.float16): 12}.get(a.dtype, 2)
    hdr = bytearray(1024)
    struct.pack_into("<4i", hdr, 0, nx, ny, nz, mode)
    struct.pack_into("<3i", hdr, 28, nx, ny, nz)
    a_ang = 10.0 * float(pixelsize_nm)
    struct.pack_into("<3f", hdr, 40, a_ang * nx, a_ang * ny, a_ang * nz)
    struct.pack_into("<3f", hdr, 52, 90.0, 90.0, 90.0)
    struct.pack_into("<3i", hdr, 64, 1, 2, 3)
    if mode in (2, 12):
        body = a.astype(np.float32 if mode == 2 else np.float16)
    elif mode == 6:
        body = a.astype(np.uint16)
    elif mode == 1:
        body = a.astype(np.int16)
    elif mode == 0:
        body = a.astype(np.uint8)
    elif mode == 101:
        body = pack4(a)
    else:
        raise NotImplementedError(f"writing MRC mode {mode}")
    fa = body.astype(np.float64) if mode != 101 else a.astype(np.float64)
    struct.pack_into("<3f", hdr, 76, float(fa.min()), float(fa.max()), float(fa.mean()))
    struct.pack_into("<i", hdr, 92, 0)
    flags = 1 if mode == 0 else 0
    struct.pack_into("<ii", hdr, 152, _IMOD_STAMP, flags)
    hdr[208:212] = b"MAP "
    hdr[212:216] = bytes([0x44, 0x44, 0, 0])
    struct.pack_into("<f", hdr, 216, float(fa.std()))
    struct.pack_into("<i", hdr, 220, 1)
    lab = label.encode("ascii", "replace")[:80]
    hdr[224:224 + len(lab)] = lab
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(np.ascontiguousarray(body).astype(body.dtype.newbyteorder("<")).tobytes())
