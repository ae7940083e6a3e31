"""MRC stack files and .zor configuration files (the formats of loadData,
saveData, saveConfig and loadConfig, zorro/zorro.py:3268-4087).

mrcz (the reference's MRC library) is not vendored and no MRC or .zor file is
held by the reference, so the MRC reader is pinned on the MRC2014 header layout
written byte by byte here (words 1-4 sizes and mode, 8-10 sampling, 11-13 cell
in Angstrom, 24 extended-header bytes, 39-40 IMOD stamp and flags) and on the
IMOD 4-bit packing (first pixel in the low nibble); the .zor files on the
section/key layout of saveConfig (zorro/zorro.py:3912-4076)."""
import configparser
import json
import struct

import numpy as np
import pytest

from zorro_amd import mrc
from zorro_amd.registrator import ImageRegistrator


def _header(nx, ny, nz, mode, cell_ang=(0.0, 0.0, 0.0), ext=0, stamp=0, flags=0, endian="<"):
    h = bytearray(1024)
    struct.pack_into(endian + "4i", h, 0, nx, ny, nz, mode)
    struct.pack_into(endian + "3i", h, 28, nx, ny, nz)
    struct.pack_into(endian + "3f", h, 40, *cell_ang)
    struct.pack_into(endian + "i", h, 92, ext)
    struct.pack_into(endian + "ii", h, 152, stamp, flags)
    return bytes(h)


def test_hand_written_headers(tmp_path):
    rng = np.random.default_rng(3)
    # mode 6 uint16, 1.3 A pixels over x (cell = 1.3 * nx), 16 extended bytes
    a = rng.integers(0, 60000, (3, 5, 7)).astype("<u2")
    p = tmp_path / "u16.mrcs"
    p.write_bytes(_header(7, 5, 3, 6, (1.3 * 7, 1.3 * 5, 1.3 * 3), ext=16) + b"\x00" * 16 + a.tobytes())
    d, h = mrc.read_mrc(str(p))
    assert d.dtype == np.uint16 and np.array_equal(d, a)
    assert h["pixelsize"][-1] == pytest.approx(0.13, rel=1e-6)
    # big-endian float32
    f = rng.standard_normal((2, 4, 6)).astype(">f4")
    p = tmp_path / "be.mrc"
    p.write_bytes(_header(6, 4, 2, 2, endian=">") + f.tobytes())
    d, h = mrc.read_mrc(str(p))
    assert np.array_equal(d, f.astype("<f4")) and h["pixelsize"][-1] == 0.0
    # mode 0: signed without the IMOD flag, unsigned with flag bit 0
    b = rng.integers(0, 256, (2, 3, 4)).astype(np.uint8)
    p = tmp_path / "s8.mrc"
    p.write_bytes(_header(4, 3, 2, 0) + b.tobytes())
    assert np.array_equal(mrc.read_mrc(str(p))[0], b.view(np.int8))
    p.write_bytes(_header(4, 3, 2, 0, stamp=1146047817, flags=1) + b.tobytes())
    assert np.array_equal(mrc.read_mrc(str(p))[0], b)


def test_four_bit_packing(tmp_path):
    # pixel values 1 and 2 in one byte: low nibble first -> 0x21
    assert mrc.pack4(np.array([[1, 2, 15, 0]], np.uint8)).tolist() == [[0x21, 0x0F]]
    rng = np.random.default_rng(5)
    c = rng.integers(0, 16, (3, 4, 10)).astype(np.uint8)
    packed = mrc.pack4(c)
    # mode 101 (nx counts pixels) and SerialEM mode 0 + flag 16 (nx counts bytes)
    for hdr in (_header(10, 4, 3, 101), _header(5, 4, 3, 0, stamp=1146047817, flags=16 | 1)):
        p = tmp_path / "c4.mrc"
        p.write_bytes(hdr + packed.tobytes())
        d, h = mrc.read_mrc(str(p))
        assert h["packed4"] and d.shape == (3, 4, 10) and np.array_equal(d, c)
        raw, _ = mrc.read_mrc(str(p), unpack4=False)
        assert np.array_equal(raw, packed)
    # odd widths pad each row to whole bytes
    c7 = rng.integers(0, 16, (2, 3, 7)).astype(np.uint8)
    p = tmp_path / "odd.mrc"
    p.write_bytes(_header(7, 3, 2, 101) + mrc.pack4(c7).tobytes())
    assert np.array_equal(mrc.read_mrc(str(p))[0], c7)
    with pytest.raises(ValueError):
        mrc.pack4(np.array([16], np.uint8))


@pytest.mark.parametrize("dtype,mode", [(np.float32, 2), (np.uint16, 6), (np.int16, 1), (np.uint8, 0),
                                        (np.float16, 12), (np.uint8, 101)])
def test_write_read_round_trip(tmp_path, dtype, mode):
    rng = np.random.default_rng(mode)
    hi = 16 if mode == 101 else 200
    a = rng.integers(0, hi, (4, 6, 8)).astype(dtype)
    p = str(tmp_path / "rt.mrc")
    mrc.write_mrc(p, a, pixelsize_nm=0.1325, mode=mode)
    d, h = mrc.read_mrc(p)
    assert h["mode"] == mode and d.shape == a.shape
    assert np.array_equal(d.astype(np.float64), a.astype(np.float64))
    assert h["pixelsize"][-1] == pytest.approx(0.1325, rel=1e-6)
    with open(p, "rb") as fh:
        head = fh.read(1024)
    assert struct.unpack("<4i", head[:16]) == (8, 6, 4, mode)
    assert head[208:212] == b"MAP "


def test_mrcz_refused(tmp_path):
    with pytest.raises(NotImplementedError):
        mrc.read_mrc(str(tmp_path / "x.mrcz"))


def test_load_data_stack_and_gain(tmp_path):
    rng = np.random.default_rng(9)
    counts = rng.integers(0, 8, (3, 6, 8)).astype(np.uint8)
    gain = rng.uniform(0.9, 1.1, (8, 6)).astype(np.float32)       # rotated 90 degrees on load
    mrc.write_mrc(str(tmp_path / "movie.mrcs"), counts, pixelsize_nm=0.05)
    mrc.write_mrc(str(tmp_path / "gain.mrc"), gain)
    reg = ImageRegistrator()
    reg.gainInfo = {"Diagonal": True}
    reg.files["gainRef"] = str(tmp_path / "gain.mrc")
    reg.loadData(str(tmp_path / "movie.mrcs"))
    assert reg.images.dtype == np.uint8 and np.array_equal(reg.images, counts)
    assert reg.pixelsize == pytest.approx(0.05, rel=1e-6)
    assert np.array_equal(reg.gainRef, np.rot90(gain, k=1))
    with pytest.raises(NotImplementedError):
        reg.loadData(str(tmp_path / "movie.dm4"))


def test_save_data_writes_sums(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    reg = ImageRegistrator()
    reg.files["stack"] = "raw/m1.mrcs"
    reg.pixelsize = 0.1
    reg.imageSum = np.arange(12, dtype=np.float32).reshape(3, 4)
    reg.filtSum = -reg.imageSum
    reg.filterMode = "dose"
    reg.images = np.ones((2, 3, 4), np.float32)
    reg.saveData()
    s, h = mrc.read_mrc("sum/m1_zorro.mrc")
    assert np.array_equal(s[0], reg.imageSum) and h["pixelsize"][-1] == pytest.approx(0.1, rel=1e-6)
    assert np.array_equal(mrc.read_mrc("sum/m1_zorro_filt.mrc")[0][0], reg.filtSum)
    assert mrc.read_mrc("align/m1_zorro_movie.mrcs")[0].shape == (2, 3, 4)


def test_zor_round_trip(tmp_path):
    reg = ImageRegistrator()
    reg.files["stack"] = str(tmp_path / "m1.mrcs")
    reg.pixelsize, reg.maxShift, reg.triMode, reg.Bmode = 0.13, 40, "tri", "conv"
    reg.shapePadded, reg.fouCrop, reg.preShift, reg.filterMode = [4096, 4096], np.array([3072, 3072]), True, "dose"
    reg.n_threads = 8
    reg.translations = np.array([[0.0, 0.0], [1.5, -2.25]])
    reg.FRC = np.linspace(1, 0, 5)
    reg.hotpixInfo["guessHotpix"] = np.int64(3)
    reg.errorDictList = [{"shiftsTriMat": np.arange(4.0).reshape(2, 2), "M": 1, "solverInfo": {"a": np.float64(2)}}]
    path = reg.saveConfig()
    assert path == reg.files["stack"] + ".zor"
    cp = configparser.RawConfigParser(allow_no_value=True)
    cp.optionxform = str
    cp.read(path)
    # the reference's section and key layout
    assert cp.get("registration", "triMode") == "tri"
    assert json.loads(cp.get("registration", "fouCrop")) == [3072, 3072]
    assert json.loads(cp.get("results", "translations")) == [[0.0, 0.0], [1.5, -2.25]]
    assert json.loads(cp.get("errorDict0", "keyList")) == ["shiftsTriMat", "M", "solverInfo"]
    assert cp.get("registration", "BFiltType") == "gaussian"
    back = ImageRegistrator().loadConfig(path)
    assert back.pixelsize == 0.13 and back.maxShift == 40 and back.n_threads == 8 and back.preShift is True
    assert back.triMode == "tri" and back.Bmode == "conv" and back.filterMode == "dose"
    assert np.array_equal(back.shapePadded, [4096, 4096]) and np.array_equal(back.translations, reg.translations)
    assert np.allclose(back.FRC, reg.FRC) and back.hotpixInfo["guessHotpix"] == 3
    assert back.BfiltType == "gaussian" and back.corrThres is None
    ed = back.errorDictList[0]
    assert np.array_equal(ed["shiftsTriMat"], np.arange(4.0).reshape(2, 2)) and ed["M"] == 1
    assert ed["solverInfo"] == {"a": 2.0}
    with pytest.raises(IOError):
        ImageRegistrator().loadConfig(str(tmp_path / "missing.zor"))


def _tiff_bytes(pages, bits, sample_format):
    """A baseline little-endian multi-page TIFF written tag by tag (TIFF 6.0:
    one uncompressed strip per page, BitsPerSample, SampleFormat 1 = unsigned,
    3 = IEEE float), independent of the library under test."""
    n, h, w = pages.shape
    out = bytearray(b"II*\x00") + struct.pack("<I", 0)
    first_ifd_pos = 4
    prev_next = first_ifd_pos
    tags = lambda data_off: [(256, 3, 1, w), (257, 3, 1, h), (258, 3, 1, bits), (259, 3, 1, 1), (262, 3, 1, 1),
                             (273, 4, 1, data_off), (277, 3, 1, 1), (278, 3, 1, h),
                             (279, 4, 1, h * w * bits // 8), (339, 3, 1, sample_format)]
    for k in range(n):
        data_off = len(out)
        out += pages[k].astype(pages.dtype.newbyteorder("<")).tobytes()
        if len(out) % 2:
            out += b"\x00"
        ifd = len(out)
        struct.pack_into("<I", out, prev_next, ifd)
        t = tags(data_off)
        out += struct.pack("<H", len(t))
        for tag, typ, cnt, val in t:
            out += struct.pack("<HHI", tag, typ, cnt) + (struct.pack("<HH", val, 0) if typ == 3 else struct.pack("<I", val))
        prev_next = len(out)
        out += struct.pack("<I", 0)
    return bytes(out)


@pytest.mark.parametrize("dtype,bits,fmt", [(np.uint16, 16, 1), (np.uint8, 8, 1), (np.float32, 32, 3)])
def test_tiff_stack_read_from_hand_written_file(tmp_path, dtype, bits, fmt):
    """loadData of a .tif stack (zorro/zorro.py:3330-3344): every page, samples
    unchanged; 8/16-bit counts stay integers for the GPU input stage."""
    pytest.importorskip("PIL")
    rng = np.random.default_rng(bits)
    top = 256 if dtype == np.uint8 else 60000
    pages = (rng.integers(0, top, (3, 5, 7)) if dtype != np.float32 else rng.normal(0, 1e3, (3, 5, 7))).astype(dtype)
    path = tmp_path / "movie.tif"
    path.write_bytes(_tiff_bytes(pages, bits, fmt))
    reg = ImageRegistrator()
    reg.pixelsize = 0.1
    reg.loadData(str(path))
    assert reg.images.dtype == (np.float32 if dtype == np.float32 else dtype)
    assert np.array_equal(reg.images, pages)
    assert reg.pixelsize == 0.1          # the reference's TIFF path sets no pixel size


def test_tiff_write_read_round_trip(tmp_path):
    pytest.importorskip("PIL")
    from zorro_amd.tiff import read_tiff, write_tiff
    a = np.random.default_rng(3).integers(0, 4000, (4, 9, 6)).astype(np.uint16)
    write_tiff(str(tmp_path / "s.tiff"), a)
    b = read_tiff(str(tmp_path / "s.tiff"))
    assert b.dtype == np.uint16 and np.array_equal(a, b)
