"""``.zor`` configuration/result files of ``ImageRegistrator``: INI sections
``initialization``, ``calibration``, ``registration``, ``ctf``, ``io``, ``plot``,
``results``, ``data`` and one ``errorDict<k>`` per entry of ``errorDictList``,
values as text or JSON, read and written with ``configparser`` (case-preserving
keys).  The layout is the one ``saveConfig``/``loadConfig`` use
(``zorro/zorro.py:3682-4087``) so files pass between the two implementations.

Differences from the reference, deliberate: integers (``n_threads``,
``subPixReg``, ``maxShift``, ...) are written (the reference casts them with
``np.int``, which current numpy lacks, and silently drops them); NumPy scalars
and arrays anywhere in a JSON value are converted to plain numbers and lists;
``doLazyFRC`` is read back into ``doLazyFRC`` (the reference assigns it to an
unused ``doFRC``, ``:3860``).
"""
from __future__ import annotations

import configparser
import json
import os

import numpy as np

# (section, key, kind) in the order the reference writes them; kind: s text,
# i int, f float, b bool, j JSON, a JSON array (read back as np.ndarray)
_FIELDS = [
    ("initialization", "METAstatus", "s"), ("initialization", "verbose", "i"), ("initialization", "umask", "i"),
    ("initialization", "fftw_effort", "s"), ("initialization", "n_threads", "i"), ("initialization", "saveC", "b"),
    ("initialization", "cachePath", "s"),
    ("calibration", "pixelsize", "f"), ("calibration", "voltage", "f"), ("calibration", "C3", "f"),
    ("calibration", "gain", "f"), ("calibration", "detectorPixelSize", "f"), ("calibration", "gainInfo", "j"),
    ("registration", "xcorrMode", "s"), ("registration", "triMode", "s"), ("registration", "shapePadded", "a"),
    ("registration", "shapeOriginal", "a"), ("registration", "shapeBinned", "a"), ("registration", "fouCrop", "a"),
    ("registration", "subPixReg", "i"), ("registration", "shiftMethod", "s"), ("registration", "maxShift", "i"),
    ("registration", "preShift", "b"), ("registration", "diagStart", "i"), ("registration", "diagWidth", "i"),
    ("registration", "autoMax", "i"), ("registration", "startFrame", "i"), ("registration", "endFrame", "i"),
    ("registration", "peaksigThres", "f"), ("registration", "corrThres", "f"), ("registration", "velocityThres", "f"),
    ("registration", "Brad", "f"), ("registration", "Bmode", "s"), ("registration", "BfiltType", "s"),
    ("registration", "originMode", "s"), ("registration", "weightMode", "s"), ("registration", "logisticK", "f"),
    ("registration", "logisticNu", "f"), ("registration", "suppressOrigin", "b"), ("registration", "filterMode", "s"),
    ("registration", "doLazyFRC", "b"), ("registration", "doEvenOddFRC", "b"), ("registration", "doseFiltParam", "j"),
    ("ctf", "CTFProgram", "s"), ("ctf", "CTFInfo", "j"),
    ("io", "savePNG", "b"), ("io", "compress_ext", "s"), ("io", "saveMovie", "b"), ("io", "doCompression", "b"),
    ("io", "files", "j"),
    ("plot", "plotDict", "j"),
    ("results", "translations", "a"), ("results", "transEven", "a"), ("results", "transOdd", "a"),
    ("results", "rotations", "a"), ("results", "scales", "a"), ("results", "velocities", "a"), ("results", "FRC", "a"),
    ("data", "hotpixInfo", "j"), ("data", "trackCorrStats", "b"), ("data", "corrStats", "j"), ("data", "bench", "j"),
]
# the writer spells this key 'BFiltType'; the reader looks for 'BfiltType' (zorro/zorro.py:3849, :3981)
_WRITE_NAME = {("registration", "BfiltType"): "BFiltType"}
_ALWAYS = {"METAstatus", "verbose", "umask", "fftw_effort", "saveC", "cachePath", "pixelsize", "voltage", "C3",
           "gain", "detectorPixelSize", "gainInfo", "xcorrMode", "triMode", "shiftMethod", "preShift", "peaksigThres",
           "corrThres", "velocityThres", "Brad", "Bmode", "BfiltType", "originMode", "weightMode", "logisticK",
           "logisticNu", "suppressOrigin", "filterMode", "doLazyFRC", "doEvenOddFRC", "CTFProgram", "CTFInfo",
           "savePNG", "compress_ext", "saveMovie", "doCompression", "files", "plotDict", "hotpixInfo",
           "trackCorrStats", "corrStats", "bench"}


def _plain(v):
    """NumPy values -> JSON-serialisable Python values."""
    if isinstance(v, dict):
        return {str(k): _plain(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_plain(x) for x in v]
    if isinstance(v, np.ndarray):
        return _plain(v.tolist())
    if isinstance(v, np.generic):
        return v.item()
    return v


def _text(v, kind):
    if kind in ("j", "a"):
        return json.dumps(_plain(v))
    if kind == "i" and v is not None:
        return str(int(v))
    if kind == "f" and isinstance(v, np.generic):
        return str(v.item())
    return str(v) if v is not None else "None"


def save_config(reg, path: str) -> str:
    cp = configparser.RawConfigParser(allow_no_value=True)
    cp.optionxform = str
    for sec in ("initialization", "calibration", "registration", "ctf", "io", "plot", "results", "data"):
        cp.add_section(sec)
    for sec, key, kind in _FIELDS:
        v = getattr(reg, key, None)
        if key not in _ALWAYS and v is None:
            continue
        if key == "doseFiltParam" and not (np.any(np.asarray([x for x in v if x is not None], dtype=float))
                                           and reg.filterMode):
            continue
        cp.set(sec, _WRITE_NAME.get((sec, key), key), _text(v, kind))
    for k, ed in enumerate(reg.errorDictList):
        name = f"errorDict{k}"
        cp.add_section(name)
        keys = list(ed.keys())
        cp.set(name, "keyList", json.dumps(keys))
        for key in keys:
            cp.set(name, key, json.dumps(_plain(ed[key])))
    d = os.path.dirname(os.path.realpath(path))
    if d and not os.path.isdir(d):
        os.makedirs(d)
    with open(path, "w") as fh:
        cp.write(fh)
    return path


def _none_or(s, conv):
    return None if s in (None, "None", "") else conv(s)


def load_config(reg, path: str) -> None:
    if not os.path.isfile(path):
        raise IOError(f"loadConfig: could not load config file {path}")
    cp = configparser.RawConfigParser(allow_no_value=True)
    cp.optionxform = str
    cp.read(path)
    for sec, key, kind in _FIELDS:
        opt = key if cp.has_option(sec, key) else _WRITE_NAME.get((sec, key))
        if opt is None or not cp.has_option(sec, opt):
            continue
        s = cp.get(sec, opt)
        try:
            if kind == "s":
                v = None if s == "None" else s
            elif kind == "i":
                v = _none_or(s, int)
            elif kind == "f":
                v = _none_or(s, float)
            elif kind == "b":
                v = cp.getboolean(sec, opt) if s not in ("None", "") else None
            elif kind == "a":
                v = json.loads(s)
                v = None if v is None else np.array(v)
            else:
                v = json.loads(s)
        except (ValueError, json.JSONDecodeError):
            continue                # unreadable values are skipped, as the reference does
        setattr(reg, key, v)
    k = 0                       # appended to the existing list, as loadConfig does (zorro/zorro.py:3778-3795)
    while cp.has_section(f"errorDict{k}") and cp.has_option(f"errorDict{k}", "keyList"):
        name = f"errorDict{k}"
        ed = {}
        for key in json.loads(cp.get(name, "keyList")):
            raw = json.loads(cp.get(name, key))
            try:
                a = np.array(raw)
            except ValueError:          # ragged lists stay lists
                ed[key] = raw
                continue
            ed[key] = a.item(0) if a.size == 1 else a
        reg.errorDictList.append(ed)
        k += 1
