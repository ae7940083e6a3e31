"""The C-ABI library builds, loads and exports every symbol include/zorro_hip.h
declares (CPU only: no compute calls)."""
import ctypes
import os

import pytest

from zorro_amd import _lib as L


def test_library_exports_every_header_symbol():
    if not os.path.exists(L.LIB_PATH):
        pytest.fail("libzorro_hip.so not built (run python -m zorro_amd.build)")
    lib = ctypes.CDLL(L.LIB_PATH)
    names = L.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header_and_params_layout(tmp_path):
    assert set(L.header_symbols()) == set(L.SIGNATURES)
    # zr_params: every ctypes field offset equals the C compiler's offsetof
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler to check the struct layout")
    names = [f[0] for f in L.ZrParams._fields_]
    src = tmp_path / "layout.c"
    inc = os.path.join(os.path.dirname(L.LIB_PATH), "..", "include")
    body = "".join(f'printf("%zu\\n", offsetof(zr_params, {n}));' for n in names)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "zorro_hip.h"\n'
                   'int main(void){' + body + 'printf("%zu\\n", sizeof(zr_params)); return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", inc, str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[:-1] == [getattr(L.ZrParams, n).offset for n in names]
    assert out[-1] == ctypes.sizeof(L.ZrParams)


def test_error_path_without_gpu_compute():
    lib = L.load()
    assert lib.zr_last_error() is not None
    # a NULL movie is rejected with an argument error, no device work
    assert lib.zr_xcorr(None) == L.ZR_ERR_ARG
    assert b"NULL" in lib.zr_last_error()
