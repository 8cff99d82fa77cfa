"""The C-ABI library loads and exports every function include/*.h declares (CPU only:
no compute call is made), and the ctypes struct layouts match the C headers."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(INC)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INC, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        # C++-only inline sections are not part of the exported ABI
        text = re.sub(r"#ifdef __cplusplus\s*\n(?!\s*extern).*?#endif", "", text, flags=re.S)
        for m in re.finditer(r"^[ \t]*(?:const\s+)?[A-Za-z_][\w \t\*]*?\b([A-Za-z_]\w*)\s*\([^;{}]*\)\s*;", text, re.M):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof"):
                continue
            names.add(name)
    return names


def test_library_exports_every_declared_function():
    from fft_demo_amd import _lib

    lib = _lib.load()
    names = declared_functions()
    assert {"mfft_plan_create_1d", "mfft_exec", "tcfftCreate", "tcfftExec", "setup", "tcfftExec_2d"} <= names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes signature in the binding
    assert names <= set(_lib.SIGNATURES), sorted(names - set(_lib.SIGNATURES))


def test_version_string():
    from fft_demo_amd import _lib

    assert b"gfx950" in _lib.load().mfft_version()


def test_struct_layouts_match_headers(tmp_path):
    from fft_demo_amd import _lib

    src = tmp_path / "sz.c"
    src.write_text('#include "mfft.h"\n#include "tcfft_half.h"\n#include "tcfft_half_2d.h"\n#include <stdio.h>\n'
                   'int main(){printf("%zu %zu %zu %zu %zu\\n", sizeof(mfft_desc), sizeof(mfft_info), '
                   'sizeof(mfft_pass_info), sizeof(tcfftHandle), sizeof(tcfftHandle_2d));}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", INC, str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(t) for t in (_lib.Desc, _lib.Info, _lib.PassInfo, _lib.TcfftHandle, _lib.TcfftHandle2d)]
    assert got == want


def test_missing_library_fails_loudly(tmp_path):
    from fft_demo_amd import _lib

    saved = _lib._lib
    try:
        _lib._lib = None
        with pytest.raises(ImportError):
            _lib.load(str(tmp_path / "nope.so"))
    finally:
        _lib._lib = saved
