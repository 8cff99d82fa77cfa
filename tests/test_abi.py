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
    """Every ctypes mirror has the C struct's size and each field's offset (a field added
    into padding keeps the size, so offsets are checked too)."""
    from fft_demo_amd import _lib

    pairs = [("mfft_desc", _lib.Desc), ("mfft_info", _lib.Info), ("mfft_pass_info", _lib.PassInfo),
             ("mfft_conv2d_desc", _lib.ConvDesc), ("mfft_conv2d_info", _lib.ConvInfo),
             ("mfft_batch_item", _lib.BatchItem), ("mfft_tune_entry", _lib.TuneEntry),
             ("mfft_server_desc", _lib.ServerDesc), ("mfft_request", _lib.Request),
             ("mfft_completion", _lib.Completion), ("mfft_server_stats", _lib.ServerStats),
             ("tcfftHandle", _lib.TcfftHandle), ("tcfftHandle_2d", _lib.TcfftHandle2d),
             ("mfft_launch_record", _lib.LaunchRecord)]
    lines = []
    want = []
    for cname, py in pairs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        want.append(ctypes.sizeof(py))
        for f in py._fields_:
            cfield = {"passes": "pass", "in_": "in"}.get(f[0], f[0])  # Python keywords renamed
            lines.append(f'printf("%zu\\n", offsetof({cname}, {cfield}));')
            want.append(getattr(py, f[0]).offset)
    # every C member has a ctypes field (a missing one can hide in padding)
    text = "".join(open(os.path.join(INC, h)).read() for h in ("mfft.h", "tcfft_half.h", "tcfft_half_2d.h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    for cname, py in pairs:
        body = re.search(r"typedef struct " + cname + r"\s*\{(.*?)\}\s*" + cname + ";", text, re.S).group(1)
        members = set()
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            for part in decl.split(","):
                members.add(re.findall(r"[A-Za-z_]\w*", re.sub(r"\[.*?\]", "", part))[-1])
        have = {{"passes": "pass", "in_": "in"}.get(f[0], f[0]) for f in py._fields_}
        assert members == have, (cname, members ^ have)
    src = tmp_path / "sz.c"
    src.write_text('#include "mfft.h"\n#include "tcfft_half.h"\n#include "tcfft_half_2d.h"\n#include <stdio.h>\n'
                   '#include <stddef.h>\nint main(){' + "".join(lines) + "}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", INC, str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
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
