"""C-ABI checks that need no GPU: the HIP library loads, exports every entry
point include/grhd_abi.h declares, and the ctypes structs match the C
layout (sizeof / offsetof printed by a host program compiled with gcc
against the header)."""
import ctypes
import os
import re
import subprocess

import pytest

from grdzhadzha_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "grhd_abi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(grhd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers exactly the declared entry points
    assert set(abi.SIGNATURES) == set(names)


STRUCTS = {
    "grhd_box": abi.GrhdBox,
    "grhd_params": abi.GrhdParams,
    "grhd_background": abi.GrhdBackground,
    "grhd_diagnostics": abi.GrhdDiagnostics,
    "grhd_error": abi.GrhdError,
    "grhd_level_box": abi.GrhdLevelBox,
}


def test_struct_layouts_match_header(tmp_path):
    src = ['#include <stdio.h>', '#include <stddef.h>',
           '#include "%s"' % HEADER, 'int main(void) {']
    for cname, cls in STRUCTS.items():
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for field, _ in cls._fields_:
            src.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));'
                       % (cname, field, cname, field))
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(c)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    for cname, cls in STRUCTS.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert int(got[cname + "." + field]) == \
                getattr(cls, field).offset, (cname, field)


def test_constants_match_header():
    text = open(HEADER).read()
    for name in ("GRHD_OK", "GRHD_ERR_RECOVERY_GHOST_FILL",
                 "GRHD_ERR_RECOVERY_RHS", "GRHD_ERR_RECOVERY_CFL",
                 "GRHD_ERR_INVALID_ARGUMENT", "GRHD_MASK_FAILED",
                 "GRHD_MASK_FLOORED", "GRHD_MASK_RESET",
                 "GRHD_MASK_SOURCE_ATMOSPHERE", "GRHD_BG_KERR_SCHILD",
                 "GRHD_BG_CIRCULAR_BINARY_PN", "GRHD_TAU", "GRHD_V3"):
        m = re.search(r"\b%s\s*=\s*(\d+)" % name, text)
        assert m, name
        assert int(m.group(1)) == getattr(abi, name), name


def test_product_fails_loudly_without_library(tmp_path):
    with pytest.raises(RuntimeError, match="not built"):
        abi.load_library(str(tmp_path / "missing.so"))


def test_bench_refuses_multi_gpu():
    res = subprocess.run(["python", os.path.join(ROOT, "bench.py"), "--gpus",
                          "2"], capture_output=True, text=True)
    assert res.returncode == 2
    assert "one GPU" in res.stderr
