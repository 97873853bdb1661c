"""The C-ABI library loads and exports every function include/gpucell/*.h
declares (no GPU needed: nothing here launches a kernel)."""
import ctypes
import glob
import os
import re

from gpucell_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDRS = sorted(glob.glob(os.path.join(ROOT, "include", "gpucell", "*.h")))


def declared_functions():
    text = "".join(open(h).read() for h in HDRS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gc_[A-Za-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = abi.load()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python prototypes cover exactly the declared set
    assert sorted(abi.PROTOTYPES) == names


def test_reference_entry_points_present():
    names = set(declared_functions())
    assert len(abi.REFERENCE_ENTRY_POINTS) == 20
    assert set(abi.REFERENCE_ENTRY_POINTS) <= names


def test_struct_layouts_match_header():
    assert ctypes.sizeof(abi.Structconst) == 96
    assert abi.Structconst.iSeed.offset == 24
    assert abi.Structconst.CutRadius.offset == 32
    assert abi.Structconst.Temp.offset == 88
    assert ctypes.sizeof(abi.gc_ext_consts) == 56
    assert abi.gc_ext_consts.noise_counter.offset == 48
    assert ctypes.sizeof(abi.int6) == 24 and ctypes.sizeof(abi.tfloat4) == 16


def test_host_only_calls():
    lib = abi.load()
    assert b"gfx950" in lib.gc_Version()
    # grid sizing is host code: the reference's default domain -> 20 x 22 x 22
    bbox = (ctypes.c_float * 6)(0, 20, -10, 10, -10, 10)
    per = (ctypes.c_int * 3)(1, 0, 0)
    nc = (ctypes.c_int * 3)()
    outer = (ctypes.c_float * 3)()
    inv = (ctypes.c_float * 3)()
    K = ctypes.c_int()
    assert lib.gc_ConfigDomain(bbox, 1.0, per, nc, outer, inv, ctypes.byref(K)) == 0
    assert list(nc) == [20, 22, 22] and K.value == 9680
    per2 = (ctypes.c_int * 3)(1, 1, 1)
    bbox2 = (ctypes.c_float * 6)(0, 2, 0, 10, 0, 10)
    assert lib.gc_ConfigDomain(bbox2, 1.0, per2, nc, outer, inv, ctypes.byref(K)) == -3
    assert b"periodic" in lib.gc_GetLastError()
    # a bad parameter block is rejected before any launch
    c = abi.Structconst()
    assert lib.gc_ConstsDataUp(ctypes.byref(c)) == -1
    e = abi.default_ext()
    assert list(e.periodic) == [1, 0, 0] and abs(e.wlc_ratio - 2.2) < 1e-6


def test_engine_struct_layouts_match_c(tmp_path):
    """gc_case_desc and gc_engine_info as the C compiler lays them out
    (the ctypes mirrors in gpucell_amd/abi.py must agree field by field)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    if cc is None:
        import pytest
        pytest.skip("gcc not found")
    src = tmp_path / "layout.c"
    src.write_text(r"""
#include <stddef.h>
#include <stdio.h>
#include "gpucell/gpucell.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(gc_case_desc), offsetof(gc_case_desc, consts),
           offsetof(gc_case_desc, ext), offsetof(gc_case_desc, skin), sizeof(gc_engine_info),
           offsetof(gc_engine_info, nrebuild), offsetof(gc_engine_info, skin));
    return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    want = [ctypes.sizeof(abi.gc_case_desc), abi.gc_case_desc.consts.offset, abi.gc_case_desc.ext.offset,
            abi.gc_case_desc.skin.offset, ctypes.sizeof(abi.gc_engine_info),
            abi.gc_engine_info.nrebuild.offset, abi.gc_engine_info.skin.offset]
    assert got == want
