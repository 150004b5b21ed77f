"""The C-ABI library loads without a GPU and exports every entry point that
include/apcloud.h declares (no compute calls here)."""
import ctypes
import os
import re

import apcloud

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(REPO, "include", "apcloud.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(apc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = apcloud.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) == set(apcloud.EXPORTS)


def test_library_is_built_for_gfx950():
    L = apcloud.lib()
    assert b"gfx950" in L.apc_version()
    blob = open(apcloud.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_create_without_device_fails_cleanly():
    if apcloud.device_count() > 0:
        return  # only meaningful on a machine without a GPU
    h = ctypes.c_void_p()
    r = apcloud.lib().apc_create(ctypes.byref(apcloud.make_params()), ctypes.byref(h))
    assert r == -10 and not h.value


def test_struct_layouts_match_header(tmp_path):
    # the ctypes mirrors must have the C compiler's sizes and offsets
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "apcloud.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(apc_params),'
        ' sizeof(apc_ksp_opts), sizeof(apc_solve_info), sizeof(apc_comp_view),'
        ' offsetof(apc_params, bv_G), offsetof(apc_ksp_opts, precision)); return 0;}\n')
    exe = tmp_path / "sz"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(apcloud.Params), ctypes.sizeof(apcloud.KspOpts),
            ctypes.sizeof(apcloud.SolveInfo), ctypes.sizeof(apcloud.CompView),
            apcloud.Params.bv_G.offset, apcloud.KspOpts.precision.offset]
    assert got == want
