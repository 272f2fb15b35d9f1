"""The C-ABI library loads and exports every symbol the public headers
declare (CPU only; no compute calls)."""
import ctypes
import os
import re

import ptamd
from ptamd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in ("pt_api.h", "pt_scenes.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"^int\s+(pt_\w+)\s*\(", text, re.M))
    return names


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 25
    lib = ctypes.CDLL(ptamd.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names <= set(ptamd.EXPORTED)


def test_struct_layouts():
    assert ptamd.TRI_DTYPE.itemsize == 40
    assert ptamd.INFO_DTYPE.itemsize == 64
    assert ctypes.sizeof(ptamd.Settings) == 16


def test_no_device_is_a_status_code():
    if hip.device_count() > 0:
        return
    try:
        ptamd.Context(0)
    except ptamd.PtError as e:
        assert e.code == ptamd.PT_E_NO_DEVICE
    else:
        raise AssertionError("pt_create succeeded without a device")


def test_tuning_knobs():
    for name in ("trace.blocks", "trace.refill"):
        v = ptamd.tune_get(name)
        ptamd.tune_set(name, v + 1)
        assert ptamd.tune_get(name) == v + 1
        ptamd.tune_set(name, v)
    assert ptamd._lib.pt_tune_set(b"no.such.knob", 1) == ptamd.PT_E_INVALID_ARG
    assert ptamd._lib.pt_tune_set(b"trace.blocks", 0) == ptamd.PT_E_INVALID_ARG
    for v in (8, 16, 32):
        ptamd.tune_set("trace.lds_stack", v)
        assert ptamd.tune_get("trace.lds_stack") == v
    for v in (0, 4, 12, 64):
        assert ptamd._lib.pt_tune_set(b"trace.lds_stack", v) == ptamd.PT_E_INVALID_ARG
    ptamd.tune_set("trace.lds_stack", 16)
    for v in (63, 4097, 0):
        assert ptamd._lib.pt_tune_set(b"trace.pool", v) == ptamd.PT_E_INVALID_ARG
    assert ptamd.tune_get("trace.pool") == 256
    chains = ptamd.tune_get("render.chains")
    for v in (1, 8):
        ptamd.tune_set("render.chains", v)
        assert ptamd.tune_get("render.chains") == v
    for v in (0, 9):
        assert ptamd._lib.pt_tune_set(b"render.chains", v) == ptamd.PT_E_INVALID_ARG
    ptamd.tune_set("render.chains", chains)
    for name in ("build.staged", "trace.filter", "render.paths"):
        v = ptamd.tune_get(name)
        assert v == 1
        for ok in (0, 1):
            ptamd.tune_set(name, ok)
            assert ptamd.tune_get(name) == ok
        assert ptamd._lib.pt_tune_set(name.encode(), 2) == ptamd.PT_E_INVALID_ARG
        ptamd.tune_set(name, v)


def test_null_arguments_rejected():
    assert ptamd._lib.pt_upload_scene(None, None, 0, None, 0, None, 0, None, None, None) == ptamd.PT_E_INVALID_ARG
    assert ptamd._lib.pt_scene_bounds(None, 0, None, None) == ptamd.PT_E_INVALID_ARG


def test_cpp_facade_header_compiles(tmp_path):
    """include/pt_tracer.hpp compiles warning-free on its own (C++14 and 17)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "t.cpp"
    src.write_text('#include "pt_tracer.hpp"\nint main() { ptamd::vec3 v{1, 2, 3}; return v.x > 0 ? 0 : 1; }\n')
    for std in ("c++14", "c++17"):
        r = subprocess.run(["g++", f"-std={std}", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                            "-I" + os.path.join(root, "include"), str(src)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
