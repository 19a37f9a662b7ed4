"""The C-ABI library: built for gfx950, loadable without a GPU, and exporting
every entry point include/xdem.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xdem.h")
LIB = os.path.join(ROOT, "xdem_amd", "libxdem_hip.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(xdem_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("xdem_create", "xdem_set_particles", "xdem_step", "xdem_get_stress"):
        assert s in syms
    from xdem_amd.engine import SYMBOLS
    assert sorted(SYMBOLS) == syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libxdem_hip.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    """The embedded offload bundle targets gfx950 (our own build)."""
    data = open(LIB, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_engine_rejects_bad_config_without_gpu():
    """Argument validation happens before any device call."""
    from xdem_amd.engine import XdemConfig, load_library
    L = load_library()
    cfg = (XdemConfig * 1)()
    cfg[0].xmin, cfg[0].xmax, cfg[0].ymin, cfg[0].ymax, cfg[0].dt = 0.0, -1.0, 0.0, 1.0, 1e-4
    h = ctypes.c_void_p()
    assert L.xdem_create(ctypes.byref(h), 1, cfg) == -1


def test_cpp_adapter_compiles():
    """The reference-style C++ adapter (xdem_world.hpp) builds and links
    against the C-ABI library with plain g++ (host code only)."""
    import subprocess
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp"), "-B"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(os.path.join(ROOT, "tests", "cpp", "world_demo"))
