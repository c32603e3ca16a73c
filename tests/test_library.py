"""CPU checks of the native boundary: the C-ABI library loads and exports
every entry point include/sc_lbm_amd.h declares (no compute calls)."""
import ctypes
import os
import re

from sc_lbm_amd import native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "sc_lbm_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(scl_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert header_symbols() == sorted(nat.EXPORTS)


def test_library_loads_and_exports_every_symbol():
    lib = nat.load()
    for name in header_symbols():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


def test_errors_are_codes_not_exceptions():
    lib = nat.load()
    h = ctypes.c_void_p()
    # bad sizes are rejected before any device call
    assert lib.scl_geometry_create(0, 5, None, ctypes.byref(h)) == nat.SCL_ERR_ARG
    assert b"bad geometry size" in lib.scl_last_error()
    assert lib.scl_set_tile(None, 64, 4) == nat.SCL_ERR_ARG


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N with N > 1 exits with a message
    (checked before any device is touched)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stdout
