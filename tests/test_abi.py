"""The C-ABI library loads and exports every symbol include/vps.h declares,
with the struct layouts the bindings assume (no GPU calls)."""
import ctypes
import os
import re

from vps import _abi, core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "vps.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vps_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = core.load_library()
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) >= set(_abi.EXPORTED)


def test_struct_layouts_and_version():
    lib = core.load_library()
    assert lib.vps_abi_version() == _abi.ABI_VERSION
    assert lib.vps_sizeof_params() == ctypes.sizeof(_abi.Params)
    assert lib.vps_sizeof_stats() == ctypes.sizeof(_abi.Stats)


def test_create_rejects_bad_params_without_gpu():
    # argument checks run before any device call
    lib = core.load_library()
    p = _abi.Params()
    p.dim = 4
    h = ctypes.c_void_p()
    assert lib.vps_create(ctypes.byref(p), ctypes.byref(h)) == _abi.VPS_ERR_INVALID_ARG
    assert b"dim" in lib.vps_last_error(None)
    assert lib.vps_create(None, ctypes.byref(h)) == _abi.VPS_ERR_INVALID_ARG
