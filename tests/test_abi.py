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


def test_cpp_adapter_header_compiles_and_links(tmp_path):
    """The C++ adapter compiles standalone (no Eigen, no torch) and the
    native driver links against the library."""
    import subprocess
    src = tmp_path / "use.cpp"
    src.write_text('#include "vps_virtual_particle.hpp"\n'
                   'int main() { vps_params p{}; (void)p;\n'
                   '  using VP3 = vps::VirtualParticleGpu<3>; using VP2 = vps::VirtualParticleGpu<2>;\n'
                   '  return sizeof(VP3) > 0 && sizeof(VP2) > 0 ? 0 : 1; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    exe = os.path.join(ROOT, "sand-water-coupling_amd", "lib", "vps_adapter_check")
    assert os.path.exists(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
