"""CPU-side checks of the C ABI: the library loads and exports every entry
point include/pyticles_hip.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

from pyticles_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "pyticles_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_bound_symbols():
    decl = _declared()
    assert len(decl) >= 30
    assert sorted(_lib.EXPORTED) == decl


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_version_and_error_strings_without_gpu():
    assert b"gfx950" in _lib.lib.sph_version()
    assert isinstance(_lib.lib.sph_last_error(), bytes)


def test_params_struct_layout():
    # 14 doubles + unsigned + int, no padding surprises
    assert ctypes.sizeof(_lib.Params) == 14 * 8 + 8


def test_library_is_gfx950_code_object():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_reference_c_level_lucy_kernel_exported():
    # c_properties.pxd:1: void lucy_kernel_3d(float r, float dx[3], float h, float* w, float dwdx[3])
    lib = ctypes.CDLL(_lib.LIB_PATH)
    assert hasattr(lib, "lucy_kernel_3d")
    src = open(os.path.join(ROOT, "include", "pyticles_hip.h")).read()
    assert "void lucy_kernel_3d(float r, const float *dx, float h, float *w, float *dwdx);" in src
