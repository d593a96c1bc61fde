"""The C-ABI library loads and exports every entry point include/mray_amd.h declares."""
import ctypes
import os
import re

from mray_amd import abi, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mray_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mray_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    L = native.lib()
    decl = declared_symbols()
    assert len(decl) >= 20
    missing = [s for s in decl if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(native.EXPORTS) == decl


def test_struct_layouts():
    assert ctypes.sizeof(abi.MRayMaterial) == 52
    assert ctypes.sizeof(abi.MRayCamera) == 52
    assert ctypes.sizeof(abi.MRayRenderParams) == 48
    assert ctypes.sizeof(abi.MRayRenderStats) == 80
    assert ctypes.sizeof(abi.MRayNrcConfig) == 48
    assert ctypes.sizeof(abi.MRayBatchView) == 76
    assert ctypes.sizeof(abi.MRayBatchTune) == 24
    assert ctypes.sizeof(abi.MRayNrcRenderParams) == 24


def test_error_path_without_gpu_work():
    """Invalid arguments are rejected before any device call."""
    L = native.lib()
    assert L.mray_accel_create(None, None) == 1
    assert b"NULL" in L.mray_last_error()
    assert L.mray_renderer_iterate(None, 1, None) == 1
    assert L.mray_radix_sort_temp_bytes(10000) > 10000 * 12
