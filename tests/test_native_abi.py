"""The C-ABI library loads and exports every entry point include/fdtd_hip.h
declares (no device calls)."""
import os
import re

import pytest

from sar_pixel_simulator_amd import _native as native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, 'include', 'fdtd_hip.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(fdtd_[a-z0-9_]+)\s*\(', src)))


def test_header_lists_every_export():
    assert declared() == sorted(native.EXPORTS)


def test_library_exports_declared_symbols():
    L = native.lib()
    missing = [s for s in declared() if not hasattr(L, s)]
    assert not missing


def test_abi_version_and_error_string():
    L = native.lib()
    src = open(os.path.join(ROOT, 'include', 'fdtd_hip.h')).read()
    want = int(re.search(r'#define FDTD_ABI_VERSION (\d+)', src).group(1))
    assert L.fdtd_abi_version() == want
    assert isinstance(L.fdtd_last_error(), bytes)


def test_struct_layouts_match_header():
    import ctypes
    assert ctypes.sizeof(native.FdtdGeom) == 10 * 4
    assert ctypes.sizeof(native.FdtdCoeffs) == 4 * 4 + 3 * 8 + 8 * 16 * 4
    assert ctypes.sizeof(native.FdtdDeviceInfo) == 4 * 4 + 2 * 8 + 64 + 128


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(native, 'LIB_PATH', str(tmp_path / 'nope.so'))
    monkeypatch.setattr(native, '_lib', None)
    with pytest.raises(native.NativeLibraryError):
        native.lib()


def test_python_constants_match_header():
    src = open(os.path.join(ROOT, 'include', 'fdtd_hip.h')).read()

    def define(name):
        return int(re.search(rf'#define {name} (\d+)', src).group(1))
    assert native.N2F_MAX_THETA == define('FDTD_N2F_MAX_THETA')
    assert native.N2F_MAX_SIDE == define('FDTD_N2F_MAX_SIDE')
    assert (native.BATCH_PACKED, native.BATCH_SERIAL, native.BATCH_HYBRID) == \
        (define('FDTD_BATCH_PACKED'), define('FDTD_BATCH_SERIAL'), define('FDTD_BATCH_HYBRID'))
