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


# Geometry validation at the boundary (fdtd_sim_footprint runs the same
# checks as fdtd_sim_create, on the host, with no device call).

def _geom(nx=33, ny=31, nz=35, npml=10, tfsf_z=20, read_out=0, nt=60, nt_inc=60,
          mode=0, dft_accum=0):
    return native.FdtdGeom(nx, ny, nz, npml, tfsf_z, read_out, nt, nt_inc, mode, dft_accum)


def test_footprint_counts_the_resident_state():
    g = _geom()
    b = native.footprint(g)
    pitch = 48                                   # 33 cells rounded up to 16 floats
    ncell = 35 * 31 * pitch
    # two buffer sets of six fields, Ca and Cb: a lower bound in closed form
    assert b >= (2 * 6 + 2) * ncell * 4
    gi = _geom(mode=1, read_out=17)
    assert native.footprint(gi) < b              # INCIDENT: one set, no Ca/Cb
    big = native.footprint(_geom(nx=761, ny=761, nz=203, tfsf_z=156, nt=1714, nt_inc=1372))
    ncell = 203 * 761 * 768
    # two sets of six fields, Ca and Cb (56 B per cell), plus CPML state,
    # phasor sums and the x-face copies (under 10 % more here)
    assert 56 * ncell < big < 1.1 * 56 * ncell


@pytest.mark.parametrize('kw,match', [
    (dict(nx=21), 'too small for npml'),                  # 2 npml + 2 = 22 cells at least
    (dict(npml=1), 'npml'),
    (dict(npml=17, nx=40, ny=40, nz=40), 'npml'),
    (dict(nx=2048, ny=2048, nz=256, tfsf_z=100), 'limit is 2\\^30'),
    (dict(tfsf_z=35), 'tfsf_z'),
    (dict(tfsf_z=-1), 'tfsf_z'),
    (dict(mode=1, read_out=0), 'read-out'),
    (dict(nt=-1), 'negative'),
    (dict(mode=2), 'mode'),
    (dict(dft_accum=5), 'dft_accum'),
])
def test_geometry_rejected_with_reason(kw, match):
    with pytest.raises(native.FdtdError, match=match):
        native.footprint(_geom(**kw))


@pytest.mark.parametrize('kw', [dict(nx=22, ny=22, nz=22, tfsf_z=11),       # smallest interior
                                dict(nx=2048, ny=2048, nz=255, tfsf_z=100),  # largest nz under 2^30
                                dict(npml=16, nx=34, ny=34, nz=34, tfsf_z=20)])
def test_geometry_limits_accepted(kw):
    assert native.footprint(_geom(**kw)) > 0
