"""ctypes binding of libfdtd_hip.so (include/fdtd_hip.h).

The HIP library is the only compute path of this package: importing this
module does not touch the GPU, but ``lib()`` raises ``NativeLibraryError``
when the shared library has not been built, and every call raises
``FdtdError`` with the library's message when it fails.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libfdtd_hip.so')

FDTD_MAX_PML = 16
MODE_MAIN, MODE_INCIDENT = 0, 1
DFT_COMPENSATED, DFT_REFERENCE = 0, 1

PHASOR_NAMES = [
    'ex_ph_d', 'ey_ph_d', 'hx_ph_d', 'hy_ph_d',
    'ex_ph_u', 'ey_ph_u', 'hx_ph_u', 'hy_ph_u',
    'ey_ph_l', 'ez_ph_l', 'hy_ph_l', 'hz_ph_l',
    'ey_ph_r', 'ez_ph_r', 'hy_ph_r', 'hz_ph_r',
    'ex_ph_f', 'ez_ph_f', 'hx_ph_f', 'hz_ph_f',
    'ex_ph_b', 'ez_ph_b', 'hx_ph_b', 'hz_ph_b',
]
INC_PHASOR_NAMES = ['ex_inc_phasor', 'ey_inc_phasor', 'hx_inc_phasor',
                    'hy_inc_phasor']
PSI_NAMES = [
    'psiExz_d', 'psiEyz_d', 'psiExz_u', 'psiEyz_u',
    'psiEyx_l', 'psiEzx_l', 'psiEyx_r', 'psiEzx_r',
    'psiExy_f', 'psiEzy_f', 'psiExy_b', 'psiEzy_b',
    'psiHxz_d', 'psiHyz_d', 'psiHxz_u', 'psiHyz_u',
    'psiHyx_l', 'psiHzx_l', 'psiHyx_r', 'psiHzx_r',
    'psiHxy_f', 'psiHzy_f', 'psiHxy_b', 'psiHzy_b',
]
FIELD_NAMES = ['ex', 'ey', 'ez', 'hx', 'hy', 'hz']

#: Every symbol include/fdtd_hip.h declares.
EXPORTS = [
    'fdtd_abi_version', 'fdtd_last_error', 'fdtd_init', 'fdtd_device_info',
    'fdtd_sim_footprint', 'fdtd_sim_create', 'fdtd_sim_destroy',
    'fdtd_sim_set_materials', 'fdtd_sim_set_uniform_material', 'fdtd_sim_build_materials', 'fdtd_sim_get_materials',
    'fdtd_sim_set_source', 'fdtd_sim_attach_beam', 'fdtd_sim_reset',
    'fdtd_sim_run', 'fdtd_sim_sync', 'fdtd_sim_step_count',
    'fdtd_sim_get_phasor', 'fdtd_sim_get_field', 'fdtd_sim_set_field',
    'fdtd_sim_get_psi', 'fdtd_sim_set_stepper', 'fdtd_sim_set_launch',
    'fdtd_sim_time_steps',
    'fdtd_sim_near2far', 'fdtd_sim_near2far_noalloc', 'fdtd_near2far_host',
    'fdtd_sim_near2far_grid', 'fdtd_near2far_host_grid',
    'fdtd_beam_create', 'fdtd_beam_upload', 'fdtd_beam_get', 'fdtd_beam_release',
    'fdtd_beam_synthesize',
    'fdtd_batch_create', 'fdtd_batch_destroy', 'fdtd_batch_set_launch',
    'fdtd_batch_run', 'fdtd_batch_sync', 'fdtd_batch_time_steps', 'fdtd_batch_set_mode', 'fdtd_batch_busy',
    'fdtd_probe_copy', 'fdtd_probe_streams', 'fdtd_probe_stencil',
]


class NativeLibraryError(RuntimeError):
    """libfdtd_hip.so is missing or cannot be loaded."""


class FdtdError(RuntimeError):
    """A libfdtd_hip call returned an error code."""

    def __init__(self, code, message):
        super().__init__(f'[{code}] {message}')
        self.code = code


class FdtdGeom(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ('nx', 'ny', 'nz', 'npml', 'tfsf_z', 'read_out', 'nt', 'nt_inc',
                 'mode', 'dft_accum')]


class FdtdCoeffs(ctypes.Structure):
    _fields_ = [('Db', ctypes.c_float), ('dtm0', ctypes.c_float),
                ('dx', ctypes.c_float), ('dte0dx', ctypes.c_float),
                ('omega', ctypes.c_double), ('dt', ctypes.c_double),
                ('cell_size', ctypes.c_double),
                ('prof', (ctypes.c_float * FDTD_MAX_PML) * 8)]


MAX_LAYERS = 4
LAYER_CONST, LAYER_RANDOM, LAYER_GRADIENT = 0, 1, 2


class FdtdMatLayer(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int32), ('seed', ctypes.c_uint32),
                ('eps', ctypes.c_double), ('sigma', ctypes.c_double),
                ('eps_std', ctypes.c_double), ('sigma_hi', ctypes.c_double),
                ('mv_bottom', ctypes.c_double), ('mv_top', ctypes.c_double),
                ('freq', ctypes.c_double), ('eps0', ctypes.c_double),
                ('f0', ctypes.c_double), ('f1', ctypes.c_double),
                ('poly', ((ctypes.c_double * 3) * 2) * 2)]


class FdtdMatCylinder(ctypes.Structure):
    _fields_ = [('c', ctypes.c_double * 3), ('u', ctypes.c_double * 3),
                ('radius', ctypes.c_double), ('length', ctypes.c_double),
                ('eps', ctypes.c_double), ('sigma', ctypes.c_double),
                ('lo', ctypes.c_int32 * 3), ('hi', ctypes.c_int32 * 3)]


class FdtdMatSpec(ctypes.Structure):
    _fields_ = [('nz_mat', ctypes.c_int32), ('nlayers', ctypes.c_int32),
                ('ncyl', ctypes.c_int32), ('vox_z1', ctypes.c_int32),
                ('vox_frac', ctypes.c_double), ('vox_eps', ctypes.c_double),
                ('vox_sigma', ctypes.c_double), ('vox_seed', ctypes.c_uint32),
                ('pad0', ctypes.c_uint32), ('dt', ctypes.c_float), ('dx', ctypes.c_float),
                ('eps0', ctypes.c_float), ('pad1', ctypes.c_float),
                ('layer', FdtdMatLayer * MAX_LAYERS)]


class FdtdBeamSpec(ctypes.Structure):
    _fields_ = [('polarization', ctypes.c_int32), ('cb_vacuum', ctypes.c_float),
                ('theta_deg', ctypes.c_double), ('phi_deg', ctypes.c_double),
                ('gauss_radius', ctypes.c_double),
                ('pulse_fft', ctypes.POINTER(ctypes.c_double)),
                ('all_k', ctypes.POINTER(ctypes.c_double))]


POLARIZATIONS = {'HH': 0, 'VV': 1, 'RHCP': 2, 'LHCP': 3}


class FdtdDeviceInfo(ctypes.Structure):
    _fields_ = [('device', ctypes.c_int32), ('cu_count', ctypes.c_int32),
                ('xcd_count', ctypes.c_int32), ('warp_size', ctypes.c_int32),
                ('total_mem', ctypes.c_int64), ('free_mem', ctypes.c_int64),
                ('arch', ctypes.c_char * 64), ('name', ctypes.c_char * 128)]


_lib = None
_lock = threading.Lock()
_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_FP = ctypes.POINTER(ctypes.c_float)
_DP = ctypes.POINTER(ctypes.c_double)

_SIGS = {
    'fdtd_abi_version': ([], _I),
    'fdtd_last_error': ([], ctypes.c_char_p),
    'fdtd_init': ([_I], _I),
    'fdtd_device_info': ([ctypes.POINTER(FdtdDeviceInfo)], _I),
    'fdtd_sim_footprint': ([ctypes.POINTER(FdtdGeom), ctypes.POINTER(ctypes.c_int64)], _I),
    'fdtd_sim_create': ([ctypes.POINTER(FdtdGeom), ctypes.POINTER(FdtdCoeffs),
                         ctypes.POINTER(_P)], _I),
    'fdtd_sim_destroy': ([_P], None),
    'fdtd_sim_set_materials': ([_P, _FP, _FP], _I),
    'fdtd_sim_get_materials': ([_P, _FP, _FP], _I),
    'fdtd_sim_build_materials': ([_P, ctypes.POINTER(FdtdMatSpec), ctypes.POINTER(ctypes.c_int32),
                                  ctypes.POINTER(FdtdMatCylinder)], _I),
    'fdtd_sim_set_uniform_material': ([_P, _F, _F], _I),
    'fdtd_sim_set_source': ([_P, _FP, _FP, _I], _I),
    'fdtd_sim_attach_beam': ([_P, _P], _I),
    'fdtd_sim_reset': ([_P], _I),
    'fdtd_sim_run': ([_P, _I], _I),
    'fdtd_sim_sync': ([_P], _I),
    'fdtd_sim_step_count': ([_P, ctypes.POINTER(_I)], _I),
    'fdtd_sim_get_phasor': ([_P, _I, _FP], _I),
    'fdtd_sim_get_field': ([_P, _I, _FP], _I),
    'fdtd_sim_set_field': ([_P, _I, _FP], _I),
    'fdtd_sim_get_psi': ([_P, _I, _FP], _I),
    'fdtd_sim_set_stepper': ([_P, _I], _I),
    'fdtd_sim_set_launch': ([_P, _I, _I, _I, _I, _I], _I),
    'fdtd_sim_time_steps': ([_P, _I, _FP, _FP, _FP, _FP], _I),
    'fdtd_sim_near2far': ([_P, _DP, _I, _D, _D, _D, _DP], _I),
    'fdtd_sim_near2far_noalloc': ([_P, _DP, _I, _D, _D, _D, _DP], _I),
    'fdtd_near2far_host': ([_FP, _FP, _FP, _FP, _I, _I, _I, _D, _DP, _I, _D,
                            _D, _D, _DP], _I),
    'fdtd_sim_near2far_grid': ([_P, _DP, _I, _DP, _I, _D, _D, _DP], _I),
    'fdtd_near2far_host_grid': ([_FP, _FP, _FP, _FP, _I, _I, _I, _D, _DP, _I, _DP, _I,
                                 _D, _D, _DP], _I),
    'fdtd_beam_create': ([_I, _I, _I, _FP, _FP, _FP, _FP, ctypes.POINTER(_P)], _I),
    'fdtd_beam_upload': ([_P, _I, _I, _FP, _FP, _FP, _FP], _I),
    'fdtd_beam_get': ([_P, _I, _FP], _I),
    'fdtd_beam_release': ([_P], None),
    'fdtd_beam_synthesize': ([ctypes.POINTER(FdtdGeom), ctypes.POINTER(FdtdCoeffs),
                              ctypes.POINTER(FdtdBeamSpec), ctypes.POINTER(_P), _FP], _I),
    'fdtd_batch_busy': ([_P], _I),
    'fdtd_batch_set_mode': ([_P, _I, ctypes.c_int64, _I], _I),
    'fdtd_batch_create': ([ctypes.POINTER(_P), _I, ctypes.POINTER(_P)], _I),
    'fdtd_batch_destroy': ([_P], None),
    'fdtd_batch_set_launch': ([_P, _I, _I, _I, _I, _I], _I),
    'fdtd_batch_run': ([_P, _I, _I], _I),
    'fdtd_batch_sync': ([_P], _I),
    'fdtd_batch_time_steps': ([_P, _I, _I, _FP], _I),
    'fdtd_probe_copy': ([ctypes.c_int64, _I, _I, _FP], _I),
    'fdtd_probe_streams': ([ctypes.c_int64, _I, _I, _I, _FP], _I),
    'fdtd_probe_stencil': ([_I, _I, _I, _I, _I, _I, _FP], _I),
}


def lib():
    """Load libfdtd_hip.so (once).  Raises NativeLibraryError if absent."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise NativeLibraryError(
                    f'{LIB_PATH} is not built; run '
                    '`python -m sar_pixel_simulator_amd.build_native`')
            try:
                L = ctypes.CDLL(LIB_PATH)
            except OSError as e:
                raise NativeLibraryError(f'cannot load {LIB_PATH}: {e}') from e
            for name, (args, res) in _SIGS.items():
                f = getattr(L, name)
                f.argtypes = args
                f.restype = res
            _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().fdtd_last_error()
        raise FdtdError(rc, msg.decode() if msg else 'unknown error')


def _fptr(a):
    return a.ctypes.data_as(_FP)


def _dptr(a):
    return a.ctypes.data_as(_DP)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != tuple(shape):
        raise ValueError(f'expected shape {tuple(shape)}, got {a.shape}')
    return a


_inited = set()


def init(device=0):
    """Select the HIP device for handles created on this thread."""
    _check(lib().fdtd_init(int(device)))
    _inited.add(int(device))


def device_info():
    info = FdtdDeviceInfo()
    _check(lib().fdtd_device_info(ctypes.byref(info)))
    return {'device': info.device, 'cu_count': info.cu_count,
            'xcd_count': info.xcd_count, 'warp_size': info.warp_size,
            'total_mem': info.total_mem, 'free_mem': info.free_mem,
            'arch': info.arch.decode(), 'name': info.name.decode()}


def probe_copy(nbytes, vec=1, reps=5):
    """Average HIP-event time (ms) of one plain device copy of nbytes with
    vec floats per thread; returns (ms, GB/s counting read + write)."""
    ms = ctypes.c_float()
    _check(lib().fdtd_probe_copy(int(nbytes), int(vec), int(reps), ctypes.byref(ms)))
    return ms.value, 2 * nbytes / (ms.value * 1e-3) / 1e9


def probe_streams(n, nread, nwrite, reps=5):
    """Average time (ms) of one pass of the multi-stream probe over n floats
    per stream; returns (ms, GB/s over all streams)."""
    ms = ctypes.c_float()
    _check(lib().fdtd_probe_streams(int(n), int(nread), int(nwrite), int(reps), ctypes.byref(ms)))
    return ms.value, (nread + nwrite) * 4 * n / (ms.value * 1e-3) / 1e9


def probe_stencil(nx, ny, nz, zchunk, kind, reps=5):
    """Average time (ms) of one pass of the stencil-pattern probe (kind 0:
    split H half, 9 streams per cell; 1: split E half, 11; 2: the halo
    stepper's fused step, 14); returns (ms, GB/s over the algorithmic
    streams)."""
    ms = ctypes.c_float()
    _check(lib().fdtd_probe_stencil(int(nx), int(ny), int(nz), int(zchunk), int(kind), int(reps),
                                    ctypes.byref(ms)))
    streams = {0: 9, 1: 11, 2: 14}[int(kind)]
    return ms.value, streams * 4 * nx * ny * nz / (ms.value * 1e-3) / 1e9


def make_geom(nx, ny, nz, npml, tfsf_z, nt, nt_inc, mode=MODE_MAIN,
              read_out=0, dft_accum=DFT_COMPENSATED):
    g = FdtdGeom()
    g.nx, g.ny, g.nz, g.npml = nx, ny, nz, npml
    g.tfsf_z, g.read_out, g.nt, g.nt_inc = tfsf_z, read_out, nt, nt_inc
    g.mode, g.dft_accum = mode, dft_accum
    return g


def make_coeffs(Db, dtm0, dx, dte0dx, omega, dt, profiles, cell_size=None):
    """``profiles``: the eight 1-D CPML profiles (be_dlf, ce_dlf, bh_dlf,
    ch_dlf, be_urb, ce_urb, bh_urb, ch_urb), each npml float32 values."""
    c = FdtdCoeffs()
    c.Db, c.dtm0, c.dx, c.dte0dx = (float(np.float32(v))
                                    for v in (Db, dtm0, dx, dte0dx))
    c.omega, c.dt = float(omega), float(dt)
    c.cell_size = float(dx if cell_size is None else cell_size)
    for k, prof in enumerate(profiles):
        prof = np.asarray(prof, dtype=np.float32)
        if prof.size > FDTD_MAX_PML:
            raise ValueError(f'CPML profile longer than {FDTD_MAX_PML}')
        for i, v in enumerate(prof):
            c.prof[k][i] = float(v)
    return c


def footprint(geom):
    out = ctypes.c_int64()
    _check(lib().fdtd_sim_footprint(ctypes.byref(geom), ctypes.byref(out)))
    return out.value


class Beam:
    """Incident beam (ex, ey, hx, hy) x (nt, ny, nx) float32 resident in HBM."""

    def __init__(self, nt, ny, nx, ex=None, ey=None, hx=None, hy=None):
        self.nt, self.ny, self.nx = nt, ny, nx
        h = _P()
        if ex is None:
            _check(lib().fdtd_beam_create(nt, ny, nx, None, None, None, None,
                                          ctypes.byref(h)))
        else:
            shp = (nt, ny, nx)
            arrs = [_f32(a, shp) for a in (ex, ey, hx, hy)]
            _check(lib().fdtd_beam_create(nt, ny, nx, *[_fptr(a) for a in arrs],
                                          ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def upload(self, t0, ex, ey, hx, hy):
        n = ex.shape[0]
        shp = (n, self.ny, self.nx)
        arrs = [_f32(a, shp) for a in (ex, ey, hx, hy)]
        _check(lib().fdtd_beam_upload(self._h, int(t0), n, *[_fptr(a) for a in arrs]))

    def component(self, comp):
        out = np.empty((self.nt, self.ny, self.nx), np.float32)
        _check(lib().fdtd_beam_get(self._h, int(comp), _fptr(out)))
        return out

    def release(self):
        if self._h:
            lib().fdtd_beam_release(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def beam_synthesize(geom, coeffs, polarization, theta_deg, phi_deg, gauss_radius,
                    pulse_fft, all_k, cb_vacuum):
    """Device beam pipeline (fdtd_beam_synthesize): returns (Beam, incident
    phasors dict).  ``geom`` is the INCIDENT-mode geometry."""
    pf = np.ascontiguousarray(np.asarray(pulse_fft, np.complex128))
    ak = np.ascontiguousarray(all_k, dtype=np.float64)
    if pf.size != geom.nt or ak.size != geom.nt:
        raise ValueError('pulse spectrum and wavenumbers need geom.nt entries')
    spec = FdtdBeamSpec()
    spec.polarization = POLARIZATIONS[polarization]
    spec.cb_vacuum = float(np.float32(cb_vacuum))
    spec.theta_deg, spec.phi_deg = float(theta_deg), float(phi_deg)
    spec.gauss_radius = float(gauss_radius)
    spec.pulse_fft = pf.view(np.float64).ctypes.data_as(_DP)
    spec.all_k = _dptr(ak)
    nyp, nxp = geom.ny - 2 * geom.npml, geom.nx - 2 * geom.npml
    ph = np.empty((4, nyp, nxp), np.complex64)
    h = _P()
    _check(lib().fdtd_beam_synthesize(ctypes.byref(geom), ctypes.byref(coeffs), ctypes.byref(spec),
                                      ctypes.byref(h), ph.view(np.float32).ctypes.data_as(_FP)))
    beam = Beam.__new__(Beam)
    beam.nt, beam.ny, beam.nx = geom.nt, geom.ny, geom.nx
    beam._h = h
    return beam, {n: ph[i] for i, n in enumerate(INC_PHASOR_NAMES)}


class Sim:
    """One simulation's device state (fields, CPML psi, phasor sums)."""

    def __init__(self, geom, coeffs):
        self.geom = geom
        self.coeffs = coeffs
        h = _P()
        _check(lib().fdtd_sim_create(ctypes.byref(geom), ctypes.byref(coeffs),
                                     ctypes.byref(h)))
        self._h = h
        self._beam = None
        nx, ny, nz, np_ = geom.nx, geom.ny, geom.nz, geom.npml
        self.shape = (nz, ny, nx)
        nxp, nyp, nzp = nx - 2 * np_, ny - 2 * np_, nz - 2 * np_
        face = {'d': (nyp, nxp), 'u': (nyp, nxp), 'l': (nzp, nyp),
                'r': (nzp, nyp), 'f': (nzp, nxp), 'b': (nzp, nxp)}
        self.phasor_shapes = [face[n[-1]] for n in PHASOR_NAMES]
        psi = {'d': (np_, ny, nx), 'u': (np_, ny, nx), 'l': (nz, ny, np_),
               'r': (nz, ny, np_), 'f': (nz, np_, nx), 'b': (nz, np_, nx)}
        self.psi_shapes = [psi[n[-1]] for n in PSI_NAMES]

    @property
    def handle(self):
        return self._h

    def set_materials(self, Ca, Cb):
        Ca, Cb = _f32(Ca, self.shape), _f32(Cb, self.shape)
        _check(lib().fdtd_sim_set_materials(self._h, _fptr(Ca), _fptr(Cb)))

    def build_materials(self, spec, surfaces=None, cylinders=None):
        """Form Ca, Cb on the device (fdtd_sim_build_materials) from an
        FdtdMatSpec, its int32 (nlayers, ny, nx) surface maps and an array
        of FdtdMatCylinder."""
        surf_p = ctypes.POINTER(ctypes.c_int32)()
        if spec.nlayers > 0:
            surfaces = np.ascontiguousarray(surfaces, dtype=np.int32)
            if surfaces.shape != (spec.nlayers,) + self.shape[1:]:
                raise ValueError(f'surfaces of shape {surfaces.shape}, expected '
                                 f'{(spec.nlayers,) + self.shape[1:]}')
            surf_p = surfaces.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        cyl_p = ctypes.POINTER(FdtdMatCylinder)()
        if spec.ncyl > 0:
            if cylinders is None or len(cylinders) != spec.ncyl:
                raise ValueError('cylinder list does not match spec.ncyl')
            cyl_p = cylinders
        _check(lib().fdtd_sim_build_materials(self._h, ctypes.byref(spec), surf_p, cyl_p))

    def materials(self):
        """The device's (Ca, Cb), each (nz, ny, nx) float32."""
        Ca = np.empty(self.shape, np.float32)
        Cb = np.empty(self.shape, np.float32)
        _check(lib().fdtd_sim_get_materials(self._h, _fptr(Ca), _fptr(Cb)))
        return Ca, Cb

    def set_uniform_material(self, ca, cb):
        _check(lib().fdtd_sim_set_uniform_material(self._h, float(ca), float(cb)))

    def set_source(self, src_ex, src_ey):
        n = src_ex.shape[0]
        shp = (n,) + self.shape[1:]
        a, b = _f32(src_ex, shp), _f32(src_ey, shp)
        _check(lib().fdtd_sim_set_source(self._h, _fptr(a), _fptr(b), n))

    def attach_beam(self, beam):
        _check(lib().fdtd_sim_attach_beam(self._h, beam.handle))
        self._beam = beam

    def reset(self):
        _check(lib().fdtd_sim_reset(self._h))

    def run(self, nsteps):
        _check(lib().fdtd_sim_run(self._h, int(nsteps)))

    def sync(self):
        _check(lib().fdtd_sim_sync(self._h))

    def step_count(self):
        t = ctypes.c_int()
        _check(lib().fdtd_sim_step_count(self._h, ctypes.byref(t)))
        return t.value

    def phasor(self, index):
        shp = self.phasor_shapes[index] if self.geom.mode == MODE_MAIN \
            else self.phasor_shapes[0]
        out = np.empty(shp, np.complex64)
        _check(lib().fdtd_sim_get_phasor(self._h, int(index),
                                         out.view(np.float32).ctypes.data_as(_FP)))
        return out

    def phasors(self):
        if self.geom.mode == MODE_MAIN:
            return {n: self.phasor(i) for i, n in enumerate(PHASOR_NAMES)}
        return {n: self.phasor(i) for i, n in enumerate(INC_PHASOR_NAMES)}

    def field(self, comp):
        comp = FIELD_NAMES.index(comp) if isinstance(comp, str) else comp
        out = np.empty(self.shape, np.float32)
        _check(lib().fdtd_sim_get_field(self._h, int(comp), _fptr(out)))
        return out

    def set_field(self, comp, values):
        comp = FIELD_NAMES.index(comp) if isinstance(comp, str) else comp
        v = _f32(values, self.shape)
        _check(lib().fdtd_sim_set_field(self._h, int(comp), _fptr(v)))

    def psi(self, index):
        index = PSI_NAMES.index(index) if isinstance(index, str) else index
        out = np.empty(self.psi_shapes[index], np.float32)
        _check(lib().fdtd_sim_get_psi(self._h, int(index), _fptr(out)))
        return out

    def set_launch(self, ty, zchunk=2, xvec=0, xwaves=0, xcd_order=0):
        """Tile shape (fdtd_sim_set_launch).  The defaults are the library's
        own: 2-plane z chunks, lane width and block width chosen from the
        grid (0), plain block order; ty 0 chooses the rows per block."""
        _check(lib().fdtd_sim_set_launch(self._h, int(ty), int(zchunk), int(xvec), int(xwaves),
                                         int(xcd_order)))

    def set_stepper(self, fused):
        """0 / False: split H and E kernels; 1 / True: fused H+E kernel per
        step on row strips; 2: fused H+E kernel with one row per
        wavefront; 3: fused H+E kernel with one row per wavefront that
        recomputes its neighbours' H (the split stepper's tiles)."""
        _check(lib().fdtd_sim_set_stepper(self._h, int(fused)))
        self.stepper = int(fused)

    def time_steps(self, nsteps, per_kernel=True):
        """Run ``nsteps`` and return HIP-event times (ms) measured on the
        library's stream: total, and per-kernel averages when asked."""
        tot, h, e, d = (ctypes.c_float() for _ in range(4))
        if per_kernel:
            _check(lib().fdtd_sim_time_steps(self._h, int(nsteps), ctypes.byref(tot),
                                             ctypes.byref(h), ctypes.byref(e),
                                             ctypes.byref(d)))
            return {'total': tot.value, 'h': h.value, 'e': e.value, 'dft': d.value}
        _check(lib().fdtd_sim_time_steps(self._h, int(nsteps), ctypes.byref(tot),
                                         None, None, None))
        return {'total': tot.value}

    def near2far(self, thetas_deg, phi_deg, r, freq, noalloc=False):
        """Far field of the top-face phasors.  ``noalloc`` uses the entry
        point that allocates no device memory (at most N2F_MAX_THETA angles
        and N2F_MAX_SIDE phasor samples per side; bit-identical results), so
        the call does not wait for other streams' work as a device free
        does."""
        th = np.ascontiguousarray(thetas_deg, dtype=np.float64)
        out = np.empty((4, th.size), np.complex128)
        fn = lib().fdtd_sim_near2far_noalloc if noalloc else lib().fdtd_sim_near2far
        _check(fn(self._h, _dptr(th), th.size, float(phi_deg), float(r), float(freq),
                  out.view(np.float64).ctypes.data_as(_DP)))
        return {'E_theta': out[0], 'E_phi': out[1], 'H_theta': out[2],
                'H_phi': out[3]}

    def near2far_grid(self, thetas_deg, phis_deg, r, freq):
        """Far field over the theta x phi grid in one launch: each component
        (nphi, ntheta) complex128."""
        th = np.ascontiguousarray(thetas_deg, dtype=np.float64)
        ph = np.ascontiguousarray(phis_deg, dtype=np.float64)
        out = np.empty((4, ph.size, th.size), np.complex128)
        _check(lib().fdtd_sim_near2far_grid(self._h, _dptr(th), th.size, _dptr(ph), ph.size,
                                            float(r), float(freq),
                                            out.view(np.float64).ctypes.data_as(_DP)))
        return {'E_theta': out[0], 'E_phi': out[1], 'H_theta': out[2], 'H_phi': out[3]}

    def destroy(self):
        if self._h:
            lib().fdtd_sim_destroy(self._h)
            self._h = _P()
            self._beam = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def near2far_host(ex_u, ey_u, hx_u, hy_u, nx_ph, ny_ph, nz_ph, dx, thetas_deg,
                  phi_deg, r, freq):
    """Far field of caller-supplied top-face phasors (complex, (ny_ph, nx_ph))."""
    arrs = [np.ascontiguousarray(a, dtype=np.complex64) for a in (ex_u, ey_u, hx_u, hy_u)]
    for a in arrs:
        if a.shape != (ny_ph, nx_ph):
            raise ValueError(f'phasor plane shape {a.shape} != {(ny_ph, nx_ph)}')
    th = np.ascontiguousarray(thetas_deg, dtype=np.float64)
    out = np.empty((4, th.size), np.complex128)
    _check(lib().fdtd_near2far_host(
        *[a.view(np.float32).ctypes.data_as(_FP) for a in arrs],
        int(nx_ph), int(ny_ph), int(nz_ph), float(dx), _dptr(th), th.size,
        float(phi_deg), float(r), float(freq),
        out.view(np.float64).ctypes.data_as(_DP)))
    return {'E_theta': out[0], 'E_phi': out[1], 'H_theta': out[2], 'H_phi': out[3]}


BATCH_PACKED, BATCH_SERIAL, BATCH_HYBRID = 0, 1, 2
N2F_MAX_THETA, N2F_MAX_SIDE = 181, 2047


def near2far_host_grid(ex_u, ey_u, hx_u, hy_u, nx_ph, ny_ph, nz_ph, dx, thetas_deg,
                       phis_deg, r, freq):
    """Far field of caller-supplied top-face phasors over the theta x phi
    grid in one launch: each component (nphi, ntheta) complex128."""
    arrs = [np.ascontiguousarray(a, dtype=np.complex64) for a in (ex_u, ey_u, hx_u, hy_u)]
    for a in arrs:
        if a.shape != (ny_ph, nx_ph):
            raise ValueError(f'phasor plane shape {a.shape} != {(ny_ph, nx_ph)}')
    th = np.ascontiguousarray(thetas_deg, dtype=np.float64)
    ph = np.ascontiguousarray(phis_deg, dtype=np.float64)
    out = np.empty((4, ph.size, th.size), np.complex128)
    _check(lib().fdtd_near2far_host_grid(
        *[a.view(np.float32).ctypes.data_as(_FP) for a in arrs],
        int(nx_ph), int(ny_ph), int(nz_ph), float(dx), _dptr(th), th.size, _dptr(ph), ph.size,
        float(r), float(freq), out.view(np.float64).ctypes.data_as(_DP)))
    return {'E_theta': out[0], 'E_phi': out[1], 'H_theta': out[2], 'H_phi': out[3]}


class Batch:
    """Several simulations stepped by one launch per kernel per step."""

    def __init__(self, sims):
        self.sims = list(sims)
        arr = (_P * len(self.sims))(*[s.handle for s in self.sims])
        h = _P()
        _check(lib().fdtd_batch_create(arr, len(self.sims), ctypes.byref(h)))
        self._h = h

    def set_launch(self, ty, zchunk=2, xvec=0, xwaves=0, xcd_order=0):
        """Tile shape of the batch (fdtd_batch_set_launch), defaults as
        Sim.set_launch; ty, xvec and xwaves all 0 give each member its own
        block width."""
        _check(lib().fdtd_batch_set_launch(self._h, int(ty), int(zchunk), int(xvec), int(xwaves),
                                           int(xcd_order)))

    def set_mode(self, mode, small_cells=0, lanes=1):
        """BATCH_PACKED: every member in one tile list; BATCH_SERIAL: each
        member's own launches, one member after another on each of ``lanes``
        concurrent streams; BATCH_HYBRID: members below ``small_cells`` cells
        packed, the others serial."""
        _check(lib().fdtd_batch_set_mode(self._h, int(mode), int(small_cells), int(lanes)))

    def run(self, nsteps, use_graph=True):
        _check(lib().fdtd_batch_run(self._h, int(nsteps), 1 if use_graph else 0))

    def sync(self):
        _check(lib().fdtd_batch_sync(self._h))

    def busy(self):
        """True while queued steps are still running (does not wait)."""
        rc = lib().fdtd_batch_busy(self._h)
        if rc < 0:
            _check(rc)
        return rc == 1

    def time_steps(self, nsteps, use_graph=True):
        tot = ctypes.c_float()
        _check(lib().fdtd_batch_time_steps(self._h, int(nsteps), 1 if use_graph else 0,
                                           ctypes.byref(tot)))
        return tot.value

    def destroy(self):
        if self._h:
            lib().fdtd_batch_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
