"""ORACLE — test infrastructure only.

CPU restatement of the reference's hot path.  Only tests/, the smoke
check and bench.py's cpu_baseline leg use it, to compare the HIP library's
results against (and, in bench.py, to time the CPU path).  The product
never imports this module.

* ``OracleSim`` drives the C loops in fdtd_oracle_impl.h (FDTD time loop of
  fdtdsarsim/cpu/main_sim.py:85-195 and cpu/incident_sim.py:67-171).
* ``far_field`` restates FarFieldTransform (fdtdsarsim/near2far.py:28-156)
  with scipy.integrate.simpson in FP64.
* ``farfield_power``, ``incident_power`` and ``sigma0`` restate
  processresults.py:260-271, :144-155 and :104-121.
"""
import ctypes
import os
import subprocess

import numpy as np
import scipy.integrate

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, '_build', 'liboracle.so')

PHASOR_NAMES = [
    'ex_ph_d', 'ey_ph_d', 'hx_ph_d', 'hy_ph_d',
    'ex_ph_u', 'ey_ph_u', 'hx_ph_u', 'hy_ph_u',
    'ey_ph_l', 'ez_ph_l', 'hy_ph_l', 'hz_ph_l',
    'ey_ph_r', 'ez_ph_r', 'hy_ph_r', 'hz_ph_r',
    'ex_ph_f', 'ez_ph_f', 'hx_ph_f', 'hz_ph_f',
    'ex_ph_b', 'ez_ph_b', 'hx_ph_b', 'hz_ph_b',
]
PSI_NAMES = [
    'psiExz_d', 'psiEyz_d', 'psiExz_u', 'psiEyz_u',
    'psiEyx_l', 'psiEzx_l', 'psiEyx_r', 'psiEzx_r',
    'psiExy_f', 'psiEzy_f', 'psiExy_b', 'psiEzy_b',
    'psiHxz_d', 'psiHyz_d', 'psiHxz_u', 'psiHyz_u',
    'psiHyx_l', 'psiHzx_l', 'psiHyx_r', 'psiHzx_r',
    'psiHxy_f', 'psiHzy_f', 'psiHxy_b', 'psiHzy_b',
]
FIELD_NAMES = ['ex', 'ey', 'ez', 'hx', 'hy', 'hz']


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(['make', '-s', '-C', HERE], check=True)


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        _LIB = ctypes.CDLL(LIB_PATH)
    return _LIB


def _structs(ct):
    P = ctypes.POINTER(ct)

    class State(ctypes.Structure):
        _fields_ = ([(n, P) for n in FIELD_NAMES]
                    + [(n, P) for n in PSI_NAMES]
                    + [('ph', P * 24)])

    class Geom(ctypes.Structure):
        _fields_ = [('nx', ctypes.c_int), ('ny', ctypes.c_int),
                    ('nz', ctypes.c_int), ('npml', ctypes.c_int),
                    ('tfsf_z', ctypes.c_int), ('read_out', ctypes.c_int),
                    ('nt_inc', ctypes.c_int),
                    ('Db', ct), ('dtm0', ct), ('dx', ct), ('dte0dx', ct),
                    ('omega', ctypes.c_double), ('dt', ctypes.c_double),
                    ('Ca', P), ('Cb', P), ('prof', P), ('inc', P),
                    ('rec', P), ('nthreads', ctypes.c_int)]
    return State, Geom


_STRUCTS = {np.dtype('float32'): _structs(ctypes.c_float),
            np.dtype('float64'): _structs(ctypes.c_double)}
_SUFFIX = {np.dtype('float32'): 'f32', np.dtype('float64'): 'f64'}


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class OracleSim:
    """Fields, psi and phasor state of one simulation, stepped by the C
    oracle.  Shapes are the reference's (cpu/fdtd_sim.py:142-187,
    cpu/main_sim.py:261-307).

    ``geom`` keys: nx, ny, nz, npml, tfsf_z, nt_inc, Db, dtm0, dx, dte0dx,
    omega, dt, and for the incident mode read_out.  ``prof`` is the eight
    CPML profiles concatenated (be_dlf ce_dlf bh_dlf ch_dlf be_urb ce_urb
    bh_urb ch_urb).
    """

    def __init__(self, geom, Ca, Cb, prof, inc=None, dtype='float32',
                 nthreads=None, incident=False):
        self.dtype = np.dtype(dtype)
        ft = self.dtype
        ct = ctypes.c_float if ft == np.float32 else ctypes.c_double
        self._ct = ct
        State, Geom = _STRUCTS[ft]
        nx, ny, nz, npml = geom['nx'], geom['ny'], geom['nz'], geom['npml']
        self.nx, self.ny, self.nz, self.npml = nx, ny, nz, npml
        self.incident = incident
        self.fields = {n: np.zeros((nz, ny, nx), ft) for n in FIELD_NAMES}
        shapes = {'d': (npml, ny, nx), 'u': (npml, ny, nx),
                  'l': (nz, ny, npml), 'r': (nz, ny, npml),
                  'f': (nz, npml, nx), 'b': (nz, npml, nx)}
        self.psi = {n: np.zeros(shapes[n[-1]], ft) for n in PSI_NAMES}
        nxp, nyp, nzp = nx - 2 * npml, ny - 2 * npml, nz - 2 * npml
        ph_shape = {'d': (nyp, nxp), 'u': (nyp, nxp), 'l': (nzp, nyp),
                    'r': (nzp, nyp), 'f': (nzp, nxp), 'b': (nzp, nxp)}
        self.ph = [np.zeros(ph_shape[n[-1]] + (2,), ft) for n in PHASOR_NAMES]
        self.Ca = np.ascontiguousarray(Ca, ft)
        self.Cb = np.ascontiguousarray(Cb, ft)
        self.prof = np.ascontiguousarray(prof, ft)
        self.inc = None if inc is None else np.ascontiguousarray(inc, ft)
        self.rec = None

        st = State()
        for n in FIELD_NAMES:
            setattr(st, n, _ptr(self.fields[n], ct))
        for n in PSI_NAMES:
            setattr(st, n, _ptr(self.psi[n], ct))
        for i in range(24):
            st.ph[i] = _ptr(self.ph[i], ct)
        g = Geom()
        g.nx, g.ny, g.nz, g.npml = nx, ny, nz, npml
        g.tfsf_z = geom['tfsf_z']
        g.read_out = geom.get('read_out', 0)
        g.nt_inc = geom['nt_inc']
        g.Db, g.dtm0, g.dx, g.dte0dx = (geom['Db'], geom['dtm0'], geom['dx'],
                                        geom.get('dte0dx', 0.0))
        g.omega, g.dt = geom['omega'], geom['dt']
        g.Ca, g.Cb, g.prof = (_ptr(self.Ca, ct), _ptr(self.Cb, ct),
                              _ptr(self.prof, ct))
        g.inc = _ptr(self.inc, ct) if self.inc is not None else None
        g.rec = None
        g.nthreads = nthreads or os.cpu_count() or 1
        self._st, self._g = st, g
        self.t = 0
        sfx = _SUFFIX[ft]
        L = lib()
        self._run = getattr(L, f'oracle_main_run_{sfx}')
        self._inc_run = getattr(L, f'oracle_incident_run_{sfx}')
        self._h_half = getattr(L, f'oracle_h_half_{sfx}')
        self._e_half = getattr(L, f'oracle_e_half_{sfx}')
        for f in (self._run, self._inc_run):
            f.argtypes = [ctypes.POINTER(Geom), ctypes.POINTER(State),
                          ctypes.c_int, ctypes.c_int]
            f.restype = None
        for f in (self._h_half, self._e_half):
            f.argtypes = [ctypes.POINTER(Geom), ctypes.POINTER(State)]
            f.restype = None

    def run(self, nsteps):
        """Advance ``nsteps`` main-simulation steps from the current t."""
        self._run(ctypes.byref(self._g), ctypes.byref(self._st),
                  self.t, self.t + nsteps)
        self.t += nsteps

    def h_half(self):
        self._h_half(ctypes.byref(self._g), ctypes.byref(self._st))

    def e_half(self):
        self._e_half(ctypes.byref(self._g), ctypes.byref(self._st))

    def run_incident(self, nsteps, record=True):
        """Advance ``nsteps`` incident-simulation steps; returns the recorded
        read-out planes (4, nsteps, ny, nx) when ``record``."""
        rec = None
        if record:
            rec = np.zeros((4, nsteps, self.ny, self.nx), self.dtype)
            self._g.rec = _ptr(rec, self._ct)
        self._inc_run(ctypes.byref(self._g), ctypes.byref(self._st),
                      self.t, self.t + nsteps)
        self._g.rec = None
        self.t += nsteps
        return rec

    def phasors(self):
        """Dict name -> complex array with the reference's shapes."""
        cdt = np.complex64 if self.dtype == np.float32 else np.complex128
        out = {}
        for n, a in zip(PHASOR_NAMES, self.ph):
            out[n] = (a[..., 0] + 1j * a[..., 1].astype(np.float64)).astype(cdt)
        return out

    def incident_phasors(self):
        """ex, ey, hx, hy incident phasors (first four slots)."""
        cdt = np.complex64 if self.dtype == np.float32 else np.complex128
        names = ['ex_inc_phasor', 'ey_inc_phasor', 'hx_inc_phasor',
                 'hy_inc_phasor']
        return {n: (self.ph[i][..., 0] + 1j * self.ph[i][..., 1]
                    .astype(np.float64)).astype(cdt)
                for i, n in enumerate(names)}


# --------------------------------------------------------------------------
# Post-processing restatements (FP64)
# --------------------------------------------------------------------------

def far_field(ph, nx, ny, nz, dx, r, thetas_deg, phis_deg, freq):
    """FarFieldTransform.__call__ (near2far.py:28-112) over the top face.

    Returns dict E_theta, E_phi, H_theta, H_phi, each squeezed
    (nphi, ntheta) complex128.
    """
    eta0 = 376.730313668
    c0 = 299792458
    lam = c0 / freq
    k = 2.0 * np.pi / lam
    thetas = np.deg2rad(np.asarray(thetas_deg))
    phis = np.deg2rad(np.asarray(phis_deg))
    P = {n: np.asarray(ph[n], dtype='complex128')
         for n in ('ex_ph_u', 'ey_ph_u', 'hx_ph_u', 'hy_ph_u')}
    x_loc = np.arange(-(nx - 1) // 2, ((nx - 1) // 2) + 1) * dx
    y_loc = np.arange(-(ny - 1) // 2, ((ny - 1) // 2) + 1) * dx
    z_loc = np.arange(-(nz - 1) // 2, ((nz - 1) // 2) + 1) * dx
    w = 1j * k * np.exp(-1j * k * r) / (4.0 * np.pi * r)
    out = {n: np.zeros((len(phis), len(thetas)), 'complex128')
           for n in ('E_theta', 'E_phi', 'H_theta', 'H_phi')}
    Jx, Jy = -P['hy_ph_u'], P['hx_ph_u']
    Mx, My = P['ey_ph_u'], -P['ex_ph_u']
    xx, yy = np.meshgrid(x_loc, y_loc)
    zt = z_loc[-1]
    simpson = scipy.integrate.simpson
    for pi, phi in enumerate(phis):
        for ti, th in enumerate(thetas):
            e = np.exp(1j * k * ((xx * np.sin(th) * np.cos(phi))
                                 + (yy * np.sin(th) * np.sin(phi))
                                 + (zt * np.cos(th))))
            n_t = (Jx * np.cos(th) * np.cos(phi)
                   + Jy * np.cos(th) * np.sin(phi)) * e
            n_p = (-Jx * np.sin(phi) + Jy * np.cos(phi)) * e
            l_t = (Mx * np.cos(th) * np.cos(phi)
                   + My * np.cos(th) * np.sin(phi)) * e
            l_p = (-Mx * np.sin(phi) + My * np.cos(phi)) * e
            N_t = simpson(simpson(n_t, dx=dx, axis=1), dx=dx, axis=0)
            N_p = simpson(simpson(n_p, dx=dx, axis=1), dx=dx, axis=0)
            L_t = simpson(simpson(l_t, dx=dx, axis=1), dx=dx, axis=0)
            L_p = simpson(simpson(l_p, dx=dx, axis=1), dx=dx, axis=0)
            out['E_theta'][pi, ti] = w * (-L_p - eta0 * N_t)
            out['E_phi'][pi, ti] = w * (L_t - eta0 * N_p)
            out['H_theta'][pi, ti] = w * (N_p - L_t / eta0)
            out['H_phi'][pi, ti] = w * (-N_t - L_p / eta0)
    return {n: np.squeeze(v) for n, v in out.items()}


def farfield_power(e_theta, e_phi, h_theta, h_phi):
    """processresults.py:260-271."""
    return (0.5 * np.real(-e_phi * np.conj(h_theta)),
            0.5 * np.real(e_theta * np.conj(h_phi)))


def incident_power(inc, dx):
    """Incident power through the read-out plane, processresults.py:144-155."""
    E = np.stack([inc['ex_inc_phasor'], inc['ey_inc_phasor'],
                  np.zeros_like(inc['ey_inc_phasor'])], axis=-1)
    H = np.stack([inc['hx_inc_phasor'], inc['hy_inc_phasor'],
                  np.zeros_like(inc['hy_inc_phasor'])], axis=-1)
    pxy = np.dot(np.real(0.5 * np.cross(E, np.conj(H))),
                 np.asarray([0.0, 0.0, -1.0]))
    py = scipy.integrate.simpson(pxy, dx=dx, axis=1)
    return scipy.integrate.simpson(py, dx=dx, axis=0)


def sigma0(ff, p_inc, nx_ph, ny_ph, dx, radius=2.0e3):
    """Scattering coefficients (h, v), processresults.py:104-121."""
    ph, pv = farfield_power(ff['E_theta'], ff['E_phi'], ff['H_theta'],
                            ff['H_phi'])
    scale = 4.0 * np.pi * (radius ** 2) / (nx_ph * ny_ph * dx * dx)
    return ph / p_inc * scale, pv / p_inc * scale
