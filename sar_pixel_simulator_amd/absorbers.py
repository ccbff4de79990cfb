"""CPML coefficients as 1-D profiles (Taflove ch. 7.9).

The profile arithmetic follows fdtdsarsim/absorbers.py:35-153 operation for
operation in the configured float type, so b and c are bit-identical to the
reference's.  The reference then tiles each profile to full 3-D arrays per
face (absorbers.py:123-185); the kernels here read the 1-D profiles instead,
because every tiled array is constant along the two face-parallel axes:

* down / left / front faces use the ``*_dlf`` profiles, indexed by the
  distance into the absorber along z / x / y respectively;
* up / right / back faces use the ``*_urb`` profiles, likewise.

``tiled()`` rebuilds the reference's 3-D arrays for tests.
"""
import numpy as np

from .constants import EPS0_2018, ETA0
from .parameters import ParameterCollection


class Absorbers:
    """1-D CPML profiles for one geometry (absorbers.py:13-189)."""

    kappa_max = 1
    alpha_max = 0.2
    m_kappa = 3
    m_sigma = 3
    m_alpha = 1

    def __init__(self, params: ParameterCollection):
        self.geom_params = params.geom_params
        self.dtype = np.dtype(params.processing_params.datatype)
        self.setup()

    def setup(self):
        g = self.geom_params
        dx, dt, npml = g.cell_size, g.timestep, g.size_pml
        ft = self.dtype
        feps = np.finfo(ft).eps
        amax, kmax = self.alpha_max, self.kappa_max
        ms, mk, ma = self.m_sigma, self.m_kappa, self.m_alpha
        sigma_max = 0.8 * (ms + 1) / (ETA0 * dx)          # eqn 7.66

        inc = np.arange(npml, dtype=ft)                   # 0 .. npml-1
        dec = npml - 1 - inc                              # npml-1 .. 0
        den = npml - 1

        # down / left / front (eqns 7.79, 7.115, 7.119)
        self.alpha_e_dlf = amax * ((inc / den) ** ma).astype(ft)
        self.alpha_h_dlf = amax * (((inc + 0.5) / den) ** ma).astype(ft)
        self.sigma_e_dlf = sigma_max * ((dec / den) ** ms).astype(ft)
        self.sigma_h_dlf = sigma_max * (((dec - 0.5) / den) ** ms).astype(ft)
        self.kappa_e_dlf = 1 + (kmax - 1) * ((dec / den) ** mk).astype(ft)
        self.kappa_h_dlf = 1 + (kmax - 1) * (((dec + 0.5) / den) ** mk).astype(ft)

        # up / right / back
        self.alpha_e_urb = amax * ((dec / den) ** ma).astype(ft)
        self.alpha_h_urb = amax * (((dec - 0.5) / den) ** ma).astype(ft)
        self.alpha_h_urb[self.alpha_h_urb < 0.0] = 0.0
        self.sigma_e_urb = sigma_max * ((inc / den) ** ms).astype(ft)
        self.sigma_h_urb = sigma_max * (((inc + 0.5) / den) ** ms).astype(ft)
        self.kappa_e_urb = 1 + (kmax - 1) * ((inc / den) ** mk).astype(ft)
        self.kappa_h_urb = 1 + (kmax - 1) * (((inc + 0.5) / den) ** mk).astype(ft)

        def b_coef(s, k, a):                              # eqn 7.102
            return np.exp(-(s / k + a) * dt / EPS0_2018)

        def c_coef(s, k, a):                              # eqn 7.99
            return (s + feps) * (np.exp(-(s / k + a) * dt / EPS0_2018) - 1) \
                / (s * k + (k ** 2 * a) + feps)

        self.be_dlf = b_coef(self.sigma_e_dlf, self.kappa_e_dlf, self.alpha_e_dlf)
        self.bh_dlf = b_coef(self.sigma_h_dlf, self.kappa_h_dlf, self.alpha_h_dlf)
        self.ce_dlf = c_coef(self.sigma_e_dlf, self.kappa_e_dlf, self.alpha_e_dlf)
        self.ch_dlf = c_coef(self.sigma_h_dlf, self.kappa_h_dlf, self.alpha_h_dlf)
        self.be_urb = b_coef(self.sigma_e_urb, self.kappa_e_urb, self.alpha_e_urb)
        self.bh_urb = b_coef(self.sigma_h_urb, self.kappa_h_urb, self.alpha_h_urb)
        self.ce_urb = c_coef(self.sigma_e_urb, self.kappa_e_urb, self.alpha_e_urb)
        self.ch_urb = c_coef(self.sigma_h_urb, self.kappa_h_urb, self.alpha_h_urb)
        for name in ('be_dlf', 'bh_dlf', 'ce_dlf', 'ch_dlf',
                     'be_urb', 'bh_urb', 'ce_urb', 'ch_urb'):
            assert getattr(self, name).dtype == ft, name

    def profiles(self):
        """The eight profiles in the order the native library takes them."""
        return (self.be_dlf, self.ce_dlf, self.bh_dlf, self.ch_dlf,
                self.be_urb, self.ce_urb, self.bh_urb, self.ch_urb)

    def tiled(self, name, nz=None):
        """Reference-shaped 3-D coefficient array, e.g. ``tiled('be_l')``.

        Faces d/u tile along (ny, nx), l/r along (nz, ny), f/b along (nz, nx)
        (absorbers.py:123-185).  ``nz`` overrides size_z (the incident
        simulation uses its own height).
        """
        g = self.geom_params
        nx, ny = g.size_x, g.size_y
        nz = g.size_z if nz is None else nz
        kind, face = name.split('_')
        prof = getattr(self, f'{kind}_{"dlf" if face in "dlf" else "urb"}')
        if face in 'du':
            return np.tile(prof[:, None, None], [1, ny, nx])
        if face in 'lr':
            return np.tile(prof[None, None, :], [nz, ny, 1])
        return np.tile(prof[None, :, None], [nz, 1, nx])
