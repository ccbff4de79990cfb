"""Near-to-far transform (Taflove ch. 8) over the top face of the phasor box.

Same call signature and outputs as fdtdsarsim/near2far.py:11-112; the
surface integrals run in FP64 on the device over the whole theta x phi grid
in one launch (libfdtd_hip ``fdtd_near2far_host_grid``): J = (-Hy_u, Hx_u),
M = (Ey_u, -Ex_u), composite Simpson weights in x and y, r, theta/phi in
degrees.
"""
from typing import Mapping

import numpy as np

from . import _native as native


class FarFieldTransform:

    def __call__(self, phasors: Mapping[str, np.ndarray], nx: int, ny: int, nz: int,
                 dx: float, r: float, thetas: np.ndarray, phis: np.ndarray, freq: float):
        """Far-field E_theta, E_phi, H_theta, H_phi at (r, theta, phi), each of
        shape (len(phis), len(thetas)), squeezed (near2far.py:28-112)."""
        thetas = np.asarray(thetas, dtype=np.float64)
        phis = np.asarray(phis, dtype=np.float64)
        if thetas.ndim != 1 or phis.ndim != 1:
            raise RuntimeError('Theta and phi arguments must be 1D arrays')
        planes = [np.asarray(phasors[n]) for n in ('ex_ph_u', 'ey_ph_u', 'hx_ph_u', 'hy_ph_u')]
        # the whole theta x phi grid in one launch
        out = native.near2far_host_grid(*planes, nx, ny, nz, dx, thetas, phis, r, freq)
        return {k: np.squeeze(v) for k, v in out.items()}
