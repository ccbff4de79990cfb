"""ORACLE — test infrastructure only.

Host restatement of the incident-beam synthesis (fdtdsarsim/incidentbeam.py:
206-433): Braunisch's tapered-wave spectrum per temporal frequency bin
(``taper_k``), its 2-D FFT to the spatial field, the inverse FFT over time
and the normalisation of the peak to 10 (``create_inc_excitation``).  The
product synthesises the beam on the device (csrc/fdtd_beam.hip); the tests
check that pipeline against this one and pin this one with closed forms and
explicit DFT sums.  Nothing in the product imports this module.

``taper_k`` keeps the reference's dtype flow (k_z formed in complex64, the
rest in float64/complex128) so that its spectra equal the reference's.
"""
import numpy as np

from sar_pixel_simulator_amd.constants import C0
from sar_pixel_simulator_amd.incidentbeam import IncidentBeam


def taper_k(k, k_x, k_y, theta, phi, radius, pol):
    """Tapered-wave spectrum (Braunisch), incidentbeam.py:316-433.

    Returns the (x, y) spectral field components for wavenumber ``k`` on the
    (ny, nx) spatial-frequency grid.  Same dtype flow as the reference: k_z
    is formed in complex64, the rest in float64/complex128.
    """
    k_p = np.hypot(k_x, k_y)
    k_z = np.conj(np.sqrt((k ** 2 - k_p ** 2).astype(np.complex64)))
    theta = np.radians(theta)
    phi = np.radians(phi)
    k_ix = k * np.sin(theta) * np.cos(phi)
    k_iy = k * np.sin(theta) * np.sin(phi)
    k_ip = np.hypot(k_ix, k_iy)
    k_iz = np.sqrt(k ** 2 - k_ip ** 2)
    gauss = ((radius ** 2) / (4 * np.pi)) * np.exp(
        -(radius ** 2) * (((k_x - k_ix) ** 2) + ((k_y - k_iy) ** 2)) / 4)

    pos = k_p > 0.0
    kp = np.where(pos, k_p, 1.0)
    hor_x = np.where(pos, k_y / kp, np.sin(phi)).astype(np.complex128)
    hor_y = np.where(pos, -k_x / kp, -np.cos(phi)).astype(np.complex128)
    ver_x = np.where(pos, (k_z * k_x) / (k * kp), np.cos(phi)).astype(np.complex128)
    ver_y = np.where(pos, (k_z * k_y) / (k * kp), np.sin(phi)).astype(np.complex128)
    ver_z = np.where(pos, k_p / k, 0.0).astype(np.complex128)

    if k_ip > 0:
        hor_ix, hor_iy = k_iy / k_ip, -k_ix / k_ip
        ver_ix, ver_iy = (k_iz * k_ix) / (k * k_ip), (k_iz * k_iy) / (k * k_ip)
        ver_iz = k_ip / k
    else:
        hor_ix, hor_iy = np.sin(phi), -np.cos(phi)
        ver_ix, ver_iy, ver_iz = np.cos(phi), np.sin(phi), 0.0

    if pol == 'HH':
        Eh, Ev = 1, 0
    elif pol == 'VV':
        Eh, Ev = 0, 1
    elif pol == 'RHCP':
        Eh, Ev = -1j, 1
    elif pol == 'LHCP':
        Eh, Ev = 1j, 1
    else:
        raise ValueError('Invalid polarization value')

    eh = Eh * ((hor_ix * hor_x) + (hor_iy * hor_y)) \
        + Ev * ((ver_ix * hor_x) + (ver_iy * hor_y))
    ev = Eh * ((hor_ix * ver_x) + (hor_iy * ver_y)) \
        + Ev * ((ver_ix * ver_x) + (ver_iy * ver_y) + (ver_iz * ver_z))
    return gauss * ((eh * hor_x) + (ev * ver_x)), \
        gauss * ((eh * hor_y) + (ev * ver_y))


def create_inc_excitation(params, workdir=None, rows=16):
    """Ex, Ey excitation of the incident simulation, (nt_inc, ny, nx)
    complex64 (complex128 for datatype float64) (incidentbeam.py:206-313),
    from the pulse of IncidentBeam.create_pulse (incidentbeam.py:174-204).

    With ``workdir`` the spectra and the result live in memory-mapped files
    there and the inverse FFT over time runs ``rows`` grid rows at a time:
    the same values per pixel, for grids whose (nt_inc, ny, nx) complex
    arrays do not fit in memory several times over."""
    ib = IncidentBeam(params)
    ib.create_pulse()
    g, b = params.geom_params, params.inc_beam_params
    dtype_c = {'float32': np.complex64, 'float64': np.complex128}[
        params.processing_params.datatype]
    width = b.illumination_width * C0 / b.frequency
    gauss_radius = width / (2.0 * np.sqrt(2.0 * np.log(10.0)))
    nt_inc, nx, ny, dx = g.size_t_inc, g.size_x, g.size_y, g.cell_size
    kx = 2 * np.pi * np.repeat(np.atleast_2d(np.fft.fftfreq(nx, dx)), ny, axis=0)
    kx = np.fft.fftshift(kx.reshape((ny, nx)), axes=1)
    ky = 2 * np.pi * np.repeat(np.atleast_2d(np.fft.fftfreq(ny, dx)).transpose(),
                               nx, axis=1)
    ky = np.fft.fftshift(ky.reshape((ny, nx)), axes=0)
    shape = (nt_inc, ny, nx)
    if workdir is None:
        pulse_ex = np.zeros(shape, dtype_c)
        pulse_ey = np.zeros(shape, dtype_c)
    else:
        import os
        pulse_ex, pulse_ey = (np.lib.format.open_memmap(os.path.join(workdir, f'pulse_{c}.npy'),
                                                        mode='w+', dtype=dtype_c, shape=shape)
                              for c in ('ex', 'ey'))
    all_k = 2 * np.pi * ib.pulse_freqs / C0
    sh, ish, fft2 = np.fft.fftshift, np.fft.ifftshift, np.fft.fft2
    for k_idx in range(nt_inc):
        k_val = all_k[k_idx]
        if k_val == 0.0:
            pulse_ex[k_idx] = 0
            pulse_ey[k_idx] = 0
            continue
        sx, sy = taper_k(k_val, kx, ky, b.theta_angle, b.phi_angle,
                         gauss_radius, b.polarization)
        w = 1.0 * ib.pulse_fft[k_idx]
        pulse_ex[k_idx] = w * sh(fft2(ish(sx)))
        pulse_ey[k_idx] = w * sh(fft2(ish(sy)))
    if workdir is None:
        pulse_ex = np.fft.ifft(pulse_ex, axis=0).astype(dtype_c, copy=False)
        pulse_ey = np.fft.ifft(pulse_ey, axis=0).astype(dtype_c, copy=False)
        maxval = float(max(np.abs(pulse_ex).max(), np.abs(pulse_ey).max()))
        scale = 10.0 / maxval
        pulse_ex = (pulse_ex.astype(np.complex128) * scale).astype(dtype_c)
        pulse_ey = (pulse_ey.astype(np.complex128) * scale).astype(dtype_c)
        return pulse_ex, pulse_ey
    maxval = 0.0
    for arr in (pulse_ex, pulse_ey):
        for y0 in range(0, ny, rows):
            v = np.fft.ifft(arr[:, y0:y0 + rows], axis=0).astype(dtype_c, copy=False)
            arr[:, y0:y0 + rows] = v
            maxval = max(maxval, float(np.abs(v).max()))
    scale = 10.0 / maxval
    for arr in (pulse_ex, pulse_ey):
        for y0 in range(0, ny, rows):
            arr[:, y0:y0 + rows] = (arr[:, y0:y0 + rows].astype(np.complex128) * scale).astype(dtype_c)
        arr.flush()
    return pulse_ex, pulse_ey
