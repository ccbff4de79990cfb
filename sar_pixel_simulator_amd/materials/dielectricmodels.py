"""Hallikainen wet-soil permittivity (materials/dielectricmodels.py:11-137,
219-260).

eps(mv) = a + b mv + c mv^2 for the real and imaginary parts, where each of
a, b, c is a linear function of the sand and clay fractions (Table 2 of
Hallikainen, Ulaby, Dobson, El-Rayes and Wu, "Microwave dielectric behavior
of wet soil -- part 1", IEEE TGRS 1985), tabulated at nine frequencies from
1.4 to 18 GHz and interpolated linearly in frequency.  The returned complex
permittivity is eps' - j eps''.

``soil_eps_sigma`` gives the (eps_r, sigma) pair the soil generators store:
sigma = |eps''| eps0 2 pi f (soilmodels.py:64-66), with scipy's eps0.
"""
import numpy as np
from scipy.constants import epsilon_0

#: Table frequencies [Hz] (dielectricmodels.py:24-25).
FREQ_TABLE = np.array([1.4, 4, 6, 8, 10, 12, 14, 16, 18]) * 1.0e9

# Rows: table frequencies; columns: coefficient of (1, sand, clay).
A_RE = np.array([[2.862, -0.012, 0.001], [2.927, -0.012, -0.001], [1.993, 0.002, 0.015],
                 [1.997, 0.002, 0.018], [2.502, -0.003, -0.003], [2.200, -0.001, 0.012],
                 [2.301, 0.001, 0.009], [2.237, 0.002, 0.009], [1.912, 0.007, 0.021]])
A_IM = np.array([[0.356, -0.003, -0.008], [0.004, 0.001, 0.002], [-0.123, 0.002, 0.003],
                 [-0.201, 0.003, 0.003], [-0.070, 0.000, 0.001], [-0.142, 0.001, 0.003],
                 [-0.096, 0.001, 0.002], [-0.027, -0.001, 0.003], [-0.071, 0.000, 0.003]])
B_RE = np.array([[3.803, 0.462, -0.341], [5.505, 0.371, 0.062], [38.086, -0.176, -0.633],
                 [25.579, -0.017, -0.412], [10.101, 0.221, -0.004], [26.473, 0.013, -0.523],
                 [17.918, 0.084, -0.282], [15.505, 0.076, -0.217], [29.123, -0.190, -0.545]])
B_IM = np.array([[5.507, 0.044, -0.002], [0.951, 0.005, -0.010], [7.502, -0.058, -0.116],
                 [11.266, -0.085, -0.155], [6.620, 0.015, -0.081], [11.868, -0.059, -0.225],
                 [8.583, -0.005, -0.153], [6.179, 0.074, -0.086], [6.938, 0.029, -0.128]])
C_RE = np.array([[119.006, -0.500, 0.633], [114.826, -0.389, -0.547], [10.720, 1.256, 1.522],
                 [39.793, 0.723, 0.941], [77.482, -0.061, -0.135], [34.333, 0.284, 1.062],
                 [50.149, 0.012, 0.387], [48.260, 0.168, 0.289], [6.960, 0.822, 1.195]])
C_IM = np.array([[17.753, -0.313, 0.206], [16.759, 0.192, 0.290], [2.942, 0.452, 0.543],
                 [0.194, 0.584, 0.581], [21.578, 0.293, 0.332], [7.817, 0.570, 0.801],
                 [28.707, 0.297, 0.357], [34.126, 0.143, 0.206], [29.945, 0.275, 0.377]])


def polynomial_coefficients(sand_frac, clay_frac):
    """(a, b, c) for the real and imaginary parts at every table frequency:
    six arrays of 9, each row's dot product with (1, sand, clay)."""
    vec = np.array([1.0, sand_frac, clay_frac])
    return tuple(np.dot(t, vec) for t in (A_RE, B_RE, C_RE, A_IM, B_IM, C_IM))


def hallikainen_soil_model(mv, sand_frac, clay_frac, freq):
    """Complex permittivity eps' - j eps'' without argument checks
    (dielectricmodels.py:219-260)."""
    a_re, b_re, c_re, a_im, b_im, c_im = polynomial_coefficients(sand_frac, clay_frac)
    eps_real = a_re + (b_re * mv) + (c_re * (mv ** 2))
    eps_imag = a_im + (b_im * mv) + (c_im * (mv ** 2))
    return np.interp(freq, FREQ_TABLE, eps_real) - 1j * np.interp(freq, FREQ_TABLE, eps_imag)


class HallikainenSoilModel:
    """Hallikainen model with the reference's argument checks
    (dielectricmodels.py:11-137)."""

    freq_table = FREQ_TABLE

    def get_epsilon(self, mv: float, sand_frac: float, clay_frac: float, freq: float):
        """Complex permittivity of soil with volumetric moisture ``mv``, sand
        and clay fractions of the dry soil, at ``freq`` Hz.  Raises
        ValueError outside 0 <= mv, sand, clay <= 1 or 1.4 <= f <= 18 GHz."""
        if mv < 0.0 or mv > 1.0:
            raise ValueError('Volumetric moisture out of bounds (0 < mv < 1)')
        if sand_frac < 0.0 or sand_frac > 1.0:
            raise ValueError('Sand fraction out of bounds (0 < sand_frac < 1')
        if clay_frac < 0.0 or clay_frac > 1.0:
            raise ValueError('Clay fraction out of bounds (0 < clay_frac < 1)')
        if freq < 1.4e9 or freq > 18.0e9:
            raise ValueError('Frequency out of bounds (1.4 GHz < freq < 18 GHz)')
        return hallikainen_soil_model(mv, sand_frac, clay_frac, freq)


def soil_eps_sigma(eps_cplx, freq):
    """(eps_r, sigma [S/m]) of a complex permittivity eps' - j eps''
    (soilmodels.py:64-66)."""
    return np.real(eps_cplx), np.abs(np.imag(eps_cplx)) * epsilon_0 * 2.0 * np.pi * freq


def table_bracket(freq):
    """Index j of the table interval holding ``freq`` (FREQ_TABLE[j] <= freq
    < FREQ_TABLE[j + 1]; the last interval also holds 18 GHz), as
    numpy.interp locates it."""
    if not FREQ_TABLE[0] <= freq <= FREQ_TABLE[-1]:
        raise ValueError('Frequency out of bounds (1.4 GHz < freq < 18 GHz)')
    j = int(np.searchsorted(FREQ_TABLE, freq, side='right')) - 1
    return min(j, len(FREQ_TABLE) - 2)


__all__ = ['HallikainenSoilModel', 'hallikainen_soil_model', 'soil_eps_sigma',
           'polynomial_coefficients', 'table_bracket', 'FREQ_TABLE']
