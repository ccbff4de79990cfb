"""Gaussian-correlated rough surfaces (materials/roughsurface.py:12-68).

White noise from ``numpy.random.default_rng(seed).normal`` is convolved with
a Gaussian kernel of 6*round(l/dx) taps per axis (``fftconvolve``, 'valid'),
then shifted to zero mean, scaled to unit standard deviation and multiplied
by the rms height.  The same seed gives the same surface as the reference.
"""
import numpy as np
import scipy.signal


def gaussian_kernel(dx, corr_len_x, corr_len_y):
    """Separable Gaussian kernel exp(-(x^2/lx^2 + y^2/ly^2)) sampled on
    [-K/2, K/2) * dx with K = int(6 * round(l / dx)) per axis."""
    kx = int(6.0 * round(corr_len_x / dx))
    ky = int(6.0 * round(corr_len_y / dx))
    xs = np.arange(-kx // 2, kx // 2) * dx
    ys = np.arange(-ky // 2, ky // 2) * dx
    xx, yy = np.meshgrid(xs, ys)
    return np.exp(-(xx ** 2 / corr_len_x ** 2 + yy ** 2 / corr_len_y ** 2))


def gaussian_correlated_surface(nx, ny, dx, corr_len_x, corr_len_y,
                                height_rms, seed=None, rabus_compat=False,
                                debug_plots=False):
    """Surface height [m], shape (ny, nx), with Gaussian autocorrelation.

    ``rabus_compat`` divides the correlation lengths by sqrt(2) as the
    reference does; ``debug_plots`` is accepted and ignored (no plotting).
    """
    if rabus_compat:
        corr_len_x /= np.sqrt(2.0)
        corr_len_y /= np.sqrt(2.0)
    kernel = gaussian_kernel(dx, corr_len_x, corr_len_y)
    ky, kx = kernel.shape
    rng = np.random.default_rng(seed=seed)
    noise = rng.normal(0, 1, size=(ny + ky - 1, nx + kx - 1))
    corr = scipy.signal.fftconvolve(noise, kernel, mode='valid')
    surf = corr - np.mean(corr)
    surf /= np.std(surf)
    surf *= height_rms
    return surf
