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


def two_layer_gaussian_correlated_surface(nx, ny, dx, corr_len_x, corr_len_y, height_rms,
                                          inter_layer_corr, seed=None, rabus_compat=False,
                                          debug_plots=False):
    """Two surfaces [m], shape (2, ny, nx), each with Gaussian
    autocorrelation, correlated with each other by ``inter_layer_corr``
    (materials/roughsurface.py:116-224).

    White noise of shape (ny + ky - 1, nx + kx - 1, 2) from
    ``default_rng(seed).normal`` is mixed per pixel by the Cholesky factor
    of [[1, rho], [rho, 1]] (rho capped at 1 - 1e-5), each layer is convolved
    with the Gaussian kernel ('valid', direct convolution), shifted to zero
    mean, scaled to unit standard deviation and multiplied by the rms height.
    """
    if rabus_compat:
        corr_len_x /= np.sqrt(2.0)
        corr_len_y /= np.sqrt(2.0)
    rho = min(inter_layer_corr, 1.0 - 1e-5)
    kernel = gaussian_kernel(dx, corr_len_x, corr_len_y)
    ky, kx = kernel.shape
    rng = np.random.default_rng(seed=seed)
    noise = rng.normal(0, 1, size=(ny + ky - 1, nx + kx - 1, 2))
    chol = np.linalg.cholesky(np.array([[1.0, rho], [rho, 1.0]]))
    # per pixel: chol @ (n0, n1); chol is lower triangular
    layers = np.empty_like(noise)
    layers[..., 0] = chol[0, 0] * noise[..., 0] + chol[0, 1] * noise[..., 1]
    layers[..., 1] = chol[1, 0] * noise[..., 0] + chol[1, 1] * noise[..., 1]
    surf = np.zeros((2, ny, nx))
    for i in range(2):
        corr = scipy.signal.convolve2d(layers[..., i], kernel, mode='valid')
        surf[i] = corr - np.mean(corr)
        surf[i] /= np.std(surf[i])
        surf[i] *= height_rms
    return surf
