"""ORACLE -- test infrastructure only.

numpy restatement of the device material builder (fdtd_sim_build_materials)
that tests/ compare the device results against; the product never imports
this module.

Per cell, in order: vacuum; a voxel inclusion where z < vox_z1 and the
cell's draw 3 is below vox_frac; the last cylinder containing the cell
centre (the loop of the host canopy generator, a restatement of the
reference's cell-centre fill); the first layer k with z <= surface k:
constant values, random soil (Box-Muller normal from draws 0 and 1,
uniform sigma from draw 2), or the Hallikainen moisture gradient of
materials/soilmodels.py:233-247 evaluated with numpy.interp over the whole
table (dielectricmodels.py:219-260).  Ca, Cb then follow
materials/materials.py:23-68 in float32.

Draws: SplitMix64 (Steele, Lea, Flood 2014) of the key
(seed << 36) ^ (cell << 2) ^ draw, top 53 bits scaled to [0, 1); cell is
the unpadded index (z ny + y) nx + x.
"""
import numpy as np
from scipy.constants import epsilon_0

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def mix64(z):
    """SplitMix64 output for state z (the first output of a generator seeded
    with z)."""
    with np.errstate(over='ignore'):             # arithmetic modulo 2^64
        z = np.asarray(z, dtype=np.uint64) + GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def draw01(seed, cell, j):
    key = (np.uint64(seed) << np.uint64(36)) ^ (np.asarray(cell, dtype=np.uint64) << np.uint64(2)) \
        ^ np.uint64(j)
    return (mix64(key) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def random_soil(seed, cells, eps_mean, eps_std, sig_lo, sig_hi):
    """(eps_r, sigma) float64 of the random-soil layer at the given cells."""
    u1 = 1.0 - draw01(seed, cells, 0)
    u2 = draw01(seed, cells, 1)
    n = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    eps = np.maximum(eps_mean + eps_std * n, 1.0)
    sig = sig_lo + (sig_hi - sig_lo) * draw01(seed, cells, 2)
    return eps, sig


def paint_cylinders(eps, sig, cylinders, dx):
    """Cell-centre cylinder fill, later cylinders over earlier ones."""
    nz, ny, nx = eps.shape
    zc, yc, xc = np.arange(nz) * dx, np.arange(ny) * dx, np.arange(nx) * dx
    for cy in cylinders:
        u = np.asarray(cy.axis, dtype=np.float64)
        c = np.asarray(cy.centre, dtype=np.float64)
        half = 0.5 * cy.length * np.abs(u) + cy.radius * np.sqrt(np.maximum(0.0, 1.0 - u * u))
        lo = np.maximum(np.floor((c - half) / dx).astype(int), 0)
        hi = np.minimum(np.ceil((c + half) / dx).astype(int) + 1, [nz, ny, nx])
        if np.any(hi <= lo):
            continue
        dz = (zc[lo[0]:hi[0]] - c[0])[:, None, None]
        dy = (yc[lo[1]:hi[1]] - c[1])[None, :, None]
        dxx = (xc[lo[2]:hi[2]] - c[2])[None, None, :]
        t = dz * u[0] + dy * u[1] + dxx * u[2]
        perp2 = dz * dz + dy * dy + dxx * dxx - t * t
        inside = (np.abs(t) <= 0.5 * cy.length) & (perp2 <= cy.radius * cy.radius)
        eps[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]][inside] = cy.eps
        sig[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]][inside] = cy.sigma


def hallikainen(mv, layer, table, coeffs):
    """eps' and eps'' of the Hallikainen model at moisture mv: the table
    polynomials a + b mv + c mv^2, interpolated in frequency by numpy.interp
    (coeffs: a, b, c (real), a, b, c (imaginary) at the table rows)."""
    a_re, b_re, c_re, a_im, b_im, c_im = coeffs
    re = np.interp(layer.freq, table, a_re + (b_re * mv) + (c_re * (mv ** 2)))
    im = np.interp(layer.freq, table, a_im + (b_im * mv) + (c_im * (mv ** 2)))
    return re, im


def eps_sigma(spec, shape, dx, table=None, coeffs_of=None):
    """float32 (nz_mat, ny, nx) eps_r and sigma of a MaterialSpec.
    ``coeffs_of(layer)`` gives the Hallikainen table polynomials of a
    gradient layer."""
    nz, ny, nx = shape
    eps = np.ones(shape, np.float64)
    sig = np.zeros(shape, np.float64)
    cell = np.arange(nz * ny * nx, dtype=np.uint64).reshape(shape)
    z = np.arange(nz).reshape(nz, 1, 1)
    if spec.vox_z1 > 0:
        hit = (z < spec.vox_z1) & (draw01(spec.vox_seed, cell, 3) < spec.vox_frac)
        eps[hit] = spec.vox_eps
        sig[hit] = spec.vox_sigma
    paint_cylinders(eps, sig, spec.cylinders, dx)
    taken = np.zeros(shape, bool)
    for k, layer in enumerate(spec.layers):
        h = spec.surfaces[k].astype(np.int64)[np.newaxis]
        m = (z <= h) & ~taken
        taken |= m
        if layer.kind == 0:
            eps[m], sig[m] = layer.eps, layer.sigma
        elif layer.kind == 1:
            e, s = random_soil(layer.seed, cell[m], layer.eps, layer.eps_std, layer.sigma,
                               layer.sigma_hi)
            eps[m], sig[m] = e, s
        else:
            h0 = np.broadcast_to(spec.surfaces[k - 1].astype(np.int64)[np.newaxis], shape)[m]
            h1 = np.broadcast_to(h, shape)[m]
            zz = np.broadcast_to(z, shape)[m]
            mv = ((layer.mv_bottom * (h1 - zz)) + (layer.mv_top * (zz - h0))) / (h1 - h0)
            coeffs = coeffs_of(layer)
            uniq, inv = np.unique(mv, return_inverse=True)
            vals = [hallikainen(v, layer, table, coeffs) for v in uniq]
            re = np.array([v[0] for v in vals])[inv]
            im = np.array([v[1] for v in vals])[inv]
            eps[m] = re
            sig[m] = np.abs(-im) * epsilon_0 * 2.0 * np.pi * layer.freq
    return eps.astype(np.float32), sig.astype(np.float32)


def ca_cb(eps_mat, sig_mat, nz, dt, dx, eps0):
    """Ca, Cb float32 (nz, ny, nx): materials.py:23-68 with vacuum above."""
    nzm, ny, nx = eps_mat.shape
    perm = np.concatenate((eps_mat, np.ones((nz - nzm, ny, nx), np.float32))).astype(np.float32)
    cond = np.concatenate((sig_mat, np.zeros((nz - nzm, ny, nx), np.float32))).astype(np.float32)
    dt, dx, eps0 = np.float32(dt), np.float32(dx), np.float32(eps0)
    perm = perm * eps0
    loss = (cond * dt) / (np.float32(2.0) * perm)
    Ca = (np.float32(1.0) - loss) / (np.float32(1.0) + loss)
    Cb = (dt / (perm * dx)) / (np.float32(1.0) + loss)
    return Ca, Cb
