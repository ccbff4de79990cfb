"""Compulsory (algorithmic) HBM bytes of the stencil and DFT kernels, the
denominator of every roofline figure bench.py and hip/measure.py report."""


def cpml_cells(nx, ny, nz, np_):
    """Cells each half-step's CPML passes touch (cpu/field_updates.py:66-191,
    243-366 index ranges): d/u (np-1)(ny-1)(nx-1), l/r (nz-1)(ny-1)(np-1),
    f/b (nz-1)(np-1)(nx-1), two of each."""
    return 2 * ((np_ - 1) * (ny - 1) * (nx - 1) + (nz - 1) * (ny - 1) * (np_ - 1)
                + (nz - 1) * (np_ - 1) * (nx - 1))


def algorithmic_bytes(nx, ny, nz, np_, uni_z=None, taps=True):
    """Compulsory HBM bytes per launch of each kernel.

    H half: read E(3) + H(3), write H(3) = 36 B/cell; each CPML face cell
    reads and writes two psi values = 16 B; E_inc injection reads 2 floats
    per TFSF-plane cell.  E half: read H(3) + E(3) + Ca + Cb, write E(3) =
    44 B/cell (36 B on the uniform top planes z >= uni_z, where Ca/Cb are
    scalars), same CPML and injection terms, and (split stepper, taps) the
    six x-face column values per (face, z, y) it copies for the DFT, 24 B.
    DFT: per plane sample, four float4 running sums read and written
    (128 B) plus four field values.
    """
    n = nx * ny * nz
    nc = cpml_cells(nx, ny, nz, np_)
    plane = nx * ny
    n_uni = (nz - (nz if uni_z is None else uni_z)) * plane
    nxp, nyp, nzp = nx - 2 * np_, ny - 2 * np_, nz - 2 * np_
    samples = 2 * nyp * nxp + 2 * nzp * nyp + 2 * nzp * nxp
    return {'h': 36 * n + 16 * nc + 8 * plane,
            'e': 44 * n - 8 * n_uni + 16 * nc + 8 * plane + (48 * ny * nz if taps else 0),
            'dft': 144 * samples,
            'cells': n, 'cpml_cells': nc, 'dft_samples': samples}
