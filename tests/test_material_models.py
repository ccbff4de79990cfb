"""Soil models, the Hallikainen permittivity and the device material
builder's oracle, on the CPU.

* Hallikainen: hand-evaluated a + b mv + c mv^2 at a table frequency and
  halfway between two (published Table 2 coefficients), plus the argument
  checks of dielectricmodels.py:99-128.
* soilmodels: the vectorised fills against pure-Python loops written as the
  reference's numba loops are (soilmodels.py:43-50, 126-137, 233-247).
* materials oracle: SplitMix64 known answers (the published first outputs
  for seed 0), uniform and normal statistics, and its cylinder fill against
  the host canopy generator.
* C-ABI structs: ctypes offsets against the C compiler's.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, 'oracle'))
import materials_oracle as mo  # noqa: E402

from sar_pixel_simulator_amd import _native as native, workloads  # noqa: E402
from sar_pixel_simulator_amd.materials import device, soilmodels, synthetic  # noqa: E402
from sar_pixel_simulator_amd.materials.dielectricmodels import (  # noqa: E402
    HallikainenSoilModel, hallikainen_soil_model, soil_eps_sigma, table_bracket)


# ---------------------------------------------------------------------------
# Hallikainen
# ---------------------------------------------------------------------------
def test_hallikainen_at_a_table_frequency_by_hand():
    # 4 GHz row: a = 2.927 - 0.012 S - 0.001 C, b = 5.505 + 0.371 S + 0.062 C,
    # c = 114.826 - 0.389 S - 0.547 C (real); a = 0.004 + 0.001 S + 0.002 C,
    # b = 0.951 + 0.005 S - 0.010 C, c = 16.759 + 0.192 S + 0.290 C (imag.)
    S, C, mv = 0.3, 0.2, 0.25
    re = (2.927 - 0.012 * S - 0.001 * C) + (5.505 + 0.371 * S + 0.062 * C) * mv \
        + (114.826 - 0.389 * S - 0.547 * C) * mv * mv
    im = (0.004 + 0.001 * S + 0.002 * C) + (0.951 + 0.005 * S - 0.010 * C) * mv \
        + (16.759 + 0.192 * S + 0.290 * C) * mv * mv
    eps = HallikainenSoilModel().get_epsilon(mv, S, C, 4.0e9)
    assert abs(eps.real - 11.49286875) < 1e-9 and abs(re - 11.49286875) < 1e-9
    assert abs(-eps.imag - 1.2969875) < 1e-9 and abs(im - 1.2969875) < 1e-9


def test_hallikainen_interpolates_linearly_in_frequency():
    S, C, mv = 0.5, 0.1, 0.3
    e4 = hallikainen_soil_model(mv, S, C, 4.0e9)
    e6 = hallikainen_soil_model(mv, S, C, 6.0e9)
    e5 = hallikainen_soil_model(mv, S, C, 5.0e9)
    assert abs(e5 - 0.5 * (e4 + e6)) < 1e-12
    # dry soil: eps = a
    e0 = hallikainen_soil_model(0.0, 0.0, 0.0, 8.0e9)
    assert abs(e0 - (1.997 - 1j * -0.201)) < 1e-12


def test_hallikainen_argument_checks():
    m = HallikainenSoilModel()
    for args in [(-0.1, 0.5, 0.2, 5e9), (1.1, 0.5, 0.2, 5e9), (0.2, 1.5, 0.2, 5e9),
                 (0.2, 0.5, -0.2, 5e9), (0.2, 0.5, 0.2, 1.0e9), (0.2, 0.5, 0.2, 19e9)]:
        with pytest.raises(ValueError):
            m.get_epsilon(*args)
    assert table_bracket(1.4e9) == 0 and table_bracket(18.0e9) == 7 and table_bracket(5.4e9) == 1


def test_soil_sigma_from_imaginary_part():
    eps = 12.5 - 2.0j
    e, s = soil_eps_sigma(eps, 5.0e9)
    from scipy.constants import epsilon_0
    assert e == 12.5 and s == 2.0 * epsilon_0 * 2.0 * np.pi * 5.0e9


# ---------------------------------------------------------------------------
# soil generators vs loops
# ---------------------------------------------------------------------------
def _params(nx=23, ny=19, nz_mat=14):
    pc = workloads.small_params(nx=nx, ny=ny, nz=nz_mat + 24, nz_mat=nz_mat)
    return pc


def test_one_layer_fill_matches_loops():
    pc = _params()
    g = pc.geom_params
    dx = g.cell_size
    args = dict(height_rms=0.8 * dx, corr_len=4 * dx, mean_depth=7 * dx, count=2, seed=5)
    surfs = list(soilmodels.one_layer_surfaces(pc, **args))
    models = list(soilmodels.one_layer_gaussian(pc, [0.1, 0.3], 0.4, 0.2, **args))
    assert len(models) == 4 and models[1].name == 'mv_0.30_num_0'
    for i, m in enumerate(models):
        surf = surfs[i // 2]
        e_r, s_r = m.metadata['epsilon_r'], m.metadata['conductivity']
        eps = np.ones_like(m.relative_permittivity)
        sig = np.zeros_like(m.conductivity)
        for z in range(g.size_z_material):
            for y in range(g.size_y):
                for x in range(g.size_x):
                    if z <= surf[y, x]:
                        eps[z, y, x] = e_r
                        sig[z, y, x] = s_r
        assert np.array_equal(eps, m.relative_permittivity)
        assert np.array_equal(sig, m.conductivity)


def test_two_layer_and_gradient_fills_match_loops():
    pc = _params()
    g = pc.geom_params
    dx, f = g.cell_size, pc.inc_beam_params.frequency
    args = dict(height_rms=0.7 * dx, corr_len=4 * dx, inter_layer_corr=0.6,
                mean_total_depth=9 * dx, top_layer_height=3 * dx, count=1, seed=11)
    surf = next(soilmodels.two_layer_surfaces(pc, **args))
    two = next(soilmodels.two_layer_gaussian(pc, 0.1, [0.3], 0.4, 0.2, **args))
    grad = next(soilmodels.gradient_gaussian(pc, 0.1, [0.3], 0.4, 0.2, **args))
    e_b, s_b = soil_eps_sigma(hallikainen_soil_model(0.3, 0.4, 0.2, f), f)
    e_t, s_t = soil_eps_sigma(hallikainen_soil_model(0.1, 0.4, 0.2, f), f)
    eps2, sig2 = np.ones_like(two.relative_permittivity), np.zeros_like(two.conductivity)
    epsg, sigg = np.ones_like(eps2), np.zeros_like(sig2)
    n_grad = 0
    for z in range(g.size_z_material):
        for y in range(g.size_y):
            for x in range(g.size_x):
                h0, h1 = int(surf[0, y, x]), int(surf[1, y, x])
                if z <= h0:
                    eps2[z, y, x], sig2[z, y, x] = e_b, s_b
                    epsg[z, y, x], sigg[z, y, x] = e_b, s_b
                elif z <= h1:
                    eps2[z, y, x], sig2[z, y, x] = e_t, s_t
                    sm = ((0.3 * (h1 - z)) + (0.1 * (z - h0))) / (h1 - h0)
                    epsg[z, y, x], sigg[z, y, x] = soil_eps_sigma(
                        hallikainen_soil_model(sm, 0.4, 0.2, f), f)
                    n_grad += 1
    assert n_grad > 0
    assert np.array_equal(eps2, two.relative_permittivity)
    assert np.array_equal(sig2, two.conductivity)
    assert np.array_equal(epsg, grad.relative_permittivity)
    assert np.array_equal(sigg, grad.conductivity)


# ---------------------------------------------------------------------------
# material builder oracle
# ---------------------------------------------------------------------------
def test_splitmix64_known_answers():
    # SplitMix64 seeded with 0: first two outputs
    assert int(mo.mix64(np.uint64(0))) == 0xE220A8397B1DCDAF
    assert int(mo.mix64(np.uint64(0x9E3779B97F4A7C15))) == 0x6E789E6AA1B965F4


def test_draw_statistics():
    cells = np.arange(1 << 18, dtype=np.uint64)
    u = mo.draw01(7, cells, 2)
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.003 and abs(u.var() - 1 / 12) < 0.002
    # draws of one cell and of neighbouring streams are not correlated
    v = mo.draw01(8, cells, 2)
    w = mo.draw01(7, cells, 3)
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.01 and abs(np.corrcoef(u, w)[0, 1]) < 0.01
    e, s = mo.random_soil(3, cells, 8.0, 0.5, 0.005, 0.02)
    assert abs(e.mean() - 8.0) < 0.005 and abs(e.std() - 0.5) < 0.005
    assert 0.005 <= s.min() and s.max() < 0.02


def test_oracle_cylinders_match_the_host_canopy_generator():
    pc = workloads.base_params(polarization='VV', theta=30.0, material_fraction=0.6)
    workloads.set_geom_params(pc, 2.0, 2.0, 1.5, 12.0, 0.05, 'wavelength')
    g = pc.geom_params
    lam = workloads.lam_of(pc)
    soil = g.size_z_material // 3
    host = synthetic.canopy_over_soil(g.size_x, g.size_y, g.size_z_material, g.cell_size, lam,
                                      soil, seed=4, density=4.0)
    cyl = device.canopy_cylinders(pc, soil, 4, 4.0)
    assert len(cyl) == host.metadata['cylinders']
    shape = host.relative_permittivity.shape
    eps = np.ones(shape)
    sig = np.zeros(shape)
    mo.paint_cylinders(eps, sig, cyl, g.cell_size)
    # above the soil surface the host model holds exactly the cylinders
    h = synthetic.surface_cells(g.size_x, g.size_y, g.cell_size, lam, soil, 4)
    above = np.arange(shape[0]).reshape(-1, 1, 1) > h[np.newaxis]
    assert (eps[above] != 1.0).sum() > 100
    assert np.array_equal(eps.astype(np.float32)[above], host.relative_permittivity[above])
    assert np.array_equal(sig.astype(np.float32)[above], host.conductivity[above])


def test_oracle_const_layer_equals_host_soil_model():
    pc = _params()
    g = pc.geom_params
    dx = g.cell_size
    args = dict(height_rms=0.8 * dx, corr_len=4 * dx, mean_depth=7 * dx, count=1, seed=5)
    spec = next(device.one_layer_specs(pc, [0.2], 0.4, 0.2, **args))
    host = next(soilmodels.one_layer_gaussian(pc, [0.2], 0.4, 0.2, **args))
    e, s = mo.eps_sigma(spec, host.relative_permittivity.shape, dx)
    assert np.array_equal(e, host.relative_permittivity)
    assert np.array_equal(s, host.conductivity)


def test_oracle_ca_cb_equals_compute_material_arrays():
    from sar_pixel_simulator_amd.materials import compute_material_arrays
    from sar_pixel_simulator_amd.constants import EPS0_2018
    pc = workloads.plain_params()
    m = workloads.plain_material(pc, seed=2)
    g = pc.geom_params
    Ca, Cb = compute_material_arrays(pc, m)
    oa, ob = mo.ca_cb(m.relative_permittivity, m.conductivity, g.size_z, g.timestep, g.cell_size,
                      EPS0_2018)
    assert np.array_equal(Ca, oa) and np.array_equal(Cb, ob)


# ---------------------------------------------------------------------------
# C-ABI struct layout
# ---------------------------------------------------------------------------
def test_material_structs_match_the_c_layout(tmp_path):
    src = tmp_path / 'layout.c'
    fields = {'FdtdMatLayer': [f for f, _ in native.FdtdMatLayer._fields_],
              'FdtdMatCylinder': [f for f, _ in native.FdtdMatCylinder._fields_],
              'FdtdMatSpec': [f for f, _ in native.FdtdMatSpec._fields_]}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fdtd_hip.h"', 'int main(void) {']
    for st, fs in fields.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines += ['return 0;', '}']
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'layout'
    subprocess.run(['gcc', '-std=c11', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)],
                   check=True)
    out = dict(line.rsplit(' ', 1) for line in
               subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split('\n')
               if line)
    for st in fields:
        cls = getattr(native, st)
        assert int(out[st]) == ctypes.sizeof(cls), st
        for f in fields[st]:
            assert int(out[f'{st}.{f}']) == getattr(cls, f).offset, (st, f)
