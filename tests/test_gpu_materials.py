"""Material models built on the device (fdtd_sim_build_materials) against
the host (needs an MI355X).

Tolerances:
  the reference's soil generators (one layer, two layers, moisture
      gradient): Ca, Cb bit-identical to compute_material_arrays of the host
      models (same IEEE double and float32 operations, no contraction);
  random soil, voxel inclusions, cylinders: bit-identical to the oracle
      restatement except where the Box-Muller log/cos of the device and of
      numpy round differently: at most 1e-5 of the cells, each within 2
      float32 ulp of Ca and Cb;
  a simulation stepped on device-built materials: fields bit-identical to
      the same simulation on host-uploaded Ca, Cb.
"""
import os
import sys

import numpy as np
import pytest

import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, 'oracle'))
import materials_oracle as mo  # noqa: E402

from sar_pixel_simulator_amd import _native as native, workloads  # noqa: E402
from sar_pixel_simulator_amd.constants import EPS0_2018  # noqa: E402
from sar_pixel_simulator_amd.hip import common  # noqa: E402
from sar_pixel_simulator_amd.materials import compute_material_arrays, device, soilmodels  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _sim(pc):
    native.init(0)
    return native.Sim(common.main_geom(pc), common.coeffs_for(pc))


def _built(pc, spec):
    sim = _sim(pc)
    sim.build_materials(*device.to_c(spec, pc))
    Ca, Cb = sim.materials()
    sim.destroy()
    return Ca, Cb


def _params(nx=45, ny=39, nz_mat=22, nz=46):
    return workloads.small_params(nx=nx, ny=ny, nz=nz, nz_mat=nz_mat)


def _ulps(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    return np.abs(ai - bi)


@pytest.mark.parametrize('kind', ['one_layer', 'two_layer', 'gradient'])
def test_reference_soil_generators_bitwise(kind):
    pc = _params()
    g = pc.geom_params
    dx = g.cell_size
    say('case: device-built', kind, g.size_x, g.size_y, g.size_z)
    if kind == 'one_layer':
        args = dict(height_rms=0.8 * dx, corr_len=4 * dx, mean_depth=12 * dx, count=2, seed=5)
        specs = device.one_layer_specs(pc, [0.1, 0.35], 0.4, 0.2, **args)
        hosts = soilmodels.one_layer_gaussian(pc, [0.1, 0.35], 0.4, 0.2, **args)
    else:
        args = dict(height_rms=0.7 * dx, corr_len=4 * dx, inter_layer_corr=0.6,
                    mean_total_depth=14 * dx, top_layer_height=5 * dx, count=1, seed=11)
        if kind == 'two_layer':
            specs = device.two_layer_specs(pc, 0.1, [0.3, 0.4], 0.4, 0.2, **args)
            hosts = soilmodels.two_layer_gaussian(pc, 0.1, [0.3, 0.4], 0.4, 0.2, **args)
        else:
            specs = device.gradient_specs(pc, 0.05, [0.3, 0.4], 0.4, 0.2, **args)
            hosts = soilmodels.gradient_gaussian(pc, 0.05, [0.3, 0.4], 0.4, 0.2, **args)
    n = 0
    for spec, host in zip(specs, hosts):
        assert spec.name == host.name
        Ca, Cb = compute_material_arrays(pc, host)
        dCa, dCb = _built(pc, spec)
        assert np.array_equal(dCa, Ca) and np.array_equal(dCb, Cb), spec.name
        n += 1
    assert n >= 2


@pytest.mark.parametrize('kind', ['rough_soil', 'snow_over_soil', 'canopy_over_soil'])
def test_device_drawn_workloads_match_oracle(kind):
    pc = workloads.base_params(polarization='VV', theta=30.0, material_fraction=0.6)
    workloads.set_geom_params(pc, 2.0, 2.0, 1.5, 12.0, 0.04, 'wavelength')
    g = pc.geom_params
    nzm = g.size_z_material
    say('case: device-drawn', kind, g.size_x, g.size_y, g.size_z)
    if kind == 'rough_soil':
        spec = device.rough_soil_spec(pc, nzm // 2, seed=9)
    elif kind == 'snow_over_soil':
        spec = device.snow_over_soil_spec(pc, nzm // 3, nzm // 3, seed=9)
    else:
        spec = device.canopy_over_soil_spec(pc, nzm // 3, seed=9, density=5.0)
        assert len(spec.cylinders) > 5
    e, s = mo.eps_sigma(spec, (nzm, g.size_y, g.size_x), g.cell_size)
    Ca, Cb = mo.ca_cb(e, s, g.size_z, g.timestep, g.cell_size, EPS0_2018)
    dCa, dCb = _built(pc, spec)
    ua, ub = _ulps(dCa, Ca), _ulps(dCb, Cb)
    diff = (ua > 0) | (ub > 0)
    say('cells differing from the oracle:', int(diff.sum()), 'of', diff.size,
        'max ulp', int(max(ua.max(), ub.max())))
    assert diff.sum() <= max(1, 1e-5 * diff.size)
    assert ua.max() <= 2 and ub.max() <= 2
    # the draws are those of the stated statistics
    soil = (Ca < 1.0) & (np.arange(g.size_z).reshape(-1, 1, 1) < nzm)
    assert soil.sum() > 1000


def test_stepping_on_device_built_materials_equals_host_upload():
    pc, _, _ = H.soil_case(nx=33, ny=31, nz=35, nz_mat=14, steps=40, seed=3)
    g = pc.geom_params
    spec = device.canopy_over_soil_spec(pc, 6, seed=2, density=30.0)
    beam = H.synthetic_beam(pc)
    say('case: stepping on device-built canopy', g.size_x, g.size_y, g.size_z, len(spec.cylinders))
    a = _sim(pc)
    a.build_materials(*device.to_c(spec, pc))
    Ca, Cb = a.materials()
    b = _sim(pc)
    b.set_materials(Ca, Cb)
    bm = native.Beam(g.size_t_inc, g.size_y, g.size_x, *beam)
    for s in (a, b):
        s.attach_beam(bm)
        s.run(40)
        s.sync()
    for n in native.FIELD_NAMES:
        assert np.array_equal(a.field(n), b.field(n)), n
    for s in (a, b):
        s.destroy()
    bm.release()


def test_invalid_specs_are_rejected():
    pc = _params()
    g = pc.geom_params
    sim = _sim(pc)
    surf = np.zeros((1, g.size_y, g.size_x), np.int32)
    bad = [device.MaterialSpec(g.size_z + 1, [device.Layer.const(4.0, 0.01)], surf),
           device.MaterialSpec(g.size_z_material,
                               [device.Layer.gradient(0.1, 0.2, 0.4, 0.2, 5.4e9)], surf)]
    for spec in bad:
        with pytest.raises(native.FdtdError):
            sim.build_materials(*device.to_c(spec, pc))
    spec = device.MaterialSpec(g.size_z_material, vox_z1=g.size_z_material + 1, vox_frac=0.5)
    with pytest.raises(native.FdtdError):
        sim.build_materials(*device.to_c(spec, pc))
    sim.destroy()
