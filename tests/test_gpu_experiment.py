"""run_experiment end to end on the HIP backend, as the reference's release
example drives it (release_examples/cuda_example/runsim.py): one-layer
Hallikainen soil models, results files, post-processing (needs an MI355X).

The same models given as device MaterialSpec's (Ca, Cb formed on the device)
must give bit-identical phasors and sigma0.
"""
import glob
import os

import numpy as np
import pytest

from sar_pixel_simulator_amd import experiment, workloads
from sar_pixel_simulator_amd.materials import device, soilmodels

pytestmark = pytest.mark.gpu


def _params(project_dir):
    # the reference's parameters.json geometry: 3 x 3 x 1 lambda, material
    # fraction 1/3, 0.05 lambda cells (59 x 59 x 57), 3 periods
    pc = workloads.base_params(polarization='HH', theta=20.0, incident_duration=2.0,
                               illumination_width=1.0, simulation_duration=3.0,
                               material_fraction=1.0 / 3.0, project_dir=str(project_dir))
    workloads.set_geom_params(pc, 3.0, 3.0, 1.0, 3.0, 0.05, 'wavelength')
    return pc


def _soil_args(pc):
    g = pc.geom_params
    dx = g.cell_size
    return dict(moistures=[0.1, 0.3], sand_frac=0.4, clay_frac=0.2, height_rms=0.5 * dx,
                corr_len=4 * dx, mean_depth=g.size_z_material * dx - 6 * dx, count=1, seed=200)


def _results(d):
    out = {}
    for f in sorted(glob.glob(os.path.join(str(d), 'results', '*.npz'))):
        with np.load(f, allow_pickle=False) as z:
            out[os.path.basename(f)] = {k: z[k] for k in z.files if k != 'attrs'}
    return out


def test_run_experiment_host_models_and_device_specs_agree(tmp_path):
    print('case: run_experiment, 2 one-layer soil models, host then device specs', flush=True)
    a, b = tmp_path / 'host', tmp_path / 'device'
    pa = _params(a)
    experiment.run_experiment(pa, soilmodels.one_layer_gaussian(pa, **_soil_args(pa)))
    pb = _params(b)
    experiment.run_experiment(pb, device.one_layer_specs(pb, **_soil_args(pb)))
    ra, rb = _results(a), _results(b)
    assert sorted(ra) == ['mv_0.10_num_0.npz', 'mv_0.30_num_0.npz'], sorted(ra)
    assert sorted(ra) == sorted(rb)
    for name in ra:
        keys = [k for k in ra[name] if k.startswith('phasors') or k.startswith('scattering')]
        assert any(k.startswith('scattering') for k in keys), list(ra[name])
        for k in keys:
            assert np.array_equal(ra[name][k], rb[name][k]), (name, k)
    for name in ra:
        sig = [v for k, v in ra[name].items() if k.startswith('scattering')]
        assert all(np.all(np.isfinite(v)) for v in sig), name
