"""Parameters and derived geometry against hand-derived values of the
reference's rules (parameters.py:242-305, utils/rounding.py:4-15)."""
import json

import pytest

from sar_pixel_simulator_amd import constants, load_from_json, round_to_odd, set_geom_params
from sar_pixel_simulator_amd import parameters as P
from sar_pixel_simulator_amd import workloads

README_JSON = {
    "processing_type": "hip", "gpu_num": [0, 1, 2, 3, 4, 5, 6, 7], "threads_per_proc": 6,
    "datatype": "float32", "project_dir": "",
    "frequency_hz": 5.4e9, "polarization": "HH", "theta_angle_deg": 20.0,
    "phi_angle_deg": 0.0, "incident_duration": 10.0, "illumination_width": 12.0,
    "generate_new": False,
    "size_x": 16.0, "size_y": 16.0, "size_z_material": 3.0, "simulation_duration": 18.0,
    "cell_size": 0.021, "unit": "wavelength", "material_fraction": 0.7,
    "disp_inc_results": False, "disp_inc_sim_movie": False, "disp_debug": False,
    "disp_main_sim_movie": False, "disp_materials": False,
}


def test_round_to_odd():
    assert round_to_odd(59.0) == 59
    assert round_to_odd(60.0) == 61
    assert round_to_odd(60.9) == 61
    assert round_to_odd(202.857) == 203
    assert round_to_odd(0) == 1
    with pytest.raises(ValueError):
        round_to_odd(-1)


def test_float_floor_quirk():
    # 3.0 // 0.05 is 59.0, not 60 (0.05 is slightly above 1/20)
    assert 3.0 // 0.05 == 59.0
    pc = workloads.base_params(simulation_duration=12.0, material_fraction=1 / 3)
    set_geom_params(pc, 3.0, 3.0, 1.0, 12.0, 0.05, 'wavelength')
    g = pc.geom_params
    assert (g.size_x, g.size_y, g.size_z_material, g.size_z) == (59, 59, 19, 57)
    assert g.tfsf_z == 24                   # 19 + int(5.999...)
    assert (g.size_t, g.size_t_pulse, g.size_t_inc) == (480, 480, 480)  # 576 clamped


def test_readme_geometry(tmp_path):
    f = tmp_path / 'params.json'
    f.write_text(json.dumps(README_JSON))
    pc = load_from_json(str(f))
    g = pc.geom_params
    lam = constants.C0 / 5.4e9
    assert g.cell_size == pytest.approx(0.021 * lam, rel=1e-15)
    assert (g.size_x, g.size_y, g.size_z) == (761, 761, 203)
    assert (g.size_z_material, g.tfsf_z, g.size_pml, g.courant) == (142, 156, 10, 0.5)
    assert (g.size_t, g.size_t_inc, g.size_t_pulse) == (1714, 1372, 1143)
    assert g.timestep == pytest.approx(g.cell_size * 0.5 / constants.C0, rel=1e-15)
    assert pc.processing_params.gpu_num == [0, 1, 2, 3, 4, 5, 6, 7]
    assert pc.display_params.display_limit == 0.5
    assert pc.inc_beam_params.theta_angle == 20.0


def test_json_without_geometry(tmp_path):
    d = dict(README_JSON)
    for k in ('size_x', 'cell_size', 'gpu_num'):
        d.pop(k)
    f = tmp_path / 'p.json'
    f.write_text(json.dumps(d))
    pc = load_from_json(str(f))
    g = pc.geom_params
    assert (g.size_x, g.size_y, g.size_z) == (0, 0, 0)
    assert g.simulation_duration == 18.0
    assert pc.processing_params.gpu_num == 0


def test_missing_required_key(tmp_path):
    d = dict(README_JSON)
    d.pop('polarization')
    f = tmp_path / 'p.json'
    f.write_text(json.dumps(d))
    with pytest.raises(KeyError):
        load_from_json(str(f))


def test_metre_units_and_bad_unit():
    pc = workloads.base_params()
    lam = constants.C0 / 5.4e9
    dx = 0.05 * lam
    set_geom_params(pc, 60 * dx + dx / 2, 40 * dx + dx / 2, 20 * dx + dx / 2, 12.0, dx, 'metre')
    g = pc.geom_params
    assert g.cell_size == dx
    assert (g.size_x, g.size_y, g.size_z_material) == (61, 41, 20)
    with pytest.raises(ValueError):
        set_geom_params(pc, 1.0, 1.0, 1.0, 1.0, 0.1, 'furlong')


def test_filenames(tmp_path):
    pc = workloads.plain_params(project_dir=str(tmp_path))
    name = P.inc_filename(pc)
    assert name.endswith('incident/inc_nx81_ny81_nt480_theta20_phi0_HH.npz')
    assert P.results_filename(pc, 'soil_1').endswith('results/soil_1.npz')


def test_explicit_geometry_matches_rules():
    pc = workloads.plain_params()
    g = pc.geom_params
    assert (g.size_x, g.size_y, g.size_z, g.size_z_material, g.tfsf_z) == (81, 81, 81, 32, 37)
    assert (g.size_t, g.size_t_inc, g.size_t_pulse) == (480, 480, 480)


def test_constants_pinned():
    # the reference mixes scipy's (CODATA-2022 on scipy 1.15) and 2018 literals
    assert constants.EPS0_2018 == 8.8541878128e-12
    assert constants.ETA0 == 376.730313668
    assert constants.C0 == 299792458.0
    assert abs(constants.EPS0_SCIPY - 8.8541878188e-12) < 1e-21
