"""Results post-processing: far field, Poynting powers and the scattering
coefficient sigma0 (fdtdsarsim/processresults.py:15-271).

The far-field transform runs on the device (near2far.FarFieldTransform);
powers and normalisation are FP64 numpy with scipy's Simpson rule, as in the
reference.  Results files are the .npz written by MainSimHip; each stage is
skipped when its flag is already present, as the reference does with its
HDF5 attributes.
"""
import glob
import json
import os.path

import numpy as np
import scipy.integrate

from .near2far import FarFieldTransform

RADIUS = 2.0e3
THETAS = np.linspace(-80.0, 80.0, num=161, endpoint=True)


def process_all_results(project_dir: str, debug_plots=False):
    for fname in sorted(glob.glob(os.path.join(project_dir, 'results', '*.npz'))):
        process_results(fname, debug_plots)


def _load(fname):
    with np.load(fname, allow_pickle=False) as f:
        arrays = {k: f[k] for k in f.files}
    attrs = json.loads(str(arrays.pop('attrs')))
    return arrays, attrs


def _store(fname, arrays, attrs):
    arrays = dict(arrays)
    arrays['attrs'] = np.array(json.dumps(attrs))
    tmp = fname + '.tmp.npz'
    np.savez(tmp, **arrays)
    os.replace(tmp, fname)


def process_results(filename: str, debug_plots=False):
    """Add far field, powers and sigma0 to a results file
    (processresults.py:25-126)."""
    arrays, attrs = _load(filename)
    par, ph_at = attrs['parameters'], attrs['phasors']
    dx = par['cell_size']
    nxp, nyp, nzp = ph_at['nx_phasor'], ph_at['ny_phasor'], ph_at['nz_phasor']
    phasors = {k.split('/', 1)[1]: v for k, v in arrays.items() if k.startswith('phasors/')}
    ff_at = attrs.setdefault('farfield', {})
    if not ff_at.get('farfields_computed'):
        phis = np.asarray([par['phi_angle']])
        ff = FarFieldTransform()(phasors, nxp, nyp, nzp, dx, RADIUS, THETAS, phis,
                                 par['frequency'])
        for k, name in (('E_theta', 'e_theta'), ('E_phi', 'e_phi'),
                        ('H_theta', 'h_theta'), ('H_phi', 'h_phi')):
            arrays['farfield/' + name] = ff[k]
        ph, pv = farfield_power(ff['E_theta'], ff['E_phi'], ff['H_theta'], ff['H_phi'])
        arrays['farfield/out_pwr_far_h'] = ph
        arrays['farfield/out_pwr_far_v'] = pv
        ff_at.update(theta_angles=THETAS.tolist(), phi_angles=phis.tolist(),
                     radius=RADIUS, farfields_computed=True)
    nf_at = attrs.setdefault('near_field_power', {})
    if not nf_at.get('powers_computed'):
        nf_at.update({k: float(v) for k, v in
                      near_field_power(phasors, dx, par['size_z_material']).items()})
        nf_at['powers_computed'] = True
    sc_at = attrs.setdefault('scattering_coefficient', {})
    if not sc_at.get('scattering_computed'):
        h, v = scattering_coefficients(arrays['farfield/out_pwr_far_h'],
                                       arrays['farfield/out_pwr_far_v'],
                                       nf_at['incident'], nxp, nyp, dx)
        arrays['scattering_coefficient/scattering_coeff_h'] = h
        arrays['scattering_coefficient/scattering_coeff_v'] = v
        sc_at['scattering_computed'] = True
    _store(filename, arrays, attrs)


def scattering_coefficients(pwr_h, pwr_v, input_power, nxp, nyp, dx, radius=RADIUS):
    """sigma = P_far / P_inc * 4 pi r^2 / (nx_ph ny_ph dx^2)
    (processresults.py:104-121)."""
    scale = 4.0 * np.pi * (radius ** 2) / (nxp * nyp * dx * dx)
    return (np.asarray(pwr_h) / input_power) * scale, (np.asarray(pwr_v) / input_power) * scale


def farfield_power(e_theta, e_phi, h_theta, h_phi):
    """(horizontal, vertical) scattered power (processresults.py:260-271)."""
    return 0.5 * np.real(-e_phi * np.conj(h_theta)), 0.5 * np.real(e_theta * np.conj(h_phi))


def _face_power(E, H, normal, dx):
    pxy = np.dot(np.real(0.5 * np.cross(E, np.conj(H))), np.asarray(normal, dtype=float))
    return pxy


def _integrate(p, dx):
    return scipy.integrate.simpson(scipy.integrate.simpson(p, dx=dx, axis=1), dx=dx, axis=0)


def incident_power(phasors, dx):
    z = np.zeros_like(phasors['ey_inc_phasor'])
    E = np.stack([phasors['ex_inc_phasor'], phasors['ey_inc_phasor'], z], axis=-1)
    H = np.stack([phasors['hx_inc_phasor'], phasors['hy_inc_phasor'], np.zeros_like(z)], axis=-1)
    return _integrate(_face_power(E, H, [0.0, 0.0, -1.0], dx), dx)


def near_field_power(phasors, dx, size_z_material):
    """Poynting power through the incident plane and the six phasor faces,
    split at the material height on the side faces
    (processresults.py:128-257)."""
    power = {'incident': incident_power(phasors, dx)}

    def side(e1, e2, h1, h2, slot, normal, name):
        z = np.zeros_like(phasors[e1])
        comps_e = [z, z, z]
        comps_h = [z, z, z]
        comps_e[slot[0]], comps_e[slot[1]] = phasors[e1], phasors[e2]
        comps_h[slot[0]], comps_h[slot[1]] = phasors[h1], phasors[h2]
        p = _face_power(np.stack(comps_e, axis=-1), np.stack(comps_h, axis=-1), normal, dx)
        power[name + '_bottom'] = _integrate(p[0:size_z_material, :], dx)
        power[name + '_top'] = _integrate(p[size_z_material:-1, :], dx)

    side('ey_ph_l', 'ez_ph_l', 'hy_ph_l', 'hz_ph_l', (1, 2), [-1.0, 0.0, 0.0], 'left')
    side('ey_ph_r', 'ez_ph_r', 'hy_ph_r', 'hz_ph_r', (1, 2), [1.0, 0.0, 0.0], 'right')
    side('ex_ph_f', 'ez_ph_f', 'hx_ph_f', 'hz_ph_f', (0, 2), [0.0, -1.0, 0.0], 'front')
    side('ex_ph_b', 'ez_ph_b', 'hx_ph_b', 'hz_ph_b', (0, 2), [0.0, 1.0, 0.0], 'back')

    for face, normal in (('d', [0.0, 0.0, -1.0]), ('u', [0.0, 0.0, 1.0])):
        z = np.zeros_like(phasors['ex_ph_' + face])
        E = np.stack([phasors['ex_ph_' + face], phasors['ey_ph_' + face], z], axis=-1)
        H = np.stack([phasors['hx_ph_' + face], phasors['hy_ph_' + face], np.zeros_like(z)],
                     axis=-1)
        power['down' if face == 'd' else 'up'] = _integrate(_face_power(E, H, normal, dx), dx)

    power['top_total'] = power['up'] + power['left_top'] + power['right_top'] \
        + power['front_top'] + power['back_top']
    power['bottom_total'] = power['down'] + power['left_bottom'] + power['right_bottom'] \
        + power['front_bottom'] + power['back_bottom']
    return power


def backscatter(sigma_h, sigma_v, theta_inc_deg, thetas=THETAS):
    """sigma0 at theta = -theta_inc: (h, v) (release_examples/cuda_example/
    gather.py:39-53 reads e_phi for HH and e_theta for VV at that angle)."""
    idx = int(np.squeeze(np.where(np.asarray(thetas) == -theta_inc_deg)))
    return float(np.asarray(sigma_h)[idx]), float(np.asarray(sigma_v)[idx])
