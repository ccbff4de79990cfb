"""Incident beam: temporal pulse, tapered-wave excitation and the vacuum
FDTD run that records the beam injected at the TFSF plane.

Follows fdtdsarsim/incidentbeam.py:25-204 (pulse, metadata validation) and
cpu/incident_sim.py:22-224 (the recording run).  The tapered-wave spectra,
their FFT synthesis and the 10.0 peak normalisation (incidentbeam.py:206-433)
run on the device (``synthesize_on_device``, csrc/fdtd_beam.hip); the
recording run executes on the HIP library in its INCIDENT mode, and the
recorded beam stays resident in HBM for the main simulations that share its
key.
"""
import json
import os.path
import time

import numpy as np

from . import _native as native
from .constants import C0
from .hip import common
from .parameters import ParameterCollection, inc_filename
from .utils import log

logger = log.getLog(__name__)

#: Attributes checked on load (incidentbeam.py:105-165).
META_KEYS = ('datatype', 'frequency', 'polarization', 'theta_angle',
             'phi_angle', 'duration', 'illumination_width', 'cell_size',
             'timestep', 'size_t_inc', 'size_t_pulse', 'size_x', 'size_y',
             'size_pml', 'courant')

# beams recorded in this process, keyed by cache file name
_RESIDENT = {}


class IncidentBeam:
    """Incident beam of one parameter set (incidentbeam.py:25-69)."""

    def __init__(self, params: ParameterCollection):
        self.params = params
        self.inc_beam_params = params.inc_beam_params
        self.geom_params = params.geom_params
        self.processing_params = params.processing_params
        self.display_params = params.display_params
        self.beam = None            # native.Beam resident on the device
        self.phasors = None         # dict of the four incident phasors

    # -- public API of the reference ------------------------------------------
    def setup(self):
        """Load the cached beam or generate it (incidentbeam.py:41-69)."""
        t0 = time.monotonic()
        key = inc_filename(self.params)
        if key in _RESIDENT and not self.inc_beam_params.generate_new:
            self.beam, self.phasors = _RESIDENT[key]
        elif not self.inc_beam_params.generate_new and self.inc_beam_exists():
            logger.info('Loading existing incident beam files')
            self.load_existing()
        else:
            logger.info('Generating new incident beam on the device')
            self.beam, self.phasors = synthesize_on_device(self.params)
            _RESIDENT[key] = (self.beam, self.phasors)
            self.save()
        logger.info(f'Incident beam setup complete ({time.monotonic() - t0:.3f}s)')

    def load_existing(self):
        """Read and validate the cached beam, then upload it to the device."""
        fields, self.phasors = self.read_file(inc_filename(self.params))
        g = self.geom_params
        self.beam = native.Beam(g.size_t_inc, g.size_y, g.size_x, *fields)
        _RESIDENT[inc_filename(self.params)] = (self.beam, self.phasors)

    def read_file(self, fname):
        """(ex, ey, hx, hy) and the incident phasors of a cached beam file,
        after the reference's checks: ValueError for a wrong shape or dtype,
        a metadata attribute that differs from the current parameters, or an
        incomplete incident run (incidentbeam.py:82-165)."""
        with np.load(fname, allow_pickle=False) as f:
            fields = [f[c] for c in ('ex', 'ey', 'hx', 'hy')]
            self.check_array_sizes(*fields)
            self.check_metadata(json.loads(str(f['attrs'])))
            phasors = {n: f['phasors/' + n] for n in native.INC_PHASOR_NAMES}
        return fields, phasors

    def check_array_sizes(self, inc_ex, inc_ey, inc_hx, inc_hy):
        """ValueError on a wrong shape or dtype (incidentbeam.py:82-103)."""
        g = self.geom_params
        want = (g.size_t_inc, g.size_y, g.size_x)
        for a in (inc_ex, inc_ey, inc_hx, inc_hy):
            if a.shape != want:
                raise ValueError(f'A component of the incident beam has the wrong '
                                 f'shape. Got {a.shape}, expected {want}.')
            if a.dtype != self.processing_params.datatype:
                raise ValueError('Incident beam has incorrect datatype. Please delete '
                                 'the incident beam file and re-run the simulation.')

    def check_metadata(self, meta):
        """ValueError when a stored attribute differs from the current
        parameters or the run did not complete (incidentbeam.py:105-165)."""
        want = beam_metadata(self.params)
        for key in META_KEYS:
            if meta.get(key) != want[key]:
                # the reference names a size_t_inc mismatch "(duration)"
                label = 'duration' if key == 'size_t_inc' else key
                raise ValueError(f'Incident beam metadata inconsistency ({label})')
        if meta.get('incident_complete') is not True:
            raise ValueError('Incident beam simulation was not completed')

    def inc_beam_exists(self):
        return os.path.isfile(inc_filename(self.params))

    def create_pulse(self):
        """Gaussian-tapered complex sinusoid and its spectrum
        (incidentbeam.py:174-204)."""
        g, b = self.geom_params, self.inc_beam_params
        dt, nt_inc, nt_pulse = g.timestep, g.size_t_inc, g.size_t_pulse
        self.time_vals = np.linspace(0.0, dt * (nt_inc - 1), num=nt_inc, endpoint=True)
        sig = b.duration / b.frequency / (2.0 * np.sqrt(2.0 * np.log(10.0)))
        om = 2.0 * np.pi * b.frequency
        self.pulse = 1.0 * np.exp(-(self.time_vals - (dt * nt_pulse / 2)) ** 2
                                  / (sig ** 2)) * np.exp(1j * om * self.time_vals)
        self.pulse_fft = np.fft.fft(self.pulse)
        self.pulse_freqs = np.fft.fftfreq(self.pulse.size, dt)

    # -- storage ---------------------------------------------------------------
    def save(self):
        """Write the beam, its phasors and metadata (the incident HDF5 layout,
        cpu/incident_sim.py:95-98, :167-171, :195-224) as .npz."""
        g = self.geom_params
        fname = inc_filename(self.params)
        write_beam_file(fname, [self.beam.component(i) for i in range(4)], self.phasors,
                        self.params)
        logger.info(f'Incident beam written to {fname} ({g.size_t_inc} steps)')


def write_beam_file(fname, fields, phasors, params: ParameterCollection, complete=True):
    """Cached-beam file: ex, ey, hx, hy (nt_inc, ny, nx), the four incident
    phasors and the metadata the reader checks; written to a temporary name
    and renamed, so a reader never sees a partial file."""
    arrays = {c: a for c, a in zip(('ex', 'ey', 'hx', 'hy'), fields)}
    for n in native.INC_PHASOR_NAMES:
        arrays['phasors/' + n] = phasors[n]
    meta = beam_metadata(params)
    meta['incident_complete'] = bool(complete)
    arrays['attrs'] = np.array(json.dumps(meta))
    tmp = fname + '.tmp.npz'
    np.savez(tmp, **arrays)
    os.replace(tmp, fname)


def beam_metadata(params: ParameterCollection):
    p, b, g = params.processing_params, params.inc_beam_params, params.geom_params
    return {'datatype': p.datatype, 'frequency': b.frequency,
            'polarization': b.polarization, 'theta_angle': b.theta_angle,
            'phi_angle': b.phi_angle, 'duration': b.duration,
            'illumination_width': b.illumination_width, 'cell_size': g.cell_size,
            'timestep': g.timestep, 'size_t_inc': g.size_t_inc,
            'size_t_pulse': g.size_t_pulse, 'size_x': g.size_x, 'size_y': g.size_y,
            'size_pml': g.size_pml, 'courant': g.courant}


def run_incident(params: ParameterCollection, pulse_ex, pulse_ey, device=None):
    """Run the vacuum incident simulation on the device and record the beam.

    Returns (native.Beam, incident phasor dict).  The excitation's imaginary
    parts drive a hard source on [1:-1, 0:-1] of the feed-in plane for the
    first size_t_pulse steps (cpu/incident_sim.py:109-111, :141-144).
    """
    common.require_float32(params)
    g = params.geom_params
    native.init(params_device(params) if device is None else device)
    geom = common.incident_geom(params)
    sim = native.Sim(geom, common.coeffs_for(params))
    try:
        sim.set_uniform_material(1.0, common.incident_cb(params))
        n_src = geom.nt_inc
        src_ex = np.ascontiguousarray(np.imag(pulse_ex[:n_src]), dtype=np.float32)
        src_ey = np.ascontiguousarray(np.imag(pulse_ey[:n_src]), dtype=np.float32)
        sim.set_source(src_ex, src_ey)
        beam = native.Beam(g.size_t_inc, g.size_y, g.size_x)
        sim.attach_beam(beam)
        sim.run(g.size_t_inc)
        sim.sync()
        phasors = sim.phasors()
    finally:
        sim.destroy()
    return beam, phasors


def beam_spectrum(params: ParameterCollection):
    """Host-side inputs of the device beam pipeline: the temporal pulse's
    FFT and bin wavenumbers (incidentbeam.py:174-204, :255-257) and the
    Gaussian footprint radius (:231-233)."""
    ib = IncidentBeam(params)
    ib.create_pulse()
    b = params.inc_beam_params
    width = b.illumination_width * C0 / b.frequency
    gauss_radius = width / (2.0 * np.sqrt(2.0 * np.log(10.0)))
    all_k = 2 * np.pi * ib.pulse_freqs / C0
    return ib.pulse_fft, all_k, gauss_radius


def synthesize_on_device(params: ParameterCollection, device=None):
    """Excitation synthesis and incident simulation both on the device
    (libfdtd_hip fdtd_beam_synthesize).  Returns (native.Beam, incident
    phasor dict); the same beam as the host synthesis of the excitation
    (incidentbeam.py:206-313) followed by ``run_incident``, up to FFT
    rounding."""
    common.require_float32(params)
    native.init(params_device(params) if device is None else device)
    b = params.inc_beam_params
    pulse_fft, all_k, radius = beam_spectrum(params)
    return native.beam_synthesize(common.incident_geom(params), common.coeffs_for(params),
                                  b.polarization, b.theta_angle, b.phi_angle, radius,
                                  pulse_fft, all_k, common.incident_cb(params))


def params_device(params: ParameterCollection):
    d = params.processing_params.gpu_num
    return d[0] if isinstance(d, (list, tuple)) else int(d)


def make_resident(params: ParameterCollection, beam, phasors):
    """Register a device-resident beam (and its incident phasors) as this
    process's beam for ``params``' key, as setup() does after generating it."""
    _RESIDENT[inc_filename(params)] = (beam, phasors)


def forget_resident(params: ParameterCollection):
    """Unregister the beam of ``params``' key without releasing it (its owner
    releases it)."""
    _RESIDENT.pop(inc_filename(params), None)


def clear_resident():
    """Drop this process's device-resident beams."""
    for beam, _ in _RESIDENT.values():
        beam.release()
    _RESIDENT.clear()
