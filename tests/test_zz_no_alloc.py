"""fdtd_sim_near2far_noalloc (the near-to-far entry point that allocates
nothing: its work area is a preallocated per-device array) against
fdtd_sim_near2far, both called through the C-ABI: bit-identical far fields
on two calls in a row, the second on a smaller simulation with fewer
angles, and sizes beyond the work area refused before any launch."""
import ctypes

import numpy as np
import pytest

import helpers as H
from sar_pixel_simulator_amd import _native as native
from sar_pixel_simulator_amd.hip import common

pytestmark = pytest.mark.gpu

_DP = ctypes.POINTER(ctypes.c_double)
FDTD_EINVAL = -1


def last_error():
    return native.lib().fdtd_last_error().decode()


def _call(fn, sim, thetas, phi, r, freq):
    th = np.ascontiguousarray(thetas, dtype=np.float64)
    out = np.full((4, th.size), np.nan, np.complex128)
    rc = fn(sim.handle, th.ctypes.data_as(_DP), th.size, float(phi), float(r), float(freq),
            out.view(np.float64).ctypes.data_as(_DP))
    return rc, out


def _sim(nx, ny, nz, nz_mat, steps, seed):
    pc, Ca, Cb = H.soil_case(nx=nx, ny=ny, nz=nz, nz_mat=nz_mat, steps=steps, seed=seed)
    native.init(0)
    sim = native.Sim(common.main_geom(pc), common.coeffs_for(pc))
    g = pc.geom_params
    beam = native.Beam(g.size_t_inc, g.size_y, g.size_x, *H.synthetic_beam(pc))
    sim.attach_beam(beam)
    sim.set_materials(Ca, Cb)
    sim.run(steps)
    sim.sync()
    return pc, sim, beam


def test_noalloc_near2far_equals_allocating_entry_point():
    L = native.lib()
    cases = [((45, 39, 37, 15, 40, 5), np.linspace(-80.0, 80.0, 161), 0.0),
             ((33, 31, 35, 14, 30, 2), np.linspace(-40.0, 50.0, 37), 30.0)]
    for dims, th, phi in cases:
        print('case: near2far_noalloc', dims, len(th), 'angles, phi', phi, flush=True)
        pc, sim, beam = _sim(*dims)
        f = pc.inc_beam_params.frequency
        rc0, want = _call(L.fdtd_sim_near2far, sim, th, phi, 2000.0, f)
        assert rc0 == 0, last_error()
        rc1, got = _call(L.fdtd_sim_near2far_noalloc, sim, th, phi, 2000.0, f)
        assert rc1 == 0, last_error()
        assert np.isfinite(want).all() and np.abs(want).max() > 0
        assert np.array_equal(got, want), np.abs(got - want).max()
        sim.destroy()
        beam.release()


def test_noalloc_near2far_rejects_sizes_beyond_work_area():
    L = native.lib()
    print('case: near2far_noalloc limits', flush=True)
    pc, sim, beam = _sim(33, 31, 35, 14, 10, 1)
    f = pc.inc_beam_params.frequency
    rc, out = _call(L.fdtd_sim_near2far_noalloc, sim, np.linspace(-90.0, 90.0, 182), 0.0, 2000.0, f)
    assert rc == FDTD_EINVAL and 'ntheta 182' in last_error()
    assert np.isnan(out).all()
    rc, out = _call(L.fdtd_sim_near2far_noalloc, sim, np.zeros(0), 0.0, 2000.0, f)
    assert rc == FDTD_EINVAL
    # the largest accepted angle count still works after the refusals
    th = np.linspace(-90.0, 90.0, 181)
    rc, got = _call(L.fdtd_sim_near2far_noalloc, sim, th, 10.0, 2000.0, f)
    assert rc == 0, last_error()
    rc, want = _call(L.fdtd_sim_near2far, sim, th, 10.0, 2000.0, f)
    assert rc == 0 and np.array_equal(got, want)
    sim.destroy()
    beam.release()
