"""HIP library vs the oracle on the same seeded inputs (needs an MI355X).

Tolerances:
  fields, CPML psi, recorded beam, phasors in 'reference' accumulation:
      bit-identical (same operation order, -ffp-contract=off on both sides)
  phasors in 'compensated' accumulation: relative L2 <= 1e-5 against the
      oracle's complex64 running sum; the float-float sum is the more
      accurate one (test_compensated_dft_is_closer_to_exact_sum)
  far field (FP64, device vs scipy Simpson): relative <= 1e-9
"""
import numpy as np
import pytest

import helpers as H
import oracle
from sar_pixel_simulator_amd import _native as native
from sar_pixel_simulator_amd.hip import common

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_sim(pc, Ca=None, Cb=None, beam_planes=None, launch=None, fused=False):
    """A simulation on device 0; the default stepper is the library's
    default (split), the others are 1 (fused strips) and 2 (row-parallel)."""
    native.init(0)
    sim = native.Sim(common.main_geom(pc), common.coeffs_for(pc))
    sim.set_stepper(fused)
    if launch:
        sim.set_launch(*launch)
    g = pc.geom_params
    if beam_planes is not None:
        b = native.Beam(g.size_t_inc, g.size_y, g.size_x, *beam_planes)
        sim.attach_beam(b)
    if Ca is not None:
        sim.set_materials(Ca, Cb)
    return sim


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        i = tuple(bad[0])
        raise AssertionError(f'{what}: {len(bad)} of {a.size} differ; first {i}: '
                             f'{a[i]!r} vs {b[i]!r}')


@pytest.fixture(scope='module')
def soil60():
    pc, Ca, Cb = H.soil_case(steps=60, dft_accumulation='reference')
    beam = H.synthetic_beam(pc)
    ref = H.oracle_main(pc, Ca, Cb, beam)
    return pc, Ca, Cb, beam, ref


@pytest.mark.parametrize('fused', [False, True, 2, 3], ids=['split', 'fused', 'rows', 'halo'])
def test_fields_psi_phasors_bitwise(soil60, fused):
    pc, Ca, Cb, beam, ref = soil60
    say('case: soil 33x31x35, 60 steps, reference DFT, fused', fused)
    sim = make_sim(pc, Ca, Cb, beam, fused=fused)
    sim.run(60)
    sim.sync()
    assert sim.step_count() == 60
    for n in native.FIELD_NAMES:
        assert_same(sim.field(n), ref.fields[n], n)
    for i, n in enumerate(native.PSI_NAMES):
        assert_same(sim.psi(i), ref.psi[n], n)
    ph = sim.phasors()
    oph = ref.phasors()
    for n in native.PHASOR_NAMES:
        assert_same(ph[n], oph[n], n)
    assert np.abs(oph['ex_ph_u']).max() > 0
    sim.destroy()


# (ty, zchunk, xvec): split stepper with 1, 2 and 4 x cells per lane
# (33 cells: one partly idle tile); fused stepper with 2, 4 and 8 rows per
# wavefront, z ranges that start mid-column (recomputed plane below) and
# strips whose first row is recomputed
@pytest.mark.parametrize('fused,launch', [(False, (1, 0, 1, 1, 1)), (False, (2, 0, 2, 1, 1)),
                                          (False, (4, 0, 4, 1, 1)), (False, (4, 7, 4, 1, 1)),
                                          (False, (1, 5, 2, 1, 1)), (False, (2, 16, 4, 1, 1)),
                                          (False, (2, 3, 1, 1, 1)), (False, (1, 1, 4, 1, 1)),
                                          (True, (2, 0, 1, 1, 1)), (True, (2, 0, 4, 1, 1)),
                                          (True, (8, 7, 1, 1, 1)), (True, (4, 5, 1, 1, 1)),
                                          (True, (2, 1, 4, 1, 1)), (True, (8, 16, 0, 1, 1)),
                                          (True, (4, 3, 2, 1, 1)), (True, (4, 2, 0, 1, 1)),
                                          (False, (1, 1, 1, 4, 0)), (False, (1, 2, 2, 4, 0)),
                                          (False, (2, 4, 1, 2, 0)), (False, (1, 16, 1, 2, 1)),
                                          (False, (1, 0, 4, 4, 1)), (False, (2, 3, 2, 2, 1)),
                                          (True, (4, 5, 1, 1, 0)), (True, (2, 0, 4, 1, 0)),
                                          (False, (1, 4, 1, 4, 2)), (False, (2, 3, 1, 2, 2)),
                                          (False, (1, 0, 2, 4, 2)), (False, (4, 1, 1, 1, 2)),
                                          (2, (3, 0, 1, 1, 0)), (2, (7, 5, 1, 1, 0)),
                                          (2, (3, 1, 2, 1, 1)), (2, (7, 8, 2, 1, 0)),
                                          (2, (3, 16, 0, 1, 0)),
                                          (3, (1, 0, 1, 4, 0)), (3, (1, 2, 1, 4, 0)),
                                          (3, (2, 3, 2, 2, 1)), (3, (4, 1, 1, 1, 0)),
                                          (3, (1, 5, 2, 4, 0)), (3, (1, 7, 1, 2, 2)),
                                          (3, (2, 16, 0, 0, 0)), (3, (0, 8, 0, 0, 0)),
                                          # fine bands of 1, 2, 4, 8 row blocks per XCD (31 rows:
                                          # the last panel partly empty)
                                          (False, (1, 2, 1, 4, 3)), (False, (2, 3, 1, 2, 4)),
                                          (False, (4, 1, 1, 1, 6)), (3, (2, 8, 1, 2, 4)),
                                          (3, (1, 5, 1, 4, 3)), (3, (2, 3, 2, 2, 5)),
                                          (3, (1, 7, 1, 1, 6)),
                                          # the tuned halo shapes of the two largest workloads
                                          (3, (1, 8, 1, 2, 5)), (3, (1, 16, 1, 4, 4))])
def test_launch_shapes_bitwise(soil60, fused, launch):
    pc, Ca, Cb, beam, ref = soil60
    say('case: fused', fused, 'launch', launch)
    sim = make_sim(pc, Ca, Cb, beam, launch=launch, fused=fused)
    sim.run(60)
    sim.sync()
    for n in native.FIELD_NAMES:
        assert_same(sim.field(n), ref.fields[n], f'{n} {launch}')
    ph, oph = sim.phasors(), ref.phasors()
    for n in native.PHASOR_NAMES:
        assert_same(ph[n], oph[n], f'{n} {launch}')
    sim.destroy()


STEPPERS = pytest.mark.parametrize('fused', [False, True, 2, 3], ids=['split', 'fused', 'rows', 'halo'])


@STEPPERS
def test_compensated_phasors_close(soil60, fused):
    pc, Ca, Cb, beam, ref = soil60
    say('case: compensated DFT, stepper', fused)
    pc.processing_params.dft_accumulation = 'compensated'
    try:
        sim = make_sim(pc, Ca, Cb, beam, fused=fused)
    finally:
        pc.processing_params.dft_accumulation = 'reference'
    sim.run(60)
    sim.sync()
    ph, oph = sim.phasors(), ref.phasors()
    for n in native.PHASOR_NAMES:
        assert H.rel_l2(ph[n], oph[n]) <= 1e-5, n
    sim.destroy()


@pytest.fixture(scope='module')
def taps139():
    """139 x 27 x 29, 150 steps: the right x face (x = 128) is a wavefront's
    first lane at 1 and 2 cells per lane, so its hy, hz at x - 1 come from
    memory in the split stepper's copies; the oracle's float32 phasors and
    the exact FP64 running sums of six face components."""
    pc, Ca, Cb = H.soil_case(nx=139, ny=27, nz=29, nz_mat=11, steps=150, seed=13,
                             dft_accumulation='reference')
    beam = H.synthetic_beam(pc)
    ref = H.oracle_main(pc, Ca, Cb, beam)
    exact = _exact_dft(pc, Ca, Cb, beam, 150)
    return pc, Ca, Cb, beam, ref.phasors(), exact


@pytest.mark.parametrize('fused,launch', [(False, None), (False, (1, 2, 1, 4, 0)),
                                          (False, (1, 3, 2, 2, 0)), (False, (2, 5, 4, 2, 1)),
                                          (False, (0, 3, 1, -1, 0)),
                                          (True, None), (2, None), (3, None), (3, (1, 3, 2, 2, 0))],
                         ids=['split-auto', 'split-1-cell', 'split-2-cells', 'split-4-cells-xcd',
                              'split-flat', 'fused', 'rows', 'halo', 'halo-2-cells'])
def test_compensated_x_faces_close_and_closer_to_exact(taps139, fused, launch):
    """The shipped default (split stepper, compensated DFT) at phasor level
    where the x-face taps sit on wavefront edges: every phasor within 1e-5
    relative L2 of the oracle's complex64 running sum, and the x- and z-face
    components at least as close to the exact FP64 sum as the oracle's."""
    pc, Ca, Cb, beam, oph, exact = taps139
    say('case: compensated DFT, 139 x 27 x 29, 150 steps, stepper', fused, 'launch', launch)
    pc.processing_params.dft_accumulation = 'compensated'
    try:
        sim = make_sim(pc, Ca, Cb, beam, launch=launch, fused=fused)
    finally:
        pc.processing_params.dft_accumulation = 'reference'
    sim.run(150)
    sim.sync()
    ph = sim.phasors()
    sim.destroy()
    for n in native.PHASOR_NAMES:
        assert H.rel_l2(ph[n], oph[n]) <= 1e-5, n
    for n, want in exact.items():
        e_comp, e_ref = H.rel_l2(ph[n], want), H.rel_l2(oph[n], want)
        assert e_comp <= e_ref and e_comp < 1e-6, (n, e_comp, e_ref)


@STEPPERS
def test_steps_past_end_are_noops(soil60, fused):
    pc, Ca, Cb, beam, ref = soil60
    sim = make_sim(pc, Ca, Cb, beam, fused=fused)
    sim.run(75)
    sim.sync()
    assert sim.step_count() == 60
    assert_same(sim.field('ey'), ref.fields['ey'], 'ey after extra steps')
    sim.destroy()


@pytest.mark.parametrize('launch', [(1, 2, 1, 4, 0), (2, 3, 1, 1, 1), (1, 2, 1, 4, 2), (1, 2, 1, 4, 4)],
                         ids=['direct-grid', 'xcd-order', 'banded', 'fine-bands'])
def test_slabs_bitwise(soil60, launch):
    """H and E launched slab by slab (FDTD_SLAB_BYTES), several z chunks a
    slab: the same fields as the oracle."""
    import os
    pc, Ca, Cb, beam, ref = soil60
    g = pc.geom_params
    say('case: slabs of 3 chunks, launch', launch)
    os.environ['FDTD_SLAB_BYTES'] = str(32 * 3 * launch[1] * g.size_y * g.size_x)
    try:
        sim = make_sim(pc, Ca, Cb, beam, launch=launch, fused=False)
    finally:
        del os.environ['FDTD_SLAB_BYTES']
    sim.run(60)
    sim.sync()
    for n in native.FIELD_NAMES:
        assert_same(sim.field(n), ref.fields[n], n)
    ph, oph = sim.phasors(), ref.phasors()
    for n in native.PHASOR_NAMES:
        assert_same(ph[n], oph[n], n)
    sim.destroy()


def test_uneven_grid_bitwise():
    pc, Ca, Cb = H.soil_case(nx=37, ny=29, nz=41, nz_mat=17, steps=40, seed=11)
    say('case: uneven 37x29x41, 40 steps')
    beam = H.synthetic_beam(pc)
    ref = H.oracle_main(pc, Ca, Cb, beam)
    sim = make_sim(pc, Ca, Cb, beam, launch=(8, 9, 1, 1, 1), fused=True)
    sim.run(40)
    sim.sync()
    for n in native.FIELD_NAMES:
        assert_same(sim.field(n), ref.fields[n], n)
    sim.destroy()


@pytest.fixture(scope='module')
def wide30():
    pc, Ca, Cb = H.soil_case(nx=301, ny=25, nz=27, nz_mat=10, steps=30, seed=7)
    beam = H.synthetic_beam(pc)
    return pc, Ca, Cb, beam, H.oracle_main(pc, Ca, Cb, beam)


@pytest.mark.parametrize('fused,xwaves', [(False, 1), (True, 1), (False, 4), (False, 2), (2, 1), (3, 4),
                                          (3, 1)],
                         ids=['split', 'fused', 'split-4-waves', 'split-2-waves', 'rows', 'halo-4-waves',
                              'halo'])
@pytest.mark.parametrize('xvec', [4, 2, 1])
def test_wide_grid_lane_widths_bitwise(wide30, fused, xwaves, xvec):
    # 301 cells in x: two 256-cell tiles at 4 cells per lane (the second
    # with 12 active lanes), three at 2, five at 1; x +- 1 neighbours cross
    # lanes and tiles (the fused stepper recomputes the cell left of a tile)
    pc, Ca, Cb, beam, ref = wide30
    say('case: wide 301x25x27, 30 steps, fused', fused, 'xvec', xvec, 'xwaves', xwaves)
    launch = (2, 8, xvec) if xwaves == 1 else (4 // xwaves, 2, xvec, xwaves, 2 if xvec == 1 else 0)
    if fused == 2:
        launch = (3, 4, xvec, 1, 0)
    if fused in (2, 3) and xvec == 4:
        # the row-parallel and halo steppers take 1 or 2 cells per lane: the
        # shape is refused on the host and the simulation keeps its previous one
        sim = make_sim(pc, Ca, Cb, beam, fused=fused)
        with pytest.raises(native.FdtdError, match='xvec must be 1 or 2'):
            sim.set_launch(*launch)
    else:
        sim = make_sim(pc, Ca, Cb, beam, launch=launch, fused=fused)
    sim.run(30)
    sim.sync()
    for n in native.FIELD_NAMES:
        assert_same(sim.field(n), ref.fields[n], f'{n} xvec {xvec}')
    for i, n in enumerate(native.PSI_NAMES):
        assert_same(sim.psi(i), ref.psi[n], f'{n} xvec {xvec}')
    sim.destroy()


@pytest.mark.parametrize('nx', [139, 77])
def test_x_face_taps_bitwise(nx):
    # the split stepper's DFT reads the x faces (x = 11 and nx - 11) from
    # the E kernel's copies of the face columns; hy, hz at x - 1 come from
    # the lane itself (x = 11 at 2 and 4 cells per lane), the lane below
    # (x = 66 at 2 cells per lane) or memory (x = 128: a wavefront's first
    # lane at 1 and 2 cells per lane)
    pc, Ca, Cb = H.soil_case(nx=nx, ny=27, nz=29, nz_mat=11, steps=150, seed=13,
                             dft_accumulation='reference')
    say('case: x-face taps', nx, 'x 27 x 29, 150 steps')
    beam = H.synthetic_beam(pc)
    ref = H.oracle_main(pc, Ca, Cb, beam)
    oph = ref.phasors()
    for n in ('ey_ph_l', 'hy_ph_l', 'ez_ph_r', 'hz_ph_r'):
        assert np.abs(oph[n]).max() > 1e-4, n
    for launch in ((1, 2, 1, 4, 0), (1, 3, 1, 1, 0), (2, 2, 2, 2, 0), (1, 4, 2, 4, 0),
                   (4, 0, 4, 1, 0), (2, 5, 4, 2, 1), (1, 1, 1, 4, 2), (0, 2, 1, -1, 0), (0, 3, 0, -1, 0)):
        say('  launch', launch)
        sim = make_sim(pc, Ca, Cb, beam, launch=launch, fused=False)
        sim.run(150)
        sim.sync()
        ph = sim.phasors()
        for n in native.PHASOR_NAMES:
            assert_same(ph[n], oph[n], f'{n} {launch}')
        sim.destroy()


@pytest.mark.parametrize('nx', [22, 23])
def test_minimum_interior_bitwise(nx):
    # npml 10: nx = 22 is the smallest grid accepted (2 interior columns),
    # where both x faces lie on one column (x = 11); at 23 they are adjacent
    pc, Ca, Cb = H.soil_case(nx=nx, ny=24, nz=26, nz_mat=8, steps=60, seed=17,
                             dft_accumulation='reference')
    say('case: minimum interior', nx, 'x 24 x 26, 60 steps')
    beam = H.synthetic_beam(pc)
    ref = H.oracle_main(pc, Ca, Cb, beam)
    oph = ref.phasors()
    for fused, launch in ((False, None), (False, (1, 2, 2, 2, 0)), (False, (2, 3, 4, 1, 1)),
                          (True, None), (2, None)):
        say('  stepper', fused, 'launch', launch)
        sim = make_sim(pc, Ca, Cb, beam, launch=launch, fused=fused)
        sim.run(60)
        sim.sync()
        for n in native.FIELD_NAMES:
            assert_same(sim.field(n), ref.fields[n], f'{n} {fused} {launch}')
        for i, n in enumerate(native.PSI_NAMES):
            assert_same(sim.psi(i), ref.psi[n], f'{n} {fused} {launch}')
        ph = sim.phasors()
        for n in native.PHASOR_NAMES:
            assert_same(ph[n], oph[n], f'{n} {fused} {launch}')
        sim.destroy()


# flat tiles: wavefronts over 64 consecutive cells of a plane's row-major
# (y, x) index, running on into the next row (pitches 112, 80 and 144; at
# 139 cells the right DFT x face, x = 128, is a wavefront's first lane in
# the first row and mid-wavefront in others); None: the automatic shape,
# which takes flat tiles for these widths
@pytest.fixture(scope='module', params=[(101, 23, 27, 9), (67, 29, 31, 12), (139, 27, 29, 11)],
                ids=['101', '67', '139'])
def flat_case(request):
    nx, ny, nz, nz_mat = request.param
    pc, Ca, Cb = H.soil_case(nx=nx, ny=ny, nz=nz, nz_mat=nz_mat, steps=50, seed=nx,
                             dft_accumulation='reference')
    beam = H.synthetic_beam(pc)
    return pc, Ca, Cb, beam, H.oracle_main(pc, Ca, Cb, beam)


@pytest.mark.parametrize('launch', [None, (0, 2, 1, -1, 0), (0, 5, 0, -1, 0), (0, 0, 1, -1, 0),
                                    (0, 1, 1, -1, 0), (0, 2, 1, -1, 3), (0, 3, 1, -1, 5),
                                    (0, 4, 0, -1, 6)],
                         ids=['auto', 'flat-2', 'flat-5', 'flat-column', 'flat-1', 'flat-bands-1',
                              'flat-bands-4', 'flat-bands-8'])
def test_flat_tiles_bitwise(flat_case, launch):
    pc, Ca, Cb, beam, ref = flat_case
    g = pc.geom_params
    say('case: flat tiles', g.size_x, 'x', g.size_y, 'x', g.size_z, 'launch', launch)
    sim = make_sim(pc, Ca, Cb, beam, launch=launch, fused=False)
    sim.run(50)
    sim.sync()
    for n in native.FIELD_NAMES:
        assert_same(sim.field(n), ref.fields[n], f'{n} {launch}')
    for i, n in enumerate(native.PSI_NAMES):
        assert_same(sim.psi(i), ref.psi[n], f'{n} {launch}')
    ph, oph = sim.phasors(), ref.phasors()
    for n in native.PHASOR_NAMES:
        assert_same(ph[n], oph[n], f'{n} {launch}')
    sim.destroy()


def test_flat_slabs_bitwise(flat_case):
    """Flat tiles launched slab by slab (FDTD_SLAB_BYTES), three 2-plane
    chunks a slab: the same fields and phasors as the oracle."""
    import os
    pc, Ca, Cb, beam, ref = flat_case
    g = pc.geom_params
    say('case: flat tiles in slabs,', g.size_x, 'x', g.size_y, 'x', g.size_z)
    os.environ['FDTD_SLAB_BYTES'] = str(32 * 3 * 2 * g.size_y * g.size_x)
    try:
        sim = make_sim(pc, Ca, Cb, beam, launch=(0, 2, 1, -1, 4), fused=False)
    finally:
        del os.environ['FDTD_SLAB_BYTES']
    sim.run(50)
    sim.sync()
    for n in native.FIELD_NAMES:
        assert_same(sim.field(n), ref.fields[n], n)
    ph, oph = sim.phasors(), ref.phasors()
    for n in native.PHASOR_NAMES:
        assert_same(ph[n], oph[n], n)
    sim.destroy()


def test_flat_batch_equals_solo():
    """Members of widths 101, 67 and 139 (pitches 112, 80, 144) stepped as
    one batch on flat tiles (automatic shape, explicit flat tiles, and each
    member's own launches on lanes): bit-identical to each run alone on row
    tiles."""
    cases = [H.soil_case(nx=nx, ny=ny, nz=nz, nz_mat=m, steps=30, seed=nx)
             for nx, ny, nz, m in ((101, 23, 27, 9), (67, 29, 31, 12), (139, 27, 29, 11))]
    solo, sims = [], []
    for pc, Ca, Cb in cases:
        beam = H.synthetic_beam(pc)
        s = make_sim(pc, Ca, Cb, beam, launch=(1, 2, 1, 1, 0))
        s.run(30)
        s.sync()
        solo.append({n: s.field(n) for n in native.FIELD_NAMES})
        solo[-1].update(s.phasors())
        s.destroy()
        sims.append(make_sim(pc, Ca, Cb, beam))
    for launch, mode in ((None, None), ((0, 4, 0, -1, 0), None), ((0, 3, 0, -1, 4), None),
                         (None, (native.BATCH_SERIAL, 0, 2))):
        say('case: flat batch, launch', launch, 'mode', mode)
        for s in sims:
            s.reset()
        b = native.Batch(sims)
        if launch:
            b.set_launch(*launch)
        if mode:
            b.set_mode(*mode)
        b.run(11, use_graph=False)
        b.run(19, use_graph=True)
        b.sync()
        for k, s in enumerate(sims):
            assert s.step_count() == 30
            for n in native.FIELD_NAMES:
                assert_same(s.field(n), solo[k][n], f'member {k} {n} {launch} {mode}')
            ph = s.phasors()
            for n in native.PHASOR_NAMES:
                assert_same(ph[n], solo[k][n], f'member {k} {n} {launch} {mode}')
        b.destroy()
    for s in sims:
        s.destroy()


def test_incident_run_flat_bitwise():
    # 67 cells wide (pitch 80): the incident run's automatic shape takes flat tiles
    pc = H.soil_case(nx=67, ny=33, nz=35, nz_mat=14, dft_accumulation='reference')[0]
    say('case: incident run 67x33 (flat tiles), nt', pc.geom_params.size_t_inc)
    pex, pey = H.excitation(pc)
    rec, oph = H.oracle_incident(pc, pex, pey)
    from sar_pixel_simulator_amd.incidentbeam import run_incident
    beam, ph = run_incident(pc, pex, pey, device=0)
    for c in range(4):
        assert_same(beam.component(c), rec[c], f'beam component {c}')
    for n in native.INC_PHASOR_NAMES:
        assert_same(ph[n], oph[n], n)
    beam.release()


def test_invalid_launch_is_rejected_and_previous_kept(soil60):
    pc, Ca, Cb, beam, ref = soil60
    sim = make_sim(pc, Ca, Cb, beam, launch=(2, 0, 4, 1, 1), fused=False)
    with pytest.raises(native.FdtdError):
        sim.set_launch(8, 0, 4, 1, 1)          # 8 rows per tile is a fused-stepper shape
    with pytest.raises(native.FdtdError):
        sim.set_launch(2, 0, 3, 1, 1)
    with pytest.raises(native.FdtdError):
        sim.set_launch(2, 0, 1, 4, 1)       # 8 wavefronts per block
    with pytest.raises(native.FdtdError):
        sim.set_launch(1, 0, 1, 3, 1)
    with pytest.raises(native.FdtdError):
        sim.set_launch(1, 4, 1, 4, 7)    # block orders are 0 - 6
    with pytest.raises(native.FdtdError):
        sim.set_launch(0, 2, 1, -1, 0)   # flat tiles need a row pitch of 64 (33 cells: 48)
    fsim = make_sim(pc, Ca, Cb, beam, fused=True)
    with pytest.raises(native.FdtdError):
        fsim.set_launch(0, 2, 1, -1, 0)  # flat tiles are a split-stepper shape
    with pytest.raises(native.FdtdError):
        fsim.set_launch(8, 0, 2, 1, 1)         # 8 rows of 2-cell lanes exceed the fused LDS budget
    with pytest.raises(native.FdtdError):
        fsim.set_launch(2, 0, 1, 2, 1)      # one wavefront per strip
    with pytest.raises(native.FdtdError):
        fsim.set_launch(2, 0, 1, 1, 2)   # banded order is a split-stepper order
    fsim.destroy()
    sim.run(60)
    sim.sync()
    assert_same(sim.field('hz'), ref.fields['hz'], 'hz after rejected shapes')
    sim.destroy()


def test_incident_run_bitwise():
    pc = H.soil_case(nx=31, ny=33, nz=35, nz_mat=14, dft_accumulation='reference')[0]
    say('case: incident run 31x33, nt', pc.geom_params.size_t_inc)
    pex, pey = H.excitation(pc)
    rec, oph = H.oracle_incident(pc, pex, pey)
    from sar_pixel_simulator_amd.incidentbeam import run_incident
    beam, ph = run_incident(pc, pex, pey, device=0)
    for c in range(4):
        assert_same(beam.component(c), rec[c], f'beam component {c}')
    for n in native.INC_PHASOR_NAMES:
        assert_same(ph[n], oph[n], n)
    assert np.abs(rec[1]).max() > 0.1
    beam.release()


@pytest.fixture(scope='module')
def incident31():
    pc = H.soil_case(nx=31, ny=33, nz=35, nz_mat=14, dft_accumulation='reference')[0]
    pex, pey = H.excitation(pc)
    rec, oph = H.oracle_incident(pc, pex, pey)
    return pc, pex, pey, rec, oph


@pytest.mark.parametrize('launch', [(1, 2, 1, 4, 0), (2, 3, 2, 2, 1), (1, 1, 4, 1, 2), (4, 0, 1, 1, 0)],
                         ids=['wide', 'xcd-2-cells', 'banded-4-cells', 'whole-column'])
def test_incident_launch_shapes_bitwise(incident31, launch):
    # the incident run steps E before H (hard source, recorded read-out
    # plane, incident phasors) on the split kernels in any tile shape
    pc, pex, pey, rec, oph = incident31
    g = pc.geom_params
    say('case: incident run, launch', launch)
    native.init(0)
    geom = common.incident_geom(pc)
    sim = native.Sim(geom, common.coeffs_for(pc))
    sim.set_launch(*launch)
    sim.set_uniform_material(1.0, common.incident_cb(pc))
    n = geom.nt_inc
    sim.set_source(np.ascontiguousarray(np.imag(pex[:n]), dtype=np.float32),
                   np.ascontiguousarray(np.imag(pey[:n]), dtype=np.float32))
    beam = native.Beam(g.size_t_inc, g.size_y, g.size_x)
    sim.attach_beam(beam)
    sim.run(g.size_t_inc)
    sim.sync()
    ph = sim.phasors()
    sim.destroy()
    for c in range(4):
        assert_same(beam.component(c), rec[c], f'beam component {c} {launch}')
    for nm in native.INC_PHASOR_NAMES:
        assert_same(ph[nm], oph[nm], f'{nm} {launch}')
    beam.release()


def test_split_and_fused_agree_long_run():
    """300 steps on the plain workload's grid: all four steppers bit-identical."""
    from sar_pixel_simulator_amd import workloads
    from sar_pixel_simulator_amd.materials import compute_material_arrays
    pc = workloads.plain_params()
    Ca, Cb = compute_material_arrays(pc, workloads.plain_material(pc, 2))
    beam = H.synthetic_beam(pc)
    say('case: 81^3, 300 steps, split vs the three fused steppers')
    out = []
    for fused in (False, True, 2, 3):
        sim = make_sim(pc, Ca, Cb, beam, fused=fused)
        sim.run(300)
        sim.sync()
        out.append(({n: sim.field(n) for n in native.FIELD_NAMES}, sim.phasors()))
        sim.destroy()
    for other in out[1:]:
        for n in native.FIELD_NAMES:
            assert_same(other[0][n], out[0][0][n], n)
        for n in native.PHASOR_NAMES:
            assert_same(other[1][n], out[0][1][n], n)


def test_split_and_rows_agree_at_largest_size():
    """The largest single domain, 1523 x 1523 x 311 (7.2e8 cells), 240 steps
    on device-built snow over rough soil: the split stepper and the
    row-parallel fused stepper, two independent implementations of the
    step, and the halo-recomputing fused stepper give bit-identical fields
    and phasors."""
    from sar_pixel_simulator_amd import workloads
    from sar_pixel_simulator_amd.materials import device
    nsteps = 240
    pc = workloads.largest_params()
    g = pc.geom_params
    g.size_t = g.size_t_inc = nsteps
    snow = int(round(1.5 / 0.021))
    spec = device.snow_over_soil_spec(pc, g.size_z_material - snow, snow, seed=3)
    beam = workloads.synthetic_beam_planes(pc, 0, nsteps)
    say('case: largest', g.size_x, g.size_y, g.size_z, '240 steps, split vs row-parallel and halo')

    def run(fused):
        native.init(0)
        sim = native.Sim(common.main_geom(pc), common.coeffs_for(pc))
        sim.set_stepper(fused)
        b = native.Beam(g.size_t_inc, g.size_y, g.size_x, *beam)
        sim.attach_beam(b)
        sim.build_materials(*device.to_c(spec, pc))
        sim.run(nsteps)
        sim.sync()
        return sim, b

    sim, b = run(False)
    ref = {n: sim.field(n) for n in native.FIELD_NAMES}
    ref_ph = sim.phasors()
    sim.destroy()
    b.release()
    assert np.abs(ref['ey']).max() > 0 and np.abs(ref_ph['ey_ph_l']).max() > 0
    for fused in (2, 3):
        say('  stepper', fused)
        sim, b = run(fused)
        for n in native.FIELD_NAMES:
            assert_same(sim.field(n), ref[n], f'{n} stepper {fused}')
        ph = sim.phasors()
        for n in native.PHASOR_NAMES:
            assert_same(ph[n], ref_ph[n], f'{n} stepper {fused}')
        sim.destroy()
        b.release()


@pytest.mark.parametrize('fused,launch', [(True, None), (False, None), (False, (2, 8, 4, 1, 1)),
                                          (False, (1, 1, 1, 4, 0)), (False, (1, 4, 1, 4, 2)),
                                          (2, None), (False, 'mixed-widths'), (False, 'serial'),
                                          (False, 'serial-split'), (False, 'hybrid'), (2, 'hybrid'),
                                          (3, None), (3, 'mixed-widths'), (3, 'serial-split'),
                                          (False, (1, 2, 1, 4, 4)), (3, (2, 8, 1, 2, 3))],
                         ids=['fused', 'split-auto', 'split-4-per-lane', 'split-4-waves-flat',
                              'split-banded', 'rows', 'split-auto-mixed-widths', 'split-serial',
                              'split-serial-small-lanes', 'split-hybrid', 'rows-hybrid', 'halo-auto',
                              'halo-auto-mixed-widths', 'halo-serial-small-lanes', 'split-fine-bands',
                              'halo-fine-bands'])
def test_batch_equals_solo(fused, launch):
    cases = [H.soil_case(nx=33, ny=31, nz=35, nz_mat=14, steps=30, seed=s) for s in (1, 2)]
    cases.append(H.soil_case(nx=45, ny=39, nz=37, nz_mat=15, steps=30, seed=5))
    if launch == 'mixed-widths':
        # automatic shape: the 201-cell member gets 4-wavefront-wide blocks
        # next to the 33- and 45-cell members' 1-wavefront blocks
        cases.append(H.soil_case(nx=201, ny=23, nz=25, nz_mat=9, steps=30, seed=6))
        cases.append(H.soil_case(nx=101, ny=23, nz=27, nz_mat=9, steps=30, seed=8))
        launch = None
    mode = None
    if launch in ('serial', 'serial-split', 'hybrid'):
        # serial: every member's own launches in turn, on two concurrent
        # lanes; serial-split: the two 33-cell members on two lanes, then
        # the 45-cell one; hybrid: the two 33-cell members packed, the
        # 45-cell one serial on the batch stream
        mode = {'serial': (native.BATCH_SERIAL, 0, 2), 'serial-split': (native.BATCH_SERIAL, 40000, 2),
                'hybrid': (native.BATCH_HYBRID, 40000, 1)}[launch]
        launch = None
    say('case: batch of', len(cases), 'uneven sims, 30 steps, fused', fused, 'launch', launch)
    solo, batch_sims = [], []
    for pc, Ca, Cb in cases:
        beam = H.synthetic_beam(pc)
        s = make_sim(pc, Ca, Cb, beam, fused=fused)
        s.run(30)
        s.sync()
        solo.append({n: s.field(n) for n in native.FIELD_NAMES})
        solo[-1].update(s.phasors())
        s.destroy()
        batch_sims.append(make_sim(pc, Ca, Cb, beam, fused=fused))
    for use_graph in (False, True):
        for s in batch_sims:
            s.reset()
        b = native.Batch(batch_sims)
        if launch:
            b.set_launch(*launch)
        if mode:
            b.set_mode(*mode)
        b.run(17, use_graph=use_graph)
        b.run(13, use_graph=use_graph)
        b.sync()
        for k, s in enumerate(batch_sims):
            assert s.step_count() == 30
            for n in native.FIELD_NAMES:
                assert_same(s.field(n), solo[k][n], f'member {k} {n} graph={use_graph}')
            ph = s.phasors()
            for n in native.PHASOR_NAMES:
                assert_same(ph[n], solo[k][n], f'member {k} {n} graph={use_graph}')
        b.destroy()
    for s in batch_sims:
        s.destroy()


def test_near2far_matches_scipy(soil60):
    pc, Ca, Cb, beam, ref = soil60
    g = pc.geom_params
    oph = ref.phasors()
    nxp, nyp, nzp = g.size_x - 20, g.size_y - 20, g.size_z - 20
    th = np.linspace(-80.0, 80.0, 161)
    want = oracle.far_field(oph, nxp, nyp, nzp, g.cell_size, 2000.0, th, [0.0],
                            pc.inc_beam_params.frequency)
    native.init(0)
    got = native.near2far_host(oph['ex_ph_u'], oph['ey_ph_u'], oph['hx_ph_u'], oph['hy_ph_u'],
                               nxp, nyp, nzp, g.cell_size, th, 0.0, 2000.0,
                               pc.inc_beam_params.frequency)
    for k in ('E_theta', 'E_phi', 'H_theta', 'H_phi'):
        scale = np.abs(want[k]).max()
        assert np.abs(got[k] - want[k]).max() <= 1e-9 * scale, k
    # from the device accumulators directly
    sim = make_sim(pc, Ca, Cb, beam)
    sim.run(60)
    sim.sync()
    got2 = sim.near2far(th, 0.0, 2000.0, pc.inc_beam_params.frequency)
    for k in ('E_theta', 'E_phi'):
        scale = np.abs(want[k]).max()
        assert np.abs(got2[k] - want[k]).max() <= 1e-9 * scale, k
    sim.destroy()


def test_linearity_at_large_size():
    """Size-independent property at the single-large-run shape: doubling the
    beam doubles every field exactly (all operations commute with x2)."""
    from sar_pixel_simulator_amd import workloads
    pc = workloads.large_params()
    g = pc.geom_params
    say('case: linearity', g.size_x, g.size_y, g.size_z)
    steps = 4
    Ca = np.ones((g.size_z, g.size_y, g.size_x), np.float32)
    Cb = np.full_like(Ca, np.float32(common.update_scalars(pc)['dte0dx']))
    rng = np.random.default_rng(0)
    Ca[:g.size_z_material] = rng.uniform(0.9, 1.0, (g.size_z_material, g.size_y, g.size_x))
    planes = workloads.synthetic_beam_planes(pc, 0, steps)
    native.init(0)
    out = []
    for scale in (1.0, 2.0):
        sim = native.Sim(common.main_geom(pc), common.coeffs_for(pc))
        b = native.Beam(g.size_t_inc, g.size_y, g.size_x)
        b.upload(0, *(scale * planes))
        sim.attach_beam(b)
        sim.set_materials(Ca, Cb)
        sim.run(steps)
        sim.sync()
        out.append(sim.field('ey'))
        sim.destroy()
        b.release()
    assert np.isfinite(out[0]).all()
    assert np.abs(out[0]).max() > 0
    assert_same(out[1], 2.0 * out[0], 'ey(2 beam) vs 2 ey(beam)')


def _exact_dft(pc, Ca, Cb, beam, steps):
    """FP64 running DFT of six face components (top face ex, hy; left face
    ey, hz; right face hy, hz), accumulated from the oracle's float32 field
    history step by step, as cpu/field_updates.py:373-461 form each term
    (float32 neighbour sums, E at phase t, H at t - 1/2): the exact sum the
    two accumulation modes approximate."""
    g = pc.geom_params
    n = g.size_pml
    s = H.scalars(pc)
    sim = oracle.OracleSim(H.oracle_geom(pc), Ca, Cb, H.profiles_of(pc), inc=beam)
    nzp, nyp, nxp = g.size_z - 2 * n, g.size_y - 2 * n, g.size_x - 2 * n
    zt = g.size_z - n - 1
    xs = {'l': n + 1, 'r': g.size_x - n - 1}
    acc = {k: np.zeros((nyp, nxp), np.complex128) for k in ('ex_ph_u', 'hy_ph_u')}
    acc.update({k: np.zeros((nzp, nyp), np.complex128)
                for k in ('ey_ph_l', 'hz_ph_l', 'hy_ph_r', 'hz_ph_r')})
    Y, X = slice(n, n + nyp), slice(n, n + nxp)
    Y1, X1 = slice(n - 1, n - 1 + nyp), slice(n - 1, n - 1 + nxp)
    Z, Z1 = slice(n, n + nzp), slice(n - 1, n - 1 + nzp)
    for t in range(steps):
        sim.run(1)
        F = sim.fields
        we = np.exp(1j * ((-s['omega'] * t) * s['dt']))
        wh = np.exp(1j * ((-s['omega'] * (t - 0.5)) * s['dt']))
        ex, hy = F['ex'][zt], F['hy']
        acc['ex_ph_u'] += 0.5 * (ex[Y, X] + ex[Y, X1]).astype(np.float64) * we
        acc['hy_ph_u'] += 0.25 * (((hy[zt, Y, X] + hy[zt, Y, X1]) + hy[zt - 1, Y, X])
                                  + hy[zt - 1, Y, X1]).astype(np.float64) * wh
        for f, x in xs.items():
            ey, hz = F['ey'][:, :, x], F['hz']
            hzx = (((hz[Z, Y, x] + hz[Z, Y, x - 1]) + hz[Z, Y1, x]) + hz[Z, Y1, x - 1])
            acc['hz_ph_' + f] += 0.25 * hzx.astype(np.float64) * wh
            if f == 'l':
                acc['ey_ph_l'] += 0.5 * (ey[Z, Y] + ey[Z, Y1]).astype(np.float64) * we
            else:
                hyx = (((hy[Z, Y, x] + hy[Z, Y, x - 1]) + hy[Z1, Y, x]) + hy[Z1, Y, x - 1])
                acc['hy_ph_r'] += 0.25 * hyx.astype(np.float64) * wh
    for a in acc.values():
        a[0, :] = 0
        a[:, 0] = 0
    return acc


@STEPPERS
def test_compensated_dft_is_closer_to_exact_sum(fused):
    pc, Ca, Cb = H.soil_case(steps=120, seed=9)
    beam = H.synthetic_beam(pc)
    say('case: DFT accumulation modes vs exact sum, 120 steps, stepper', fused)
    exact = _exact_dft(pc, Ca, Cb, beam, 120)
    err = {}
    for mode in ('reference', 'compensated'):
        pc.processing_params.dft_accumulation = mode
        sim = make_sim(pc, Ca, Cb, beam, fused=fused)
        sim.run(120)
        sim.sync()
        ph = sim.phasors()
        err[mode] = {k: H.rel_l2(ph[k], want) for k, want in exact.items()}
        sim.destroy()
    say('relative L2 error vs exact sum', err)
    for k in exact:
        assert err['compensated'][k] <= err['reference'][k], k
        assert err['compensated'][k] < 1e-6 and err['reference'][k] < 1e-5, k


@pytest.mark.parametrize('pol,theta', [('HH', 20.0), ('VV', 35.0), ('RHCP', 0.0)])
def test_device_beam_synthesis_matches_host(pol, theta):
    """Device pipeline (spectra, FFTs, normalisation, incident FDTD) vs the
    host numpy synthesis followed by the same incident FDTD.  They differ by
    FFT rounding only (hipFFT vs numpy pocketfft; complex64 inverse)."""
    from sar_pixel_simulator_amd import workloads
    from sar_pixel_simulator_amd.incidentbeam import run_incident, synthesize_on_device
    pc = workloads.base_params(polarization=pol, theta=theta, incident_duration=2.0,
                               illumination_width=1.0, simulation_duration=3.0)
    workloads.explicit_geometry(pc, 37, 33, 41, 15, 0.05)
    say('case: beam synthesis', pol, theta)
    pex, pey = H.excitation(pc)
    host_beam, host_ph = run_incident(pc, pex, pey, device=0)
    dev_beam, dev_ph = synthesize_on_device(pc, device=0)
    for c in range(4):
        a, b = dev_beam.component(c), host_beam.component(c)
        assert np.abs(b).max() > 0
        assert H.rel_l2(a, b) < 1e-4, (c, H.rel_l2(a, b))
    for n in native.INC_PHASOR_NAMES:
        assert H.rel_l2(dev_ph[n], host_ph[n]) < 1e-4, n
    host_beam.release()
    dev_beam.release()


def test_bandwidth_probes_run_and_reject_bad_arguments():
    """The probes bench.py reports as practical ceilings run on small sizes
    and reject unbuilt shapes with an error, not a launch."""
    say('case: probes, small sizes')
    ms, gbs = native.probe_copy(1 << 20, vec=1, reps=2)
    assert ms > 0 and gbs > 0
    ms, gbs = native.probe_streams(1 << 16, 6, 3, reps=2)
    assert ms > 0 and gbs > 0
    for kind in (0, 1, 2):
        for zc in (1, 4, 64):
            ms, gbs = native.probe_stencil(301, 17, 9, zc, kind, reps=2)
            assert ms > 0 and gbs > 0
    with pytest.raises(native.FdtdError):
        native.probe_stencil(64, 8, 8, 4, 3)
    with pytest.raises(native.FdtdError):
        native.probe_stencil(64, 8, 8, 0, 0)
    with pytest.raises(native.FdtdError):
        native.probe_streams(1 << 16, 5, 3)


def test_near2far_grid_matches_oracle(soil60):
    """The far field over a theta x phi grid in one launch, from host
    phasors and from the device accumulators, against the oracle's FP64
    scipy-Simpson transform (relative 1e-9); each phi row equals the
    single-phi transform bit for bit."""
    pc, Ca, Cb, beam, ref = soil60
    g = pc.geom_params
    oph = ref.phasors()
    nxp, nyp, nzp = g.size_x - 20, g.size_y - 20, g.size_z - 20
    th = np.linspace(-80.0, 80.0, 17)
    phis = np.arange(0.0, 360.0, 30.0)
    f = pc.inc_beam_params.frequency
    want = oracle.far_field(oph, nxp, nyp, nzp, g.cell_size, 2000.0, th, phis, f)
    native.init(0)
    got = native.near2far_host_grid(oph['ex_ph_u'], oph['ey_ph_u'], oph['hx_ph_u'], oph['hy_ph_u'],
                                    nxp, nyp, nzp, g.cell_size, th, phis, 2000.0, f)
    sim = make_sim(pc, Ca, Cb, beam)
    sim.run(60)
    sim.sync()
    got2 = sim.near2far_grid(th, phis, 2000.0, f)
    for k in ('E_theta', 'E_phi', 'H_theta', 'H_phi'):
        scale = np.abs(want[k]).max()
        assert got[k].shape == (phis.size, th.size)
        assert np.abs(got[k] - want[k]).max() <= 1e-9 * scale, k
        assert np.abs(got2[k] - want[k]).max() <= 1e-9 * scale, k
    for p, phi in enumerate(phis):
        one = sim.near2far(th, phi, 2000.0, f)
        for k in one:
            assert np.array_equal(one[k], got2[k][p]), (phi, k)
    sim.destroy()


def test_pec_box_powers_and_hemisphere_on_device():
    """Device pipeline (device beam synthesis, split stepper, compensated
    DFT) on a flat PEC plate under a normally incident 2-lambda beam: the
    incident power returns through the top half of the phasor box, none
    reaches the bottom face, and the far field of the top face, computed
    over the upper hemisphere in one launch, carries the power crossing it."""
    from sar_pixel_simulator_amd import processresults, workloads
    from sar_pixel_simulator_amd.hip import batch as B
    from sar_pixel_simulator_amd.materials import compute_material_arrays, flat_surface
    pc = workloads.base_params(theta=0.0, illumination_width=2.0, incident_duration=4.0,
                               simulation_duration=11.0)
    workloads.explicit_geometry(pc, 101, 101, 41, 14, 0.05)
    g = pc.geom_params
    say('case: PEC box 101 x 101 x 41 on the device')
    native.init(0)
    beams = B.BeamCache()
    Ca, Cb = compute_material_arrays(pc, next(flat_surface(pc, 1.0, 1.0e10)))
    dm = B.DeviceModel(pc, Ca, Cb, beams)
    B.run_plain([dm])
    ph = dict(dm.sim.phasors())
    ph.update(dm.inc_phasors)
    pw = processresults.near_field_power(ph, g.cell_size, g.size_z_material)
    dth, dphi, r = 2.0, 10.0, 2000.0
    th = np.arange(dth / 2, 90.0, dth)
    phis = np.arange(0.0, 360.0, dphi)
    ff = dm.sim.near2far_grid(th, phis, r, pc.inc_beam_params.frequency)
    p_h, p_v = processresults.farfield_power(ff['E_theta'], ff['E_phi'], ff['H_theta'], ff['H_phi'])
    w = np.sin(np.deg2rad(th))[None, :] * np.deg2rad(dth) * np.deg2rad(dphi) * r * r
    hemi = float(((p_h + p_v) * w).sum())
    p_inc = dm.p_inc
    say('powers / P_inc: top', pw['top_total'] / p_inc, 'bottom', pw['bottom_total'] / p_inc,
        'hemisphere', hemi / p_inc)
    dm.release()
    beams.release()
    assert pw['top_total'] == pytest.approx(p_inc, rel=0.02)
    assert abs(pw['bottom_total']) < 0.01 * p_inc
    assert hemi == pytest.approx(p_inc, rel=0.02)
    assert hemi == pytest.approx(pw['up'], rel=0.01)


def test_rejected_cached_shape_falls_back_to_default(soil60, tmp_path, monkeypatch):
    """A cached tile shape the library refuses for its grid (flat tiles on a
    33-cell row, pitch 48) gives way to the stepper's default shape: the
    simulation still steps to the oracle's fields."""
    import json
    from sar_pixel_simulator_amd.hip import tuning
    pc, Ca, Cb, beam, ref = soil60
    g = pc.geom_params
    cache = tmp_path / 'cache.json'
    cache.write_text(json.dumps({f'split:gfx950:{g.size_x}x{g.size_y}x{g.size_z}':
                                 {'ty': 0, 'zchunk': 2, 'xvec': 1, 'xwaves': -1, 'xcd_order': 0}}))
    monkeypatch.setattr(tuning, 'CACHE', str(cache))
    assert tuning.launch_for(g.size_x, g.size_y, g.size_z, fused=0) == (0, 2, 1, -1, 0)
    native.init(0)
    sim = common.new_main_sim(pc, fused=0)
    b = native.Beam(g.size_t_inc, g.size_y, g.size_x, *beam)
    sim.attach_beam(b)
    sim.set_materials(Ca, Cb)
    sim.run(60)
    sim.sync()
    for n in native.FIELD_NAMES:
        assert_same(sim.field(n), ref.fields[n], n)
    sim.destroy()
    b.release()


def test_more_row_blocks_than_a_launch_grid_holds():
    """22 x 66001 x 22 with one row per block: 66001 row blocks exceed a
    launch grid's y extent (65535), so the kernels take their tiles from
    the list; fields equal the oracle's on the split and halo steppers."""
    pc, Ca, Cb = H.soil_case(nx=22, ny=66001, nz=22, nz_mat=8, steps=3, seed=2,
                             dft_accumulation='reference')
    say('case: 22 x 66001 x 22, 3 steps, one row per block')
    beam = H.synthetic_beam(pc)
    ref = H.oracle_main(pc, Ca, Cb, beam)
    for fused, launch in ((False, (1, 2, 1, 1, 0)), (3, (1, 8, 1, 1, 0))):
        sim = make_sim(pc, Ca, Cb, beam, launch=launch, fused=fused)
        sim.run(3)
        sim.sync()
        for n in native.FIELD_NAMES:
            assert_same(sim.field(n), ref.fields[n], f'{n} stepper {fused}')
        sim.destroy()
