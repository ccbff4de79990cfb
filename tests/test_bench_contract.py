"""bench.py pieces that run without a GPU: the multi-GPU refusal, the
algorithmic byte model and the PMC traffic arithmetic."""
import os
import sqlite3
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert 'one GPU' in r.stdout


def test_algorithmic_bytes_without_cpml_or_uniform_planes():
    # npml 1: no CPML cells (every face range is empty)
    ab = bench.algorithmic_bytes(10, 12, 14, 1, uni_z=None, taps=False)
    n = 10 * 12 * 14
    assert ab['cpml_cells'] == 0
    assert ab['h'] == 36 * n + 8 * 10 * 12
    assert ab['e'] == 44 * n + 8 * 10 * 12
    # all planes uniform: the E half reads no Ca/Cb
    ab0 = bench.algorithmic_bytes(10, 12, 14, 1, uni_z=0, taps=False)
    assert ab0['e'] == 36 * n + 8 * 10 * 12
    # split stepper: six x-face column values per (face, z, y) written
    abt = bench.algorithmic_bytes(10, 12, 14, 1, uni_z=None)
    assert abt['e'] - ab['e'] == 2 * 14 * 12 * 6 * 4
    assert abt['h'] == ab['h']


def test_cpml_cells_brute_force():
    nx, ny, nz, np_ = 9, 11, 13, 3
    count = 0
    # index ranges of the six faces (down/up, left/right, front/back)
    count += 2 * (np_ - 1) * (ny - 1) * (nx - 1)
    count += 2 * (nz - 1) * (ny - 1) * (np_ - 1)
    count += 2 * (nz - 1) * (np_ - 1) * (nx - 1)
    assert bench.cpml_cells(nx, ny, nz, np_) == count


def test_uniform_top():
    Ca = np.ones((6, 3, 4), np.float32)
    Cb = np.full((6, 3, 4), 0.5, np.float32)
    assert bench.uniform_top(Ca, Cb) == 0
    Ca[2, 1, 1] = 0.9
    assert bench.uniform_top(Ca, Cb) == 3
    Cb[4, 0, 0] = 0.4
    assert bench.uniform_top(Ca, Cb) == 5


def _fake_db(path, counter, kernel_kib, copy_kib):
    c = sqlite3.connect(path)
    c.execute('create table counters_collection (kernel_name text, counter_name text, value real)')
    for v in kernel_kib:
        c.execute('insert into counters_collection values (?, ?, ?)',
                  ('fdtd::k_e_half(fdtd::Desc const*, HIP_vector_type<int, 4u> const*, int)',
                   counter, v))
    c.execute('insert into counters_collection values (?, ?, ?)',
              ('fdtd::k_h_half(fdtd::Desc const*, HIP_vector_type<int, 4u> const*, int)',
               counter, 1.0))
    c.execute('insert into counters_collection values (?, ?, ?)',
              ('void fdtd::k_copy_probe<1>(float const*, float*, long long)', counter, copy_kib))
    c.commit()
    c.close()


def test_traffic_scales_each_counter_by_the_calibration_copy(tmp_path, monkeypatch):
    gib_kib = bench.PROBE_BYTES / 1024.0
    # FETCH reads the 1 GiB copy as 0.5 GiB (scale 2), WRITE exactly (scale 1)
    raw = {'FETCH_SIZE': ([1000.0, 1200.0], 0.5 * gib_kib), 'WRITE_SIZE': ([800.0], gib_kib)}

    def fake_pass(counter, child_args, out_dir):
        db = os.path.join(str(tmp_path), counter + '.db')
        _fake_db(db, counter, *raw[counter])
        return db
    monkeypatch.setattr(bench, 'pmc_pass', fake_pass)
    monkeypatch.setattr(bench, 'rocprofv3', lambda: '/bin/true')
    total, detail = bench.measure_traffic('k_e_half', [], keep_dir=str(tmp_path))
    fetch = 1100.0 * 1024 * 2.0
    write = 800.0 * 1024 * 1.0
    assert detail['fetch_scale'] == pytest.approx(2.0)
    assert detail['write_scale'] == pytest.approx(1.0)
    assert detail['fetch_bytes'] == int(fetch)
    assert total == int(fetch) + int(write)


def test_traffic_missing_kernel_reports_null(tmp_path, monkeypatch):
    def fake_pass(counter, child_args, out_dir):
        db = os.path.join(str(tmp_path), counter + '.db')
        _fake_db(db, counter, [], 1.0)
        return db
    monkeypatch.setattr(bench, 'pmc_pass', fake_pass)
    monkeypatch.setattr(bench, 'rocprofv3', lambda: '/bin/true')
    total, detail = bench.measure_traffic('k_e_half', [], keep_dir=str(tmp_path))
    assert total is None and 'no k_e_half' in detail['note'].replace('fdtd::', '')
