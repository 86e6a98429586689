"""Golgi-cell population on the GPU (global FP64 engine, grid gap junctions in
the cable matrix) through the C-ABI, against the oracle.

The Golgi cell amplifies last-bit differences the way the inferior olive
does: in the oracle itself a 1e-12 mV change of one cell's initial voltage
moves some spikes by one macro step within a second, through the step-size
controller (the rejected-step rate is about one in three).  The GPU evaluates
exp correctly rounded (exp_acc) and log with the device library; the oracle
uses the host C library's.  So:

* the first macro step agrees with the oracle to the rounding level;
* over 400 ms of a 2 x 2 grid (with a step current from 150 to 300 ms) every
  cell fires the same number of spikes; spikes before the current are at the
  oracle's macro step, later ones within two macro steps (0.25 ms).  That is
  the oracle's own spread on this workload: changing the initial voltages by
  one part in 1e15 moves the same spikes by up to 0.25 ms
  (tests/test_oracle_go.py::test_noise_floor_of_the_grid_workload);
* the reference-named entry points (cb_go_initialize with libc rand() seeded
  0) start from the oracle's initial state exactly.
"""
import ctypes
import os
import shutil

import numpy as np
import pytest

import cbrkc
import orc
from test_oracle_go import go_pop

pytestmark = pytest.mark.gpu

COMP = 131


def crossings(trace, thr=10.0):
    prev = np.full(trace.shape[1], 100.0)
    out = []
    for k in range(trace.shape[0]):
        v = trace[k]
        for c in np.nonzero((v > thr) & (prev <= thr))[0]:
            out.append((0.125 * (k + 1), int(c)))
        prev = v
    return out


def test_go_grid_matches_oracle(ref_data):
    nx, ny = 2, 2
    n = nx * ny
    v0 = np.linspace(-61.0, -59.0, n)
    stim = (30, 150.0, 300.0, 0.2e-3)
    p = cbrkc.Population(cbrkc.GO, nx, ny, morph=os.path.join(ref_data, "Go_131.txt"), v0=v0, stim=stim)
    o = go_pop(ref_data, nx, ny, v0=v0)
    o.set_stim(*stim)
    steps = 3200
    print(f"GO {nx}x{ny} grid, {steps} macro steps, GPU vs oracle", flush=True)
    tg, to = [], []
    for k in range(steps):
        p.set_current(k * 0.125)
        p.advance()
        o.set_current(k * 0.125)
        o.advance()
        vg, vo = p.read(cbrkc.F_VNEW), o.get(orc.VNEW)
        if k == 0:
            assert np.max(np.abs(vg - vo)) <= 1e-9
            for f, of in ((cbrkc.F_CA, orc.CA), (cbrkc.F_REV, orc.REV)):
                np.testing.assert_allclose(p.read(f), o.get(of), rtol=1e-12)
        tg.append(vg[30::COMP])
        to.append(vo[30::COMP])
    sg, so = crossings(np.array(tg)), crossings(np.array(to))
    print(f"GO {nx}x{ny}: {len(sg)} spikes on the GPU, {len(so)} in the oracle", flush=True)
    assert len(so) >= 2 * n
    for c in range(n):
        a = [t for t, cc in sg if cc == c]
        b = [t for t, cc in so if cc == c]
        assert len(a) == len(b), (c, a, b)
        assert [t for t in a if t < 150.0] == [t for t in b if t < 150.0], (c, a, b)
        assert np.max(np.abs(np.array(a) - np.array(b)), initial=0.0) <= 0.25, (c, a, b)
    st = p.stats()
    print("GO stats (steps, rejects, stages, rho iterations, ion updates, macro steps):", st, flush=True)


def test_go_reference_entry_points(ref_data, tmp_path, monkeypatch):
    shutil.copy(os.path.join(ref_data, "Go_131.txt"), tmp_path / "Go_131.txt")
    monkeypatch.chdir(tmp_path)
    ctypes.CDLL(None).srand(0)
    p = cbrkc.go_initialize(3, 1, "RKC")
    o = go_pop(ref_data, 3, 1, seed=0)
    np.testing.assert_array_equal(p.read(cbrkc.F_V), o.get(orc.V))
    for g in range(24):
        np.testing.assert_array_equal(p.read(cbrkc.F_GATE0 + g), o.get(orc.ION0 + g))
    memv_prev = np.full(3, 100.0)
    for k in range(8):
        cbrkc.go_set_current(p, k * 0.125)
        cbrkc.go_solve_update_v(p)
        o.set_current(k * 0.125)
        o.advance()
        v = p.read(cbrkc.F_VNEW)
        np.testing.assert_allclose(v, o.get(orc.VNEW), rtol=0, atol=1e-8)
        cbrkc.go_output_file(p, v, memv_prev, (k + 1) * 0.125, "go_v_by_RKC.csv", "go_raster_by_RKC.csv")
    rows = (tmp_path / "go_v_by_RKC.csv").read_text().splitlines()
    assert len(rows) == 8 and rows[0].startswith("0.125000,") and rows[0].count(",") == 4
    cbrkc.go_finalize(p)
