"""Granule-cell population on the GPU (global FP64 engine, 12-state Na scheme
solved per compartment in registers) through the C-ABI, against the oracle.

Two cells driven from 2 ms, 30 ms (240 macro steps, about 116 000 RKC steps,
most of them rejected as the reference's controller does at TOL 1e-9):
the first macro step agrees with the oracle to the rounding level; every cell
fires the same number of spikes, each within one macro step of the oracle's
time.  (In the oracle a one-ulp change of the initial voltages leaves these
spikes at the same macro step; the GPU's exp and the host's differ in the last
bit, hence the one-step allowance.)  The reference-named entry points start
from the oracle's initial state exactly.
"""
import ctypes
import os
import shutil

import numpy as np
import pytest

import cbrkc
import orc
from test_oracle_gr import gr_pop

pytestmark = pytest.mark.gpu

COMP = 578


def test_gr_cells_match_oracle(ref_data):
    v0 = np.array([-70.0, -72.0])
    amp = np.array([0.012e-3, 0.02e-3])
    stim = (16, 2.0, 1e9, amp)
    p = cbrkc.Population(cbrkc.GR, 2, 1, morph=os.path.join(ref_data, "Gr.txt"), v0=v0, stim=stim)
    o = gr_pop(ref_data, 2, 1, v0=v0)
    o.set_stim(16, 2.0, 1e9, amp)
    print("GR 2 cells, 240 macro steps, GPU vs oracle", flush=True)
    prev_g, prev_o = np.full(2, 100.0), np.full(2, 100.0)
    sg, so = [], []
    for k in range(240):
        p.set_current(k * 0.125)
        p.advance()
        o.set_current(k * 0.125)
        o.advance()
        vg, vo = p.read(cbrkc.F_VNEW), o.get(orc.VNEW)
        if k == 0:
            assert np.max(np.abs(vg - vo)) <= 1e-9
            # gates and Na occupancies after the 44 ion updates of the first
            # macro step, to the oracle's own spread under a one-ulp change of
            # the initial voltages (3.4e-8 for the Na states, which start out
            # of equilibrium; tests/test_oracle_gr.py::test_first_step_spread)
            for g in range(20):
                np.testing.assert_allclose(p.read(cbrkc.F_GATE0 + g), o.get(orc.ION0 + g), rtol=0, atol=1e-7)
        for v, prev, out in ((vg[77::COMP], prev_g, sg), (vo[77::COMP], prev_o, so)):
            for c in np.nonzero((v > 10.0) & (prev <= 10.0))[0]:
                out.append((0.125 * (k + 1), int(c)))
            prev[:] = v
    print(f"GR: spikes on the GPU {sg}, in the oracle {so}", flush=True)
    assert len(so) >= 4 and len(sg) == len(so)
    for (tg, cg), (to, co) in zip(sg, so):
        assert cg == co and abs(tg - to) <= 0.125
    print("GR stats (steps, rejects, stages, rho iterations, ion updates, macro steps):", p.stats(), flush=True)


def test_gr_reference_entry_points(ref_data, tmp_path, monkeypatch):
    shutil.copy(os.path.join(ref_data, "Gr.txt"), tmp_path / "Gr.txt")
    monkeypatch.chdir(tmp_path)
    ctypes.CDLL(None).srand(0)
    p = cbrkc.gr_initialize(2, 1, "RKC")
    o = gr_pop(ref_data, 2, 1, seed=0)
    np.testing.assert_array_equal(p.read(cbrkc.F_V), o.get(orc.V))
    for g in range(20):
        np.testing.assert_array_equal(p.read(cbrkc.F_GATE0 + g), o.get(orc.ION0 + g))
    prev = np.full(2, 100.0)
    for k in range(4):
        cbrkc.gr_set_current(p, k * 0.125)
        cbrkc.gr_solve_update_v(p)
        o.set_current(k * 0.125)
        o.advance()
        v = p.read(cbrkc.F_VNEW)
        np.testing.assert_allclose(v, o.get(orc.VNEW), rtol=0, atol=1e-8)
        cbrkc.gr_output_file(p, v, prev, (k + 1) * 0.125, "gr_v_by_RKC.csv", "gr_raster_by_RKC.csv")
    rows = (tmp_path / "gr_v_by_RKC.csv").read_text().splitlines()
    assert len(rows) == 4 and rows[0].startswith("0.125000,") and rows[0].count(",") == 3
    cbrkc.gr_finalize(p)
