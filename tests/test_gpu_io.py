"""Inferior-olive population on the GPU (global FP64 engine with grid gap
junctions) through the C-ABI.

* The reference's own io_raster_by_RKC.csv (8 x 8 grid, libc rand() seeded
  with 0, io_set_current's 800-820 ms pulse) is reproduced exactly.
* A 3 x 3 grid with a pulse moved to 2-12 ms follows the oracle compartment
  by compartment: after the first macro step (no ion update changes b
  before the first accepted steps) to 1e-9 mV, over 240 macro steps with
  spikes to 1e-6 mV.
"""
import ctypes
import os
import shutil

import numpy as np
import pytest

import cbrkc
import orc
from test_oracle_io import io_pop, read_raster

pytestmark = pytest.mark.gpu


def test_reference_io_raster_on_gpu(ref_data, tmp_path):
    shutil.copy(os.path.join(ref_data, "IO.txt"), tmp_path / "IO.txt")
    ctypes.CDLL(None).srand(0)
    print("IO 8x8 to 810 ms on the global engine", flush=True)
    p = cbrkc.Population(cbrkc.IO, 8, 8, morph=str(tmp_path / "IO.txt"))
    rec = p.run(6480, 0.0, 2)
    prev = np.full(64, 100.0)
    spikes = []
    for k in range(rec.shape[0]):
        v = rec[k]
        for c in np.nonzero((v > 10.0) & (prev <= 10.0))[0]:
            spikes.append((0.125 * (k + 1), int(c)))
        prev = v
    assert spikes == read_raster(os.path.join(ref_data, "io_raster_by_RKC.csv"))


def test_io_grid_matches_oracle(ref_data):
    n = 9
    v0 = np.linspace(-64.0, -56.0, n)
    stim = (0, 2.0, 12.0, 5.5)
    p = cbrkc.Population(cbrkc.IO, 3, 3, morph=os.path.join(ref_data, "IO.txt"), v0=v0, stim=stim)
    o = io_pop(ref_data, 3, 3, v0=v0)
    orc.lib().or_set_stim(o.h, 0, 2.0, 12.0, orc._dp(np.full(n, 5.5)))
    worst = 0.0
    for k in range(240):
        p.set_current(k * 0.125)
        p.advance()
        o.set_current(k * 0.125)
        o.advance()
        d = np.max(np.abs(p.read(cbrkc.F_VNEW) - o.get(orc.VNEW)))
        if k == 0:
            assert d <= 1e-9, d
        worst = max(worst, d)
    print(f"IO 3x3 vs oracle: max |dV| over 240 macro steps {worst:.3e} mV", flush=True)
    assert worst <= 1e-6
    so, sp = o.stats(), p.stats()
    assert (so["steps"], so["rejects"], so["ion_updates"]) == (sp["steps"], sp["rejects"], sp["ion_updates"])
