"""The Crank-Nicolson engine on the GPU (global CG / BiCG on the device)
through the C-ABI, against the oracle's CN path.

Both sides run the same algorithm with the same reduction order (256-wide
deterministic trees for every dot product), so they differ only by the
device's and the host's exp / log in the ion updates:

* one Purkinje cell of the reference morphology, 2 ms (80 CN steps, every
  solve the full 100 CG iterations, as in the reference): every compartment
  within 1e-8 mV of the oracle after every output interval;
* a 3 x 3 inferior-olive grid (BiCG, explicit gap currents) with a 10 ms
  pulse, 30 ms: the first interval to 1e-9 mV, then the same spike counts
  per cell within one output interval;
* a 2 x 2 Golgi grid (CG with gap junctions in the matrix), 60 ms, and two
  granule cells with a step current, 12 ms: first interval to 1e-9 mV, spike
  counts equal, times within one output interval.

Populations of at most 16 x 256 compartments (all of the above) run each
step's whole solve in one workgroup (k_kry_small); larger ones launch the
two halves of every iteration over the whole grid (k_kry_a / k_kry_b).  Both
reduce in the same order, so both match the oracle alike; the 2 x 4 granule
population (4 624 compartments, 19 blocks) checks the multi-block path.
"""
import ctypes
import os

import numpy as np
import pytest

import cbrkc
import orc
from test_oracle_cn import L as LC

pytestmark = pytest.mark.gpu


def oracle_cn(pop):
    assert LC().or_cn_init(pop.h) == 0
    return pop


def spikes(trace, thr):
    prev = np.full(trace.shape[1], 100.0)
    out = []
    for k in range(trace.shape[0]):
        v = trace[k]
        for c in np.nonzero((v > thr) & (prev <= thr))[0]:
            out.append((0.125 * (k + 1), int(c)))
        prev = v
    return out


def run_pair(p, o, steps, rec, comp, first_tol, every_tol=None):
    tg, to = [], []
    for k in range(steps):
        t = k * 0.125
        # one output interval: the reference's five CN sub-steps
        l_dt = 0.0
        while l_dt < 0.125:
            p.set_current(t + l_dt)
            p.advance()
            l_dt += 0.025
        LC().or_cn_advance(o.h, t)
        vg, vo = p.read(cbrkc.F_V), o.get(orc.V)
        if k == 0:
            assert np.max(np.abs(vg - vo)) <= first_tol
        if every_tol is not None:
            assert np.max(np.abs(vg - vo)) <= every_tol, (k, np.max(np.abs(vg - vo)))
        tg.append(vg[rec::comp])
        to.append(vo[rec::comp])
    return np.array(tg), np.array(to)


def test_cn_pkj_reference_cell(ref_data):
    v0 = np.array([-65.0])
    stim = (1599, 0.5, 1e9, 2e-3)
    p = cbrkc.Population(cbrkc.PKJ, 1, 1, solver=cbrkc.CN, morph=os.path.join(ref_data, "Pkj.txt"), v0=v0,
                         stim=stim)
    o = oracle_cn(orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=v0, libc_rand=False))
    o.set_stim(*stim)
    print("CN PKJ 1 cell, 16 output intervals", flush=True)
    run_pair(p, o, 16, 1599, 1600, 1e-9, every_tol=1e-8)
    st = p.stats()
    assert st["steps"] == 80 and st["stages"] == 8000, st   # 100 CG iterations per step, as the reference


def test_cn_io_grid(ref_data):
    from test_oracle_io import io_pop
    n = 9
    v0 = np.linspace(-64.0, -56.0, n)
    stim = (0, 2.0, 12.0, 5.5)
    p = cbrkc.Population(cbrkc.IO, 3, 3, solver=cbrkc.CN, morph=os.path.join(ref_data, "IO.txt"), v0=v0, stim=stim)
    o = oracle_cn(io_pop(ref_data, 3, 3, v0=v0))
    orc.lib().or_set_stim(o.h, 0, 2.0, 12.0, orc._dp(np.full(n, 5.5)))
    print("CN IO 3x3, 240 output intervals", flush=True)
    tg, to = run_pair(p, o, 240, 2, 3, 1e-9)
    sg, so = spikes(tg, 10.0), spikes(to, 10.0)
    print(f"CN IO: {len(sg)} spikes on the GPU, {len(so)} in the oracle", flush=True)
    assert len(so) > 0 and sorted(c for _, c in sg) == sorted(c for _, c in so)
    for c in range(n):
        a = [t for t, cc in sg if cc == c]
        b = [t for t, cc in so if cc == c]
        assert np.max(np.abs(np.array(a) - np.array(b)), initial=0.0) <= 0.125


def test_cn_go_grid_and_gr_cells(ref_data):
    from test_oracle_go import go_pop
    from test_oracle_gr import gr_pop
    v0 = np.array([-61.0, -60.5, -60.0, -59.5])
    p = cbrkc.Population(cbrkc.GO, 2, 2, solver=cbrkc.CN, morph=os.path.join(ref_data, "Go_131.txt"), v0=v0)
    o = oracle_cn(go_pop(ref_data, 2, 2, v0=v0))
    print("CN GO 2x2, 480 output intervals", flush=True)
    tg, to = run_pair(p, o, 480, 30, 131, 1e-9)
    sg, so = spikes(tg, 10.0), spikes(to, 10.0)
    print(f"CN GO: spikes GPU {sg} oracle {so}", flush=True)
    assert len(so) >= 4 and len(sg) == len(so)
    assert all(a[1] == b[1] and abs(a[0] - b[0]) <= 0.125 for a, b in zip(sorted(sg, key=lambda x: (x[1], x[0])),
                                                                        sorted(so, key=lambda x: (x[1], x[0]))))
    v0 = np.array([-70.0, -72.0])
    amp = np.array([0.012e-3, 0.02e-3])
    p = cbrkc.Population(cbrkc.GR, 2, 1, solver=cbrkc.CN, morph=os.path.join(ref_data, "Gr.txt"), v0=v0,
                         stim=(16, 2.0, 1e9, amp))
    o = oracle_cn(gr_pop(ref_data, 2, 1, v0=v0))
    o.set_stim(16, 2.0, 1e9, amp)
    print("CN GR 2 cells, 96 output intervals", flush=True)
    tg, to = run_pair(p, o, 96, 77, 578, 1e-9)
    sg, so = spikes(tg, 10.0), spikes(to, 10.0)
    print(f"CN GR: spikes GPU {sg} oracle {so}", flush=True)
    assert len(so) >= 2 and len(sg) == len(so)
    assert all(a[1] == b[1] and abs(a[0] - b[0]) <= 0.125 for a, b in zip(sg, so))


def test_cn_multiblock_gr_population(ref_data):
    from test_oracle_gr import gr_pop
    v0 = np.linspace(-72.0, -68.0, 8)
    p = cbrkc.Population(cbrkc.GR, 2, 4, solver=cbrkc.CN, morph=os.path.join(ref_data, "Gr.txt"), v0=v0)
    assert p.nc > 16 * 256
    o = oracle_cn(gr_pop(ref_data, 2, 4, v0=v0))
    print("CN GR 2x4 (multi-block Krylov launches), 4 output intervals", flush=True)
    run_pair(p, o, 4, 16, 578, 1e-9)
    st = p.stats()
    assert st["steps"] == 20 and 0 < st["stages"] <= 2000, st
