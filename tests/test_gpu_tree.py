"""The exact per-cell tree solve of the CN and BE engines (CB_ENGINE_TREE_FP64:
Hines elimination per cell, one wavefront per cell, in place of the
reference's 100-iteration CG / BiCG) through the C-ABI, against the oracle's
exact solve (or_cn_set_direct, itself pinned against numpy's dense solve).

The two eliminate in different orders (the oracle in leaf order, the GPU
level by level from each tree's centre) and differ by the device's and the
host's exp / log, so they agree to rounding, and, unlike the truncated CG,
the exact solve does not amplify last-bit differences: the Purkinje cell
stays within 1e-9 mV of the oracle after every output interval on both
paths; the granule, Golgi and olive populations agree to 1e-11 mV after the
first interval and fire the same spikes within one output interval.  Where
the reference's CG converges (a Golgi cell) the tree solve and the Krylov
engine agree to 1e-9 mV; on the Purkinje tree they differ by the CG's
truncation.  Populations that are not uncoupled trees are refused.
"""
import os

import numpy as np
import pytest

import cbrkc
import orc
from test_oracle_direct import L as LD
from test_gpu_be import spikes, same_spikes

pytestmark = pytest.mark.gpu


def oracle(pop, solver):
    lib = LD()
    assert (lib.or_cn_init(pop.h) if solver == cbrkc.CN else lib.or_be_init(pop.h)) == 0
    assert lib.or_cn_set_direct(pop.h, 1) == 0
    return pop


def run_pair(p, o, solver, steps, rec, comp, first_tol, every_tol=None):
    adv = LD().or_cn_advance if solver == cbrkc.CN else LD().or_be_advance
    tg, to = [], []
    worst = 0.0
    for k in range(steps):
        t = k * 0.125
        l_dt = 0.0
        while l_dt < 0.125:
            p.set_current(t + l_dt)
            p.advance()
            l_dt += 0.025
        adv(o.h, t)
        vg, vo = p.read(cbrkc.F_V), o.get(orc.V)
        d = np.max(np.abs(vg - vo))
        worst = max(worst, d)
        if k == 0:
            print(f"first interval: max |GPU - oracle| = {d:.3e} mV", flush=True)
            assert d <= first_tol
        if every_tol is not None:
            assert d <= every_tol, (k, d)
        tg.append(vg[rec::comp])
        to.append(vo[rec::comp])
    print(f"worst interval: max |GPU - oracle| = {worst:.3e} mV", flush=True)
    return np.array(tg), np.array(to)


@pytest.mark.parametrize("solver", [cbrkc.CN, cbrkc.BE])
def test_tree_pkj_reference_cell(ref_data, solver):
    v0 = np.array([-65.0])
    stim = (1599, 0.5, 1e9, 2e-3)
    p = cbrkc.Population(cbrkc.PKJ, 1, 1, solver=solver, engine=cbrkc.ENGINE_TREE_FP64,
                         morph=os.path.join(ref_data, "Pkj.txt"), v0=v0, stim=stim)
    o = oracle(orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=v0, libc_rand=False), solver)
    o.set_stim(*stim)
    print(f"tree solve, PKJ 1 cell, {'CN' if solver == cbrkc.CN else 'BE'}, 80 output intervals", flush=True)
    run_pair(p, o, solver, 80, 1599, 1600, 1e-11, every_tol=1e-9)
    st = p.stats()
    assert st["steps"] == 400 and st["stages"] == 0, st      # no Krylov iterations


@pytest.mark.parametrize("solver", [cbrkc.CN, cbrkc.BE])
def test_tree_gr_cells(ref_data, solver):
    from test_oracle_gr import gr_pop
    v0 = np.array([-70.0, -72.0])
    amp = np.array([0.012e-3, 0.02e-3])
    p = cbrkc.Population(cbrkc.GR, 2, 1, solver=solver, engine=cbrkc.ENGINE_TREE_FP64,
                         morph=os.path.join(ref_data, "Gr.txt"), v0=v0, stim=(16, 2.0, 1e9, amp))
    o = oracle(gr_pop(ref_data, 2, 1, v0=v0), solver)
    o.set_stim(16, 2.0, 1e9, amp)
    print(f"tree solve, GR 2 cells, {'CN' if solver == cbrkc.CN else 'BE'}, 96 output intervals", flush=True)
    tg, to = run_pair(p, o, solver, 96, 77, 578, 1e-11)
    sg, so = spikes(tg, 10.0), spikes(to, 10.0)
    print(f"spikes GPU {sg} oracle {so}", flush=True)
    assert len(so) >= 2 and same_spikes(sg, so)


def test_tree_io_grid(ref_data):
    from test_oracle_io import io_pop
    n = 9
    v0 = np.linspace(-64.0, -56.0, n)
    p = cbrkc.Population(cbrkc.IO, 3, 3, solver=cbrkc.CN, engine=cbrkc.ENGINE_TREE_FP64,
                         morph=os.path.join(ref_data, "IO.txt"), v0=v0, stim=(0, 2.0, 12.0, 5.5))
    o = oracle(io_pop(ref_data, 3, 3, v0=v0), cbrkc.CN)
    orc.lib().or_set_stim(o.h, 0, 2.0, 12.0, orc._dp(np.full(n, 5.5)))
    print("tree solve, IO 3x3 (gap currents on the right-hand side), CN, 240 output intervals", flush=True)
    tg, to = run_pair(p, o, cbrkc.CN, 240, 2, 3, 1e-11)
    sg, so = spikes(tg, 10.0), spikes(to, 10.0)
    print(f"spikes GPU {sg} oracle {so}", flush=True)
    assert len(so) >= 3 and same_spikes(sg, so)


def test_tree_go_cell_and_the_converged_krylov_engine(ref_data):
    from test_oracle_go import go_pop
    v0 = np.array([-60.0])
    p = cbrkc.Population(cbrkc.GO, 1, 1, solver=cbrkc.CN, engine=cbrkc.ENGINE_TREE_FP64,
                         morph=os.path.join(ref_data, "Go_131.txt"), v0=v0)
    q = cbrkc.Population(cbrkc.GO, 1, 1, solver=cbrkc.CN, morph=os.path.join(ref_data, "Go_131.txt"), v0=v0)
    o = oracle(go_pop(ref_data, 1, 1, v0=v0), cbrkc.CN)
    print("tree solve, GO 1 cell, CN, 16 output intervals against the oracle and the Krylov engine", flush=True)
    run_pair(p, o, cbrkc.CN, 16, 30, 131, 1e-11, every_tol=1e-9)
    for k in range(16):
        l_dt = 0.0
        while l_dt < 0.125:
            q.set_current(k * 0.125 + l_dt)
            q.advance()
            l_dt += 0.025
    d = np.max(np.abs(p.read(cbrkc.F_V) - q.read(cbrkc.F_V)))
    print(f"tree vs Krylov engine: {d:.3e} mV ({q.stats()['stages']} CG iterations in 80 steps)", flush=True)
    assert q.stats()["stages"] < 8000 and d <= 1e-9


def test_tree_and_truncated_cg_differ_on_the_purkinje_tree(ref_data):
    v0 = np.array([-65.0])
    mk = lambda e: cbrkc.Population(cbrkc.PKJ, 1, 1, solver=cbrkc.CN, engine=e,
                                    morph=os.path.join(ref_data, "Pkj.txt"), v0=v0)
    p, q = mk(cbrkc.ENGINE_TREE_FP64), mk(cbrkc.ENGINE_GLOBAL_FP64)
    for k in range(8):
        for x in (p, q):
            x.set_current(k * 0.025)
            x.advance()
    d = np.max(np.abs(p.read(cbrkc.F_V) - q.read(cbrkc.F_V)))
    print(f"PKJ CN after 8 steps: exact vs 100-iteration CG max |dV| = {d:.3e} mV", flush=True)
    assert 0.0 < d < 1.0


def test_tree_solve_refusals(ref_data):
    with pytest.raises(cbrkc.CbError):        # gap junctions between the Purkinje cells close cycles
        cbrkc.Population(cbrkc.PKJ, 2, 1, solver=cbrkc.CN, engine=cbrkc.ENGINE_TREE_FP64,
                         morph=os.path.join(ref_data, "Pkj.txt"), gj=os.path.join(ref_data, "Pkj_GJ.csv"),
                         v0=[-65.0, -64.0])
    with pytest.raises(cbrkc.CbError):        # the Golgi grid's gap junctions
        cbrkc.Population(cbrkc.GO, 2, 2, solver=cbrkc.BE, engine=cbrkc.ENGINE_TREE_FP64,
                         morph=os.path.join(ref_data, "Go_131.txt"), v0=[-60.0] * 4)
    with pytest.raises(cbrkc.CbError):        # RKC has no linear solve to replace
        cbrkc.Population(cbrkc.GR, 1, 1, solver=cbrkc.RKC, engine=cbrkc.ENGINE_TREE_FP64,
                         morph=os.path.join(ref_data, "Gr.txt"), v0=[-70.0])


def test_tree_finer_step_pkj(ref_data):
    """step_ms = 2^-7 ms (16 CN steps per output interval) against the oracle
    with the same step."""
    import ctypes
    LD().or_cn_set_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
    v0 = np.array([-65.0])
    stim = (1599, 0.5, 1e9, 2e-3)
    dt = 2.0 ** -7
    p = cbrkc.Population(cbrkc.PKJ, 1, 1, solver=cbrkc.CN, engine=cbrkc.ENGINE_TREE_FP64,
                         morph=os.path.join(ref_data, "Pkj.txt"), v0=v0, stim=stim, step_ms=dt)
    o = oracle(orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=v0, libc_rand=False), cbrkc.CN)
    assert LD().or_cn_set_dt(o.h, dt) == 0
    o.set_stim(*stim)
    print("tree solve, PKJ 1 cell, CN with 2^-7 ms steps, 16 output intervals", flush=True)
    worst = 0.0
    for k in range(16):
        p.run(1, k * 0.125, 1599, record=False)
        LD().or_cn_advance(o.h, k * 0.125)
        d = np.max(np.abs(p.read(cbrkc.F_V) - o.get(orc.V)))
        worst = max(worst, d)
        assert d <= 1e-9, (k, d)
    print(f"worst interval: {worst:.3e} mV", flush=True)
    assert p.stats()["steps"] == 256


def test_step_validation(ref_data):
    with pytest.raises(cbrkc.CbError):        # 0.125 / 0.03 is not a whole number of steps
        cbrkc.Population(cbrkc.GR, 1, 1, solver=cbrkc.CN, morph=os.path.join(ref_data, "Gr.txt"), v0=[-70.0],
                         step_ms=0.03)
    with pytest.raises(cbrkc.CbError):        # RKC chooses its own steps
        cbrkc.Population(cbrkc.GR, 1, 1, solver=cbrkc.RKC, morph=os.path.join(ref_data, "Gr.txt"), v0=[-70.0],
                         step_ms=2.0 ** -7)


def test_tree_autotune_picks_a_solve_and_leaves_the_run_unchanged(ref_data):
    # granule cells are 9 unbranched runs: both the solve by runs and the
    # level walk apply; the autotuner times both on copies of the current
    # system, keeps the faster and changes no result (the two are
    # bit-identical, and the trial does not touch the population's state)
    kw = dict(solver=cbrkc.CN, engine=cbrkc.ENGINE_TREE_FP64, morph=os.path.join(ref_data, "Gr.txt"),
              v0=np.array([-70.0, -65.0]), stim=(16, 0.0, 1e9, 0.015e-3))
    print("case: tree-solve autotune on 2 granule cells, CN, 8 output intervals", flush=True)
    p = cbrkc.Population(cbrkc.GR, 2, 1, **kw)
    q = cbrkc.Population(cbrkc.GR, 2, 1, **kw)
    a = p.run(2, 0.0, 16)
    b = q.run(2, 0.0, 16)
    assert q.autotune(16) == 64
    m = q.mapping()
    print(m, flush=True)
    assert "trial" in m and ("by unbranched runs" in m or "level by level" in m)
    a = np.vstack([a, p.run(6, 0.25, 16)])
    b = np.vstack([b, q.run(6, 0.25, 16)])
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(p.read(cbrkc.F_V), q.read(cbrkc.F_V))
