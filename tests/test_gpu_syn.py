"""Synapses on the GPU (cb_*_create / cb_*_synapse_update and the coupling
inside every engine) against the oracle's network, through the C-ABI.

One small network per solver, run the way main.cu runs it with network_sim
on: per macro step the granule, Golgi and Purkinje populations advance
(five 0.025 ms sub-steps for CN and BE), then the synapse sets update from
the presynaptic voltages.  Granule cells get a step current so they fire
near 6 ms (granule -> Golgi and granule -> Purkinje events), the Golgi cell a
step current from 6 ms so it fires too (Golgi -> granule events), and the
mossy fibres fire from the Philox draws, identical on both sides.

Bars: the first macro step within 1e-9 mV in every population; every
population's spikes (granule compartment 77 and Golgi 30 through +10 mV,
Purkinje 1599 through -10 mV) in equal numbers per cell with times within
one macro step; the mossy-fibre conductances, which depend on the draws
alone, equal at every step; every set's conductances within 1e-6 relative at
the end when the spike trains agree to the macro step (they do on these
workloads).  The granule -> Purkinje set also drives a Purkinje cell on the
per-cell engines, checked against the global FP64 engine (its own bars in
the test's docstring).
"""
import ctypes
import os

import numpy as np
import pytest

import cbrkc
import orc
from test_oracle_syn import L as LS, Syn, pkj_x
from test_oracle_cn import L as LC
from test_oracle_be import L as LB

pytestmark = pytest.mark.gpu

STEPS = 96


def spikes(trace, thr):
    prev = np.full(trace.shape[1], 100.0)
    out = []
    for k in range(trace.shape[0]):
        v = trace[k]
        for c in np.nonzero((v > thr) & (prev <= thr))[0]:
            out.append((k + 1, int(c)))
        prev = v
    return out


def network(ref_data, solver, gr_shape):
    n_gr = gr_shape[0] * gr_shape[1]
    amp = np.linspace(0.012e-3, 0.02e-3, n_gr)
    gr = cbrkc.Population(cbrkc.GR, *gr_shape, solver=solver, morph=os.path.join(ref_data, "Gr.txt"),
                          v0=np.full(n_gr, -70.0), stim=(16, 2.0, 1e9, amp))
    go = cbrkc.Population(cbrkc.GO, 1, 1, solver=solver, morph=os.path.join(ref_data, "Go_131.txt"), v0=[-60.0],
                          stim=(30, 6.0, 1e9, 3e-3))
    pk = cbrkc.Population(cbrkc.PKJ, 1, 1, solver=solver, morph=os.path.join(ref_data, "Pkj.txt"), v0=[-65.0])
    loc = os.path.join(ref_data, "Pkj_location_info.csv")
    syn = dict(mf=cbrkc.mfgr_create(gr, 99), gg=cbrkc.gogr_create(go, gr), rg=cbrkc.grgo_create(gr, go),
               rp=cbrkc.grpkj_create(gr, pk, loc, 5), ml=cbrkc.mlipkj_create(pk, gr, 7))

    from test_oracle_gr import gr_pop
    from test_oracle_go import go_pop
    ogr = gr_pop(ref_data, *gr_shape, v0=np.full(n_gr, -70.0))
    ogo = go_pop(ref_data, 1, 1, v0=[-60.0])
    opk = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=[-65.0], libc_rand=False)
    ogr.set_stim(16, 2.0, 1e9, amp)
    ogo.set_stim(30, 6.0, 1e9, np.array([3e-3]))
    for o in (ogr, ogo, opk):
        if solver == cbrkc.CN:
            assert LC().or_cn_init(o.h) == 0
        elif solver == cbrkc.BE:
            assert LB().or_be_init(o.h) == 0
    osyn = dict(mf=Syn(LS().or_mfgr_create(n_gr, 99)), gg=Syn(LS().or_gogr_create(ogo.h, ogr.h)),
                rg=Syn(LS().or_grgo_create(ogr.h, ogo.h)),
                rp=Syn(LS().or_grpkj_create(ogr.h, opk.h, orc._dp(pkj_x(ref_data)), 5)))
    for o, k in ((ogr, "mf"), (ogr, "gg"), (ogo, "rg"), (opk, "rp")):
        assert LS().or_pop_attach_syn(o.h, osyn[k].h) == 0
    for k in ("mf", "gg", "rg", "rp"):
        gp, gq = syn[k].comp()
        op, oq = osyn[k].comp()
        np.testing.assert_array_equal(gp, op)
        np.testing.assert_array_equal(gq, oq)
    return (gr, go, pk), syn, (ogr, ogo, opk), osyn


def macro_gpu(p, solver, t):
    if solver == cbrkc.RKC:
        p.set_current(t)
        p.advance()
        return
    l_dt = 0.0
    while l_dt < 0.125:
        p.set_current(t + l_dt)
        p.advance()
        l_dt += 0.025


def macro_oracle(o, solver, t):
    if solver == cbrkc.RKC:
        o.set_current(t)
        o.advance()
    elif solver == cbrkc.CN:
        LC().or_cn_advance(o.h, t)
    else:
        LB().or_be_advance(o.h, t)


def run(ref_data, solver, gr_shape):
    pops, syn, opops, osyn = network(ref_data, solver, gr_shape)
    lib = orc.lib()
    lib.or_net_update.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_double]
    field = cbrkc.F_VNEW if solver == cbrkc.RKC else cbrkc.F_V
    ofield = orc.VNEW if solver == cbrkc.RKC else orc.V
    rec = {"gr": (77, 578, 10.0), "go": (30, 131, 10.0), "pk": (1599, 1600, -10.0)}
    tg = {k: [] for k in rec}
    to = {k: [] for k in rec}
    for k in range(STEPS):
        t = k * 0.125
        for p, o in zip(pops, opops):
            macro_gpu(p, solver, t)
            macro_oracle(o, solver, t)
        cbrkc.gr_synapse_update(t, syn["mf"], syn["gg"])
        cbrkc.go_synapse_update(t, syn["rg"])
        cbrkc.pkj_synapse_update(t, syn["rp"], syn["ml"])
        lib.or_net_update(osyn["mf"].h, osyn["gg"].h, osyn["rg"].h, osyn["rp"].h, t)
        for name, p, o in zip(("gr", "go", "pk"), pops, opops):
            vg, vo = p.read(field), o.get(ofield)
            if k == 0:
                d = np.max(np.abs(vg - vo))
                print(f"{name}: first macro step max |GPU - oracle| = {d:.3e} mV", flush=True)
                assert d <= 1e-9, name
            c, comp, _ = rec[name]
            tg[name].append(vg[c::comp])
            to[name].append(vo[c::comp])
        np.testing.assert_array_equal(syn["mf"].read(cbrkc.SYN_VAL), osyn["mf"].get(4))
    same_steps = True
    for name, (_, _, thr) in rec.items():
        sg, so = spikes(np.array(tg[name]), thr), spikes(np.array(to[name]), thr)
        print(f"{name}: spikes (step, cell) GPU {sg} oracle {so}", flush=True)
        assert len(sg) == len(so), name
        key = lambda x: (x[1], x[0])
        for a, b in zip(sorted(sg, key=key), sorted(so, key=key)):
            assert a[1] == b[1] and abs(a[0] - b[0]) <= 1, name
        same_steps &= sorted(sg) == sorted(so)
    assert len(spikes(np.array(to["gr"]), 10.0)) > 0 and len(spikes(np.array(to["go"]), 10.0)) > 0
    if same_steps:
        for k in ("mf", "gg", "rg", "rp"):
            g, o = syn[k].read(cbrkc.SYN_VAL), osyn[k].get(4)
            assert np.max(o) > 0.0, k
            np.testing.assert_allclose(g, o, rtol=1e-6, atol=1e-6 * np.max(o), err_msg=k)
    for s in syn.values():
        s.close()
    return same_steps


def test_network_cn(ref_data):
    print("network: GR 2x4, GO 1x1, PKJ 1x1, CN, 12 ms", flush=True)
    assert run(ref_data, cbrkc.CN, (2, 4))


def test_network_be(ref_data):
    print("network: GR 2x4, GO 1x1, PKJ 1x1 (no BE coupling), BE, 12 ms", flush=True)
    assert run(ref_data, cbrkc.BE, (2, 4))


def test_network_rkc(ref_data):
    print("network: GR 2x2, GO 1x1, PKJ 1x1, RKC, 12 ms", flush=True)
    run(ref_data, cbrkc.RKC, (2, 2))


@pytest.mark.parametrize("engine", [cbrkc.ENGINE_CELL_FP64, cbrkc.ENGINE_CELL_MIXED])
def test_granule_purkinje_synapses_on_the_per_cell_engines(ref_data, engine):
    """A Purkinje cell on a per-cell engine driven by granule -> Purkinje
    synapses from a 2 x 2 granule population (global engine, step current,
    spikes near 6 ms), against the same Purkinje cell on the global FP64
    engine with its own set built from the same granule cells and seed (so
    the same contacts and, updated from the same presynaptic voltages, the
    same conductances).  For one cell the per-cell controller is the global
    one up to the order of its norm reductions, so the two take different
    step sequences within the 1e-6 tolerance: identical spike counts, spike
    times within 0.01 ms, subthreshold within 0.03 mV (FP64; measured 0.014 mV
    with the synaptic input moving the soma by 7 mV) / 0.1 mV (mixed).  A
    third, uncoupled Purkinje cell shows the synapses act."""
    amp = np.linspace(0.012e-3, 0.02e-3, 4)
    gr = cbrkc.Population(cbrkc.GR, 2, 2, morph=os.path.join(ref_data, "Gr.txt"), v0=np.full(4, -70.0),
                          stim=(16, 2.0, 1e9, amp))
    mk = lambda e: cbrkc.Population(cbrkc.PKJ, 1, 1, engine=e, morph=os.path.join(ref_data, "Pkj.txt"), v0=[-65.0])
    pc, pg, pn = mk(engine), mk(cbrkc.ENGINE_GLOBAL_FP64), mk(cbrkc.ENGINE_GLOBAL_FP64)
    loc = os.path.join(ref_data, "Pkj_location_info.csv")
    sc = (cbrkc.grpkj_create(gr, pc, loc, 5), cbrkc.mlipkj_create(pc, gr, 7))
    sg = (cbrkc.grpkj_create(gr, pg, loc, 5), cbrkc.mlipkj_create(pg, gr, 7))
    np.testing.assert_array_equal(sc[0].comp()[1], sg[0].comp()[1])
    print(f"granule -> Purkinje ({sc[0].n} synapses), PKJ on engine {engine} vs the global FP64 engine, 12 ms",
          flush=True)
    tc, tg, tn, tgr = [], [], [], []
    for k in range(96):
        t = k * 0.125
        for p in (gr, pc, pg, pn):
            p.set_current(t)
            p.advance()
        cbrkc.pkj_synapse_update(t, *sc)
        cbrkc.pkj_synapse_update(t, *sg)
        tc.append(pc.read(cbrkc.F_VNEW)[1599])
        tg.append(pg.read(cbrkc.F_VNEW)[1599])
        tn.append(pn.read(cbrkc.F_VNEW)[1599])
        tgr.append(gr.read(cbrkc.F_VNEW)[77::578])
    np.testing.assert_allclose(sc[0].read(cbrkc.SYN_VAL), sg[0].read(cbrkc.SYN_VAL), rtol=0, atol=0)
    assert np.max(sg[0].read(cbrkc.SYN_VAL)) > 0.0 and len(spikes(np.array(tgr), 10.0)) > 0
    from trace_metrics import compare
    a, b = np.array(tg)[:, None], np.array(tc)[:, None]
    counts, dt, sub = compare(a, b)
    effect = np.max(np.abs(np.array(tg) - np.array(tn)))
    print(f"spikes {counts.tolist()}, max spike-time diff {dt.max():.3e} ms, subthreshold {sub.max():.3e} mV; "
          f"synaptic effect on the soma {effect:.3e} mV", flush=True)
    assert effect > 1e-3
    assert np.all(counts[:, 0] == counts[:, 1]) and dt.max() <= 0.01
    assert sub.max() <= (0.03 if engine == cbrkc.ENGINE_CELL_FP64 else 0.1)
    for x in sc + sg:
        x.close()


def test_synapses_rejected_by_populations_without_them(ref_data):
    gr = cbrkc.Population(cbrkc.GR, 1, 1, morph=os.path.join(ref_data, "Gr.txt"), v0=[-70.0])
    io = cbrkc.Population(cbrkc.IO, 1, 1, morph=os.path.join(ref_data, "IO.txt"), v0=[-60.0])
    with pytest.raises(cbrkc.CbError):
        cbrkc.grgo_create(gr, io)             # not a Golgi population
    s = cbrkc.gogr_create(None, gr)           # an empty set couples nothing
    assert s.n == 0
    s.close()
