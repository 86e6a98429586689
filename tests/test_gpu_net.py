"""The fixed in-degree network on the GPU (cb_net_*) against the oracle's
network (oracle/or_net.c), through the C-ABI.

Delivery is checked bit for bit on its own: the GPU's raster, replayed
through the oracle's delivery (or_net_deliver_given: decay, arrivals after
the delay, mossy-fibre draws, each slot's sum in synapse order), must give
every (compartment, receptor) conductance of the GPU exactly.  The
population dynamics are checked against the oracle's populations.

* FP64, against the oracle: 64 granule chains (4 compartments), 8 Golgi
  trees and 2 Purkinje trees (synthetic, seeded; one granule and one Golgi
  geometry shared by their cells, the reference's format, which the
  oracle's readers take), every population on the global FP64 RKC engine
  (the reference's semantics; the oracle runs the same populations), mossy
  fibres at 5 Hz, 1 ms delay, 30 ms.  Bars: delivery exact (above);
  voltages within 1e-6 mV of the oracle after the first interval; the
  raster equal to the oracle's in spike count per cell with every spike
  time within two intervals.  Measured cause of the spread: until the first
  synaptic event the GPU and the oracle take the same steps (equal step,
  reject and ion-update counts, voltages within rounding); at the first
  mossy-fibre arrival, with bit-identical conductances, the
  population-global controller's accept / reject path forks on a rounding
  difference (granule steps 1682 vs 1677 after 6 intervals), which moves
  the operator-splitting points and so the voltages by up to ~0.01 mV; over
  30 ms one Golgi spike moves by one interval and one Purkinje spike by two.
* Mixed against FP64: the same network with 8 Purkinje cells on the per-cell
  mixed engine and on the per-cell FP64 engine (granule and Golgi cells on
  the global FP64 engine in both runs), 50 ms.  The Purkinje cells have no
  outgoing synapses, so the granule and Golgi rasters and every conductance
  must be identical; the Purkinje rasters must agree in spike count per cell
  with times within one interval.
* Full size: 2048 granule, 512 Golgi, 512 Purkinje cells, exactly 780404
  synapses, 24 intervals: counts, delivery exact against the oracle's
  replay, finite voltages.
"""
import ctypes
import os

import numpy as np
import pytest

import cbrkc
import net_oracle as NO

pytestmark = pytest.mark.gpu

H = 0.125


def gpu_pops(files, v0, pkj_engine=cbrkc.ENGINE_GLOBAL_FP64, stable=False):
    gr_f, go_f, pk_f = files
    g0, o0, p0 = v0
    gr = cbrkc.Population(cbrkc.GR, len(g0), 1, morph=gr_f, comp=4, v0=g0, stim="none")
    go = cbrkc.Population(cbrkc.GO, len(o0), 1, morph=go_f, v0=o0, stim="none")
    kw = dict(cbrkc.STABLE) if stable else {}
    pk = cbrkc.Population(cbrkc.PKJ, len(p0), 1, morph=pk_f, gj="", v0=p0, atol=1e-6, rtol=1e-6, stim="none",
                          engine=pkj_engine, **kw)
    return gr, go, pk


def v0_of(n, seed=1):
    rng = np.random.default_rng(seed)
    return (-70.0 + rng.normal(0, 1, n[0]), -60.0 + rng.normal(0, 1, n[1]), -68.0 + rng.normal(0, 1, n[2]))


def raster_set(t, p, c):
    return sorted(zip(np.rint(np.asarray(t) / H).astype(int).tolist(), np.asarray(p).tolist(),
                      np.asarray(c).tolist()))


def raster_bytes(t, p, c, n, nint):
    """[nint][granule, Golgi, Purkinje cells] spike bytes of a raster"""
    off = np.cumsum((0,) + tuple(n))
    out = np.zeros((nint, sum(n)), np.uint8)
    for ti, pi, ci in zip(t, p, c):
        out[int(round(ti / H)) - 1, off[pi] + ci] = 1
    return out


def replay_conductances(files, n, cfg, spikes, v0):
    """the oracle's delivery of a given raster: conductances after
    len(spikes) intervals, [population][receptor]"""
    lib = NO.L()
    opops = NO.oracle_pops(files, v0)
    h = lib.or_net_create(opops[0].h, opops[1].h, opops[2].h, ctypes.byref(NO.spec_of(cfg)))
    delay = int(round(cfg.delay_ms / H))
    P = ctypes.POINTER(ctypes.c_ubyte)
    for k in range(len(spikes)):
        j = k - 1 - delay
        lib.or_net_deliver_given(h, spikes[j].ctypes.data_as(P) if j >= 0 else None)
    out = []
    for p in range(3):
        row = []
        for r in (0, 1):
            g = np.zeros(opops[p].nc)
            lib.or_net_conductance(h, p, r, g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            row.append(g)
        out.append(row)
    lib.or_net_free(h)
    return out


def same_up_to(a, b, pops, tol):
    """equal spike counts per (population, cell), times within tol intervals"""
    for p in pops:
        cells = {s[2] for s in a + b if s[1] == p}
        for c in cells:
            ta = [s[0] for s in a if s[1] == p and s[2] == c]
            tb = [s[0] for s in b if s[1] == p and s[2] == c]
            if len(ta) != len(tb) or any(abs(x - y) > tol for x, y in zip(ta, tb)):
                return False, (p, c, ta, tb)
    return True, None


def test_network_fp64_against_the_oracle(tmp_path):
    n, nint = (64, 8, 2), 240
    files = NO.synth_files(tmp_path, *n, shared=True)
    v0 = v0_of(n)
    cfg = cbrkc.net_config(total=0)
    print(f"case: network {n} cells, FP64 global engines vs the oracle, {nint} intervals", flush=True)
    pops = gpu_pops(files, v0)
    net = cbrkc.Network(*pops, cfg=cfg)
    st = net.stats()
    assert st["synapses"] == [4 * n[0], 4 * n[0], 100 * n[1], 1392 * n[2]]
    marks = (1, 40, 100, nint)
    gpu_vk = {}
    done = 0
    for m in marks:
        net.run(m - done)
        done = m
        gpu_vk[m] = [p.read(cbrkc.F_VNEW) for p in pops]
        if m == 1:
            gpu_st1 = [p.stats() for p in pops]
    gpu_v1 = gpu_vk[1]
    gt, gp, gc = net.spikes()
    g_gpu = [[net.conductance(p, r) for r in (0, 1)] for p in range(3)]
    v_gpu = [p.read(cbrkc.F_VNEW) for p in pops]
    net.close()
    for p in pops:
        p.close()

    lib = NO.L()
    opops = NO.oracle_pops(files, v0)
    h = lib.or_net_create(opops[0].h, opops[1].h, opops[2].h, ctypes.byref(NO.spec_of(cfg)))
    stride = sum(n)
    sp = np.zeros((nint, stride), np.uint8)
    P = ctypes.POINTER(ctypes.c_ubyte)
    o_vk = {}
    done = 0
    for m in marks:
        assert lib.or_net_run(h, m - done, sp[done:m].ctypes.data_as(P)) == 0
        done = m
        o_vk[m] = [p.get(6) for p in opops]
        if m == 1:
            o_st1 = [p.stats() for p in opops]
    o_v1 = o_vk[1]
    g_or = []
    for p in range(3):
        row = []
        for r in (0, 1):
            out = np.zeros(opops[p].nc)
            lib.or_net_conductance(h, p, r, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            row.append(out)
        g_or.append(row)
    o_v = [p.get(6) for p in opops]
    lib.or_net_free(h)

    off = np.cumsum((0,) + n)
    ref = []
    for k in range(nint):
        for p in range(3):
            for c in np.nonzero(sp[k, off[p]:off[p + 1]])[0]:
                ref.append((k + 1, p, int(c)))
    got = raster_set(gt, gp, gc)
    print(f"spikes: gpu {len(got)}, oracle {len(ref)} (granule {sum(1 for s in ref if s[1] == 0)}, Golgi "
          f"{sum(1 for s in ref if s[1] == 1)}, Purkinje {sum(1 for s in ref if s[1] == 2)}); identical: "
          f"{got == sorted(ref)}", flush=True)
    for p in range(3):
        print(f"population {p}: |dV| after intervals " + ", ".join(
            f"{m}: {np.abs(gpu_vk[m][p] - o_vk[m][p]).max():.3g}" for m in marks) + " mV", flush=True)
    assert sum(1 for s in ref if s[1] == 0) > 0 and sum(1 for s in ref if s[1] == 1) > 0
    ok, why = same_up_to(got, sorted(ref), range(3), 2)
    assert ok, why
    for p in range(3):
        assert np.abs(gpu_v1[p] - o_v1[p]).max() < 1e-6
        for key in ("steps", "rejects", "ion_updates"):   # the same steps before any synaptic event
            assert gpu_st1[p][key] == o_st1[p][key], (p, key)
    # delivery: the oracle's replay of the GPU's raster gives its conductances exactly
    rep = replay_conductances(files, n, cfg, raster_bytes(gt, gp, gc, n, nint), v0)
    for p in range(3):
        for r in (0, 1):
            assert np.array_equal(g_gpu[p][r], rep[p][r]), (p, r)
    if got == sorted(ref):   # and with the oracle's own raster, the same conductances
        for p in range(3):
            for r in (0, 1):
                assert np.array_equal(g_gpu[p][r], g_or[p][r]), (p, r)


def test_network_mixed_purkinje_against_fp64(tmp_path):
    n, nint = (64, 8, 8), 400
    files = NO.synth_files(tmp_path, *n, seed=23)
    v0 = v0_of(n, seed=2)
    cfg = cbrkc.net_config(total=0)
    res = {}
    for eng in (cbrkc.ENGINE_CELL_FP64, cbrkc.ENGINE_CELL_MIXED):
        print(f"case: network {n} cells, Purkinje engine {eng}, {nint} intervals", flush=True)
        pops = gpu_pops(files, v0, pkj_engine=eng, stable=True)
        net = cbrkc.Network(*pops, cfg=cfg)
        net.run(nint)
        res[eng] = net.spikes()
        res[(eng, "g")] = [net.conductance(p, r) for p in (0, 1) for r in (0, 1)]
        net.close()
        for p in pops:
            p.close()
    a = raster_set(*res[cbrkc.ENGINE_CELL_FP64])
    b = raster_set(*res[cbrkc.ENGINE_CELL_MIXED])
    a_small = [s for s in a if s[1] < 2]
    b_small = [s for s in b if s[1] < 2]
    assert a_small == b_small
    for x, y in zip(res[(cbrkc.ENGINE_CELL_FP64, "g")], res[(cbrkc.ENGINE_CELL_MIXED, "g")]):
        assert np.array_equal(x, y)
    pk_a = [s for s in a if s[1] == 2]
    pk_b = [s for s in b if s[1] == 2]
    print(f"Purkinje spikes: FP64 {len(pk_a)}, mixed {len(pk_b)}; granule + Golgi {len(a_small)}", flush=True)
    assert len(pk_a) > 0
    ok, why = same_up_to(pk_a, pk_b, (2,), 1)
    assert ok, why


def test_network_with_a_crank_nicolson_granule_population(tmp_path):
    """Each population keeps its own solver: granule cells on Crank-Nicolson
    with the exact tree solve (five 0.025 ms sub-steps per interval, the
    synaptic conductance entering its system as the reference's CN path
    does), Golgi cells on the global FP64 RKC engine, Purkinje cells on the
    per-cell mixed engine.  Bars: delivery exact against the oracle's replay
    of the GPU raster; granule spikes present; voltages finite."""
    n, nint = (64, 8, 4), 160
    files = NO.synth_files(tmp_path, *n, seed=41)
    v0 = v0_of(n, seed=4)
    cfg = cbrkc.net_config(total=0, mf_rate_hz=40.0)
    print(f"case: network {n} cells, granule CN tree, {nint} intervals", flush=True)
    g0, o0, p0 = v0
    gr = cbrkc.Population(cbrkc.GR, n[0], 1, solver=cbrkc.CN, engine=cbrkc.ENGINE_TREE_FP64, morph=files[0], comp=4,
                          v0=g0, stim="none")
    go = cbrkc.Population(cbrkc.GO, n[1], 1, morph=files[1], v0=o0, stim="none")
    pk = cbrkc.Population(cbrkc.PKJ, n[2], 1, morph=files[2], gj="", v0=p0, atol=1e-6, rtol=1e-6, stim="none",
                          engine=cbrkc.ENGINE_CELL_MIXED, **cbrkc.STABLE)
    net = cbrkc.Network(gr, go, pk, cfg=cfg)
    net.run(nint)
    t, p, c = net.spikes()
    g = [[net.conductance(q, r) for r in (0, 1)] for q in range(3)]
    assert net.stats()["intervals"] == nint
    assert gr.stats()["macro_steps"] == nint
    for q in (gr, go, pk):
        assert np.all(np.isfinite(q.read(cbrkc.F_VNEW)))
    net.close()
    for q in (gr, go, pk):
        q.close()
    print(f"spikes: granule {int(np.sum(p == 0))}, Golgi {int(np.sum(p == 1))}, Purkinje {int(np.sum(p == 2))}",
          flush=True)
    assert np.sum(p == 0) > 0
    rep = replay_conductances(files, n, cfg, raster_bytes(t, p, c, n, nint), v0)
    for q in range(3):
        for r in (0, 1):
            assert np.array_equal(g[q][r], rep[q][r]), (q, r)


def test_network_full_size_780404_synapses(tmp_path):
    n, nint = (2048, 512, 512), 24
    files = NO.synth_files(tmp_path, *n, seed=31)
    v0 = v0_of(n, seed=3)
    cfg = cbrkc.net_config(mf_rate_hz=200.0)   # busier fibres: arrivals within the short run
    print(f"case: network {n} cells, 780404 synapses, {nint} intervals", flush=True)
    pops = gpu_pops(files, v0, pkj_engine=cbrkc.ENGINE_CELL_MIXED, stable=True)
    net = cbrkc.Network(*pops, cfg=cfg)
    st = net.stats()
    assert sum(st["synapses"]) == 780404
    assert st["delay"] == 8
    net.run(nint)
    assert net.stats()["intervals"] == nint
    t, p, c = net.spikes()
    g = [[net.conductance(q, r) for r in (0, 1)] for q in range(3)]
    for q in pops:
        assert np.all(np.isfinite(q.read(cbrkc.F_VNEW)))
    net.close()
    for q in pops:
        q.close()
    print(f"spikes: {len(t)} (granule {int(np.sum(p == 0))}, Golgi {int(np.sum(p == 1))}, Purkinje "
          f"{int(np.sum(p == 2))})", flush=True)
    assert g[0][0].any()   # the mossy fibres have fired
    rep = replay_conductances(files, n, cfg, raster_bytes(t, p, c, n, nint), v0)
    for q in range(3):
        for r in (0, 1):
            assert np.array_equal(g[q][r], rep[q][r]), (q, r)
