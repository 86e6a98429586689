"""The batched engine over an uneven mix of cell types (cb_mix_*), on the GPU.

* Granule (4-compartment synthetic chains), Golgi (synthetic 131-compartment
  trees) and olive cells on the small-cell lane-group path against the
  oracle, cell by cell: every cell is the reference's RKC on a population of
  one (or_gr_create_comp / or_go_create / or_io_create on that cell's block).
  Olive cells are coupled by a random junction list; the oracle cells run in
  lockstep and each interval gets the junction current computed here from
  the cells' compartment-0 voltages at the interval's start, with the
  reference's junction shape g (0.8 exp(-0.01 dV^2) + 0.2) dV and the
  per-pair conductance over compartment 0's area, injected as -I into
  compartment 0 (the engine adds +I to the frozen current).  Bar: the first
  interval to rounding (1e-9 mV), identical spike counts, spike times within
  one interval, subthreshold voltages within 0.1 mV (the same bar as the
  population engines' per-type tests; the engine is built with FMA
  contraction, the oracle without).
* Purkinje cells inside the mix run the per-cell mixed body unchanged: equal
  to the CB_ENGINE_CELL_MIXED population on the same inputs.
* Independence and replay: each type alone equals its part of the full mix
  bit for bit, and the graph-replayed chunks equal the same launches enqueued
  one by one, over several chunks with host work between them.
"""
import ctypes
import os

import numpy as np
import pytest

import cbrkc
import orc
from trace_metrics import compare

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="the batched mix engine is switched off (cb_mix_create refuses): the first device run "
                            "of these tests ended with a GPU memory fault whose cause has not been found"),
]

GR, GO, PKJ, IO = cbrkc.GR, cbrkc.GO, cbrkc.PKJ, cbrkc.IO


def _L():
    from test_oracle_go import L as LG
    from test_oracle_io import L as LI
    lib = orc.lib()
    LG()
    LI()
    dp = ctypes.POINTER(ctypes.c_double)
    lib.or_gr_create_comp.restype = ctypes.c_void_p
    lib.or_gr_create_comp.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int]
    return lib


def _blocks(path, rows_per_cell, n, outdir, stem):
    lines = [l for l in open(path) if l.strip()]
    assert len(lines) == n * rows_per_cell
    out = []
    for c in range(n):
        p = os.path.join(outdir, f"{stem}{c}.txt")
        with open(p, "w") as f:
            f.writelines(lines[c * rows_per_cell:(c + 1) * rows_per_cell])
        out.append(p)
    return out


def _gj_current(pairs, n, v0s, area0):
    """junction current of each cell from the compartment-0 voltages v0s"""
    cur = np.zeros(n)
    for a, b, g in pairs:
        gpa = g * 1e-6 / area0
        for i, j in ((a, b), (b, a)):
            dv = v0s[i] - v0s[j]
            cur[i] += gpa * (0.8 * np.exp(-0.01 * dv * dv) + 0.2) * dv
    return cur


def _check(ref, got, name, sub_tol=0.1, floor=None):
    """floor: the oracle against itself under a one-ulp change of its initial
    voltages (first-interval |dV|, spike-time and subthreshold spread); the
    bars widen to a small multiple of it where the model itself is that
    sensitive"""
    first = float(np.max(np.abs(ref[0] - got[0])))
    counts, dt, sub = compare(ref, got, thr=10.0)
    f_first, f_dt, f_sub = floor if floor is not None else (0.0, 0.0, 0.0)
    print(f"{name}: first interval max |dV| {first:.3e} mV; spikes {counts.sum(axis=0).tolist()}, "
          f"max spike-time diff {dt.max():.4f} ms, max subthreshold |dV| {sub.max():.3e} mV; "
          f"one-ulp floor {f_first:.3e} / {f_dt:.4f} / {f_sub:.3e}", flush=True)
    assert np.all(np.isfinite(got))
    assert first <= max(1e-9, 4.0 * f_first), name
    assert np.all(counts[:, 0] == counts[:, 1]), (name, counts.tolist())
    assert dt.max() <= max(0.125 + 1e-12, 2.0 * f_dt), name
    assert sub.max() <= max(sub_tol, 2.0 * f_sub), name


def _olive_oracle(lib, iof, v0, amp, pairs, steps):
    """the olive cells as oracle populations of one in lockstep, each interval
    with the junction current from the compartment-0 voltages at its start"""
    rows = [l.split() for l in open(iof) if l.strip()]
    d, ln = float(rows[0][2]), float(rows[0][3])
    area0 = np.pi * d * 1e-4 * ln * 1e-4
    n = len(v0)
    cells = [orc.Pop(lib.or_io_create(iof.encode(), 1, 1, orc._dp(v0[c:c + 1]), 1e-6, 1e-6, 0), 1, 3)
             for c in range(n)]
    ref = np.zeros((steps, n))
    v0s = np.array(v0, dtype=float)
    for k in range(steps):
        cur = _gj_current(pairs, n, v0s, area0)
        for c, o in enumerate(cells):
            iext = np.zeros(3)
            iext[0] = amp[c] - cur[c]
            o.set(105, iext)
            o.advance()
        vv = [o.get(orc.VNEW) for o in cells]
        v0s = np.array([v[0] for v in vv])
        ref[k] = [v[2] for v in vv]
    return ref


def _floor(ref, other):
    counts, dt, sub = compare(ref, other, thr=10.0)
    return float(np.max(np.abs(ref[0] - other[0]))), float(dt.max()), float(sub.max())


def test_small_types_against_the_oracle_cell_by_cell(ref_data, tmp_path):
    lib = _L()
    steps = 240   # 30 ms
    # granule chains
    ngr = 3
    grf = str(tmp_path / "gr_chains.txt")
    cbrkc.synth_gr_chains(grf, os.path.join(ref_data, "Gr.txt"), 7, ngr)
    gr_v0 = np.array([-70.0, -69.0, -71.0])
    gr_amp = np.array([0.009e-3, 0.011e-3, 0.013e-3])
    # Golgi trees
    ngo = 2
    gof = str(tmp_path / "go_trees.txt")
    cbrkc.synth_go_trees(gof, os.path.join(ref_data, "Go_131.txt"), 8, ngo)
    go_v0 = np.array([-60.0, -61.0])
    go_amp = np.array([0.2e-3, 0.3e-3])
    # olive cells, four, joined by a random list
    nio = 4
    gjf = str(tmp_path / "io_gj.csv")
    cbrkc.synth_io_gj(gjf, 9, nio, k=2, g_min_nS=0.5, g_max_nS=1.5)
    pairs = [(int(a), int(b), float(g)) for a, b, g in (l.strip().split(",") for l in open(gjf) if l.strip())]
    io_v0 = np.array([-60.0, -58.0, -62.0, -59.0])
    io_amp = np.array([5.5, 4.5, 6.0, 5.0])
    print("case: mix of 3 granule chains, 2 Golgi trees, 4 coupled olive cells, 30 ms, vs the oracle per cell",
          flush=True)
    m = cbrkc.Mix({GR: dict(n=ngr, morph=grf, v0=gr_v0, amp=gr_amp, comp=4),
                   GO: dict(n=ngo, morph=gof, v0=go_v0, amp=go_amp),
                   IO: dict(n=nio, morph=os.path.join(ref_data, "IO.txt"), v0=io_v0, amp=io_amp)},
                  io_gj=gjf)
    rec, spk = m.run(steps)

    # oracle: granule cells
    ref = np.zeros((steps, ngr))
    for c, f in enumerate(_blocks(grf, 4, ngr, str(tmp_path), "gr")):
        o = orc.Pop(lib.or_gr_create_comp(f.encode(), 4, 1, 1, orc._dp(gr_v0[c:c + 1]), 1e-9, 1e-9, 0), 1, 4)
        o.set_stim(1, 0.0, 1e300, np.array([gr_amp[c]]))
        for k in range(steps):
            o.set_current(k * 0.125)
            o.advance()
            ref[k, c] = o.get(orc.VNEW)[3]
    _check(ref, m.trace(rec, GR), "granule")
    # oracle: Golgi cells
    ref = np.zeros((steps, ngo))
    for c, f in enumerate(_blocks(gof, 131, ngo, str(tmp_path), "go")):
        o = orc.Pop(lib.or_go_create(f.encode(), 1, 1, orc._dp(go_v0[c:c + 1]), 1e-6, 1e-6, 0), 1, 131)
        o.set_stim(30, 0.0, 1e300, np.array([go_amp[c]]))
        for k in range(steps):
            o.set_current(k * 0.125)
            o.advance()
            ref[k, c] = o.get(orc.VNEW)[30]
    _check(ref, m.trace(rec, GO), "Golgi")
    # oracle: olive cells in lockstep with the junction current per interval;
    # the olive model moves its adaptive steps under last-bit changes, so its
    # bar is measured on the oracle itself (initial voltages one ulp lower)
    iof = os.path.join(ref_data, "IO.txt")
    ref = _olive_oracle(lib, iof, io_v0, io_amp, pairs, steps)
    nudged = _olive_oracle(lib, iof, np.nextafter(io_v0, 0.0), io_amp, pairs, steps)
    _check(ref, m.trace(rec, IO), "olive", floor=_floor(ref, nudged))
    # the device raster agrees with the recorded traces' crossings
    for t in (GR, GO, IO):
        tr = m.trace(rec, t)
        prev = np.vstack([np.full((1, tr.shape[1]), 100.0), tr[:-1]])
        np.testing.assert_array_equal(m.spikes(spk, t), (tr > 10.0) & (prev <= 10.0))
    s = m.stats(GR)
    assert s["steps"] > 0 and s["ion_updates"] > 0 and s["macro_steps"] == steps


def test_purkinje_part_equals_the_mixed_engine(tmp_path):
    n, steps = 3, 96
    morph = str(tmp_path / "pkj3.txt")
    cbrkc.synth_pkj_morph(morph, seed=31, ncell=n)
    v0 = np.array([-68.0, -67.0, -69.0])
    amp = np.array([0.4, 0.7, 1.0]) * 1e-3
    print("case: 3 Purkinje cells in the mix vs the per-cell mixed engine, 12 ms", flush=True)
    m = cbrkc.Mix({PKJ: dict(n=n, morph=morph, v0=v0, amp=amp)})
    rec, spk = m.run(steps)
    p = cbrkc.Population(PKJ, n, 1, engine=cbrkc.ENGINE_CELL_MIXED, morph=morph, v0=v0,
                         stim=(1599, 0.0, 1e300, amp), **cbrkc.STABLE)
    p.autotune(0, threads=512)
    ref = p.run(steps, 0.0, 1599)
    np.testing.assert_array_equal(m.trace(rec, PKJ), ref)
    prev = np.vstack([np.full((1, n), 100.0), ref[:-1]])
    np.testing.assert_array_equal(m.spikes(spk, PKJ), (ref > -10.0) & (prev <= -10.0))


def test_parts_independent_and_graph_replays_equal_eager(ref_data, tmp_path):
    grf = str(tmp_path / "gr.txt")
    cbrkc.synth_gr_chains(grf, os.path.join(ref_data, "Gr.txt"), 3, 40)
    gof = str(tmp_path / "go.txt")
    cbrkc.synth_go_trees(gof, os.path.join(ref_data, "Go_131.txt"), 4, 5)
    pkf = str(tmp_path / "pkj.txt")
    cbrkc.synth_pkj_morph(pkf, seed=5, ncell=2)
    gjf = str(tmp_path / "gj.csv")
    cbrkc.synth_io_gj(gjf, 6, 20, k=6)
    pops = {GR: dict(n=40, morph=grf, v0=np.linspace(-72, -68, 40), amp=np.linspace(0.008e-3, 0.016e-3, 40)),
            GO: dict(n=5, morph=gof, v0=np.linspace(-62, -58, 5), amp=np.linspace(0.1e-3, 0.3e-3, 5)),
            PKJ: dict(n=2, morph=pkf, v0=[-68.0, -67.0], amp=[0.5e-3, 0.8e-3]),
            IO: dict(n=20, morph=os.path.join(ref_data, "IO.txt"), v0=np.linspace(-62, -58, 20),
                     amp=np.linspace(3.0, 6.0, 20))}
    nint = 3 * 64
    print("case: full mix (40 GR, 5 GO, 2 PKJ, 20 coupled IO), graph replays vs eager, 24 ms", flush=True)
    a = cbrkc.Mix(pops, io_gj=gjf)
    ra, sa = [], []
    for _ in range(3):
        r, s = a.run(64)
        ra.append(r)
        sa.append(s)
        _ = np.sum(r)   # host work between the replays
    ra, sa = np.vstack(ra), np.vstack(sa)
    b = cbrkc.Mix(pops, io_gj=gjf)
    b.use_graph(False)
    rb, sb = b.run(nint)
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(sa, sb)
    assert np.all(np.isfinite(ra))
    for t in (GR, GO, PKJ, IO):
        print(f"type {t} alone", flush=True)
        alone = cbrkc.Mix({t: pops[t]}, io_gj=gjf if t == IO else None)
        r1, s1 = alone.run(nint)
        np.testing.assert_array_equal(alone.trace(r1, t), a.trace(ra, t))
        np.testing.assert_array_equal(alone.spikes(s1, t), a.spikes(sa, t))
