"""Device-side output (cb_pop_set_raster / cb_pop_spikes, CB_REC_ALL), on the
GPU through the C-ABI: the spike list a run builds on the device equals the
upward threshold crossings of the same compartment found on the host from a
recording of it (previous value 100 mV at the start, the reference's
bookkeeping, main.cu:162-169; spike time = the end of the output interval),
for the population-global RKC engine (granule cells, raster on compartment
77 while compartment 16 is recorded, as gr_output_file), the CN engine
(Golgi cells), the per-cell mixed engine (Purkinje cells) and the olive
(every compartment recorded with CB_REC_ALL, raster on compartment 2)."""
import os

import numpy as np
import pytest

import cbrkc

pytestmark = pytest.mark.gpu


def host_raster(tr, thr, t0=0.0):
    prev = np.vstack([np.full((1, tr.shape[1]), 100.0), tr[:-1]])
    k, c = np.nonzero((tr > thr) & (prev <= thr))
    order = np.lexsort((c, k))
    return t0 + (k[order] + 1) * cbrkc.H_MAX, c[order]


def check(p, twin, steps, rec_comp, raster_comp, thr, per=1):
    p.set_raster(raster_comp, thr)
    rec = p.run(steps, 0.0, rec_comp)
    t, c = p.spikes()
    tr = twin.run(steps, 0.0, raster_comp)     # the raster compartment of a twin population
    ht, hc = host_raster(tr, thr)
    print(f"{len(t)} spikes on the device, {len(ht)} from the recorded trace", flush=True)
    assert len(ht) > 0
    np.testing.assert_array_equal(c, hc)
    np.testing.assert_array_equal(t, ht)
    assert p.spikes()[0].size == 0           # the list was emptied
    return rec


def test_raster_global_rkc_granule(ref_data):
    kw = dict(morph=os.path.join(ref_data, "Gr.txt"), v0=[-70.0, -69.0], stim=(16, 1.0, 1e9, [0.012e-3, 0.016e-3]))
    print("case: GR 2 cells, global RKC, raster on 77 while recording 16, 12 ms", flush=True)
    p = cbrkc.Population(cbrkc.GR, 2, 1, **kw)
    twin = cbrkc.Population(cbrkc.GR, 2, 1, **kw)
    rec = check(p, twin, 96, 16, 77, 10.0)
    assert rec.shape == (96, 2)


def test_raster_cn_golgi(ref_data):
    kw = dict(solver=cbrkc.CN, morph=os.path.join(ref_data, "Go_131.txt"), v0=[-60.0, -55.0],
              stim=(30, 0.0, 1e9, [0.6e-3, 0.9e-3]))
    print("case: GO 2 x 1, CN, raster on 30, 25 ms", flush=True)
    p = cbrkc.Population(cbrkc.GO, 2, 1, **kw)
    twin = cbrkc.Population(cbrkc.GO, 2, 1, **kw)
    check(p, twin, 200, 30, 30, 10.0)


def test_raster_per_cell_mixed_purkinje(synth_dir):
    morph = str(synth_dir / "pkj_raster.txt")
    cbrkc.synth_pkj_morph(morph, seed=17, ncell=3)
    kw = dict(engine=cbrkc.ENGINE_CELL_MIXED, morph=morph, v0=[-68.0, -67.0, -69.0],
              stim=(1599, 0.0, 1e9, np.array([0.5, 0.8, 1.0]) * 1e-3), **cbrkc.STABLE)
    print("case: PKJ 3 cells, per-cell mixed engine, raster on 1599, 20 ms", flush=True)
    p = cbrkc.Population(cbrkc.PKJ, 3, 1, **kw)
    twin = cbrkc.Population(cbrkc.PKJ, 3, 1, **kw)
    check(p, twin, 160, 1599, 1599, -10.0)
    # recording another compartment than the raster one is refused
    p.set_raster(1599, -10.0)
    with pytest.raises(cbrkc.CbError):
        p.run(1, 20.0, 0)


def test_record_all_and_raster_olive(ref_data):
    kw = dict(morph=os.path.join(ref_data, "IO.txt"), v0=[-60.0, -62.0, -58.0, -61.0],
              stim=(0, 1.0, 1e9, [5.5, 6.0, 5.0, 6.5]))
    print("case: IO 2 x 2, global RKC, every compartment recorded, raster on 2, 30 ms", flush=True)
    p = cbrkc.Population(cbrkc.IO, 2, 2, **kw)
    twin = cbrkc.Population(cbrkc.IO, 2, 2, **kw)
    rec = check(p, twin, 240, cbrkc.REC_ALL, 2, 10.0)
    assert rec.shape == (240, 12)
    again = cbrkc.Population(cbrkc.IO, 2, 2, **kw)
    np.testing.assert_array_equal(again.run(240, 0.0, 2), rec[:, 2::3])
