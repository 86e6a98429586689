"""Per-cell persistent RKC engines (HIP).

* The per-cell FP64 engine is the reference algorithm applied to each cell
  as its own population; it is compared with the oracle run cell by cell.
* The mixed engine (FP32 stages/gates/currents, compensated V, FP64
  control) is compared with the per-cell FP64 engine on the same inputs.

Tolerance 1e-6 (the reference's TOL_PKJ): at the looser 1e-4 two FP64
implementations of the same algorithm already differ by ~0.4 mV between
spikes and ~0.1 ms in spike time, because a 1-ulp difference moves the
adaptive step sequence.

The synthetic trees (10 um soma on a 3 um trunk) are outside the stability
range of the reference's split scheme (frozen current for 2^log10(atol) ms,
stale fn after each ion update); they run with the 2^-8 ms ion interval and
the fn refresh, on both sides of every comparison."""
import numpy as np
import pytest

import cbrkc
import orc
from trace_metrics import compare

pytestmark = pytest.mark.gpu

STABLE = cbrkc.STABLE


def test_cell_fp64_matches_oracle_cell_by_cell(synth_dir):
    n, steps = 3, 160
    morph = str(synth_dir / "pkjc3.txt")
    cbrkc.synth_pkj_morph(morph, seed=21, ncell=n)
    v0 = np.array([-68.2, -67.4, -69.0])
    amp = np.array([6e-4, 1.2e-3, 2.0e-3])
    print(f"case: per-cell FP64 vs oracle, {n} synthetic cells, {steps} macro steps", flush=True)
    g = cbrkc.Population(cbrkc.PKJ, n, 1, morph=morph, v0=v0, atol=1e-6, rtol=1e-6,
                         stim=(1599, 1.0, 1e9, amp), engine=cbrkc.ENGINE_CELL_FP64, poisson=dict(seed=3), **STABLE)
    gv = g.run(steps, 0.0, 1599)
    files = orc.split_morph(morph, n, 1600, str(synth_dir))
    ov = np.zeros_like(gv)
    for c in range(n):
        o = orc.Pop.pkj(files[c], 1, 1, v0=v0[c:c + 1], atol=1e-6, rtol=1e-6, libc_rand=False)
        o.gid_offset(c * 1600)
        o.stable_split()
        o.enable_poisson(seed=3)
        o.set_stim(1599, 1.0, 1e9, amp[c])
        for k in range(steps):
            o.set_current(k * 0.125)
            o.advance()
            ov[k, c] = o.get(orc.VNEW)[1599]
    counts, dt, sub = compare(ov, gv)
    print("spike counts", counts.tolist(), "max spike-time diff [ms]", dt, "subthreshold |dV| [mV]", sub, flush=True)
    assert np.all(counts[:, 0] == counts[:, 1]) and counts.sum() > 0
    assert np.all(dt <= 0.01)
    assert np.all(sub <= 0.01)


def test_mixed_matches_fp64_per_cell(synth_dir):
    n, steps = 8, 400
    morph = str(synth_dir / "pkjc8.txt")
    cbrkc.synth_pkj_morph(morph, seed=33, ncell=n)
    rng = np.random.default_rng(3)
    v0 = -68.0 + rng.standard_normal(n)
    amp = rng.uniform(0.2e-3, 1.0e-3, n)
    print(f"case: mixed vs FP64 per-cell, {n} cells, {steps} macro steps, Poisson drive", flush=True)
    res = {}
    for eng in (cbrkc.ENGINE_CELL_FP64, cbrkc.ENGINE_CELL_MIXED):
        p = cbrkc.Population(cbrkc.PKJ, n, 1, morph=morph, v0=v0, atol=1e-6, rtol=1e-6,
                             stim=(1599, 2.0, 1e9, amp), engine=eng, poisson=dict(seed=7), **STABLE)
        res[eng] = (p.run(steps, 0.0, 1599), p.read(cbrkc.F_VNEW), p.stats())
        print(eng, res[eng][2], flush=True)
    a, b = res[cbrkc.ENGINE_CELL_FP64], res[cbrkc.ENGINE_CELL_MIXED]
    counts, dt, sub = compare(a[0], b[0])
    print("spike counts", counts.tolist(), "max spike-time diff [ms]", dt, "subthreshold |dV| [mV]", sub, flush=True)
    assert np.all(counts[:, 0] == counts[:, 1]) and counts.sum() > 0
    assert np.all(dt <= 0.01)
    assert np.all(sub <= 0.1)


def test_single_macro_steps_equal_one_long_launch(synth_dir):
    # advance() one macro step at a time and run(k) in one persistent launch
    # must give the same trajectory (the launch boundary saves and restores
    # the whole per-cell state)
    n = 4
    morph = str(synth_dir / "pkjc4.txt")
    cbrkc.synth_pkj_morph(morph, seed=4, ncell=n)
    kw = dict(morph=morph, v0=np.full(n, -67.0), atol=1e-6, rtol=1e-6, stim=(1599, 0.5, 1e9, 1e-3),
              engine=cbrkc.ENGINE_CELL_MIXED, **STABLE)
    p1 = cbrkc.Population(cbrkc.PKJ, n, 1, **kw)
    p2 = cbrkc.Population(cbrkc.PKJ, n, 1, **kw)
    tr = p1.run(24, 0.0, 1599)
    rows = []
    for k in range(24):
        p2.set_current(k * 0.125)
        p2.advance()
        rows.append(p2.read(cbrkc.F_VNEW)[1599::1600])
    assert np.array_equal(np.array(rows), tr)
    assert np.array_equal(p1.read(cbrkc.F_VNEW), p2.read(cbrkc.F_VNEW))
