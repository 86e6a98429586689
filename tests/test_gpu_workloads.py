"""The plain and the scale-stress workloads at full size, on the GPU through the C-ABI.

* Plain method, small: 9 seeded synthetic Purkinje cells of 1600
  compartments (a 3 x 3 population without gap junctions: Pkj_GJ.csv's
  pairs name the reference morphology's compartments), FP64 RKC with the
  population-global controller at rtol 1e-4 / atol 1e-6, V0 = -68 + N(0, 1)
  mV, 0.5 nA at the soma over [20, 80) ms, 100 ms (800 macro steps),
  against the oracle (the same algorithm in C, same reduction order).  The
  two differ only through exp / log rounding (device vs glibc), and at this
  loose tolerance a last-bit difference moves the adaptive step sequence;
  the oracle's own spread under a one-ulp change of one initial voltage is
  measured in the same test and the GPU must stay within it (with a floor
  of one output interval for the spike times).
* Scale stress: 8192 synthetic Purkinje cells with diameters x U(0.5, 2.0),
  50 ms, mixed per-cell engine against the per-cell FP64 engine, checked by
  size-independent properties: every voltage finite, spike counts per cell
  equal to the FP64 run's except for as many cells as the FP64 model itself
  moves under a change of its reduction order (the FP64 engine at another
  thread mapping), spike-time shifts and subthreshold differences within
  that same noise floor.

The synthetic trees run with the safeguards they need (ion interval 2^-8 ms,
fn refresh, spectral-radius floor) on both sides of every comparison.
"""
import importlib.util
import math
import os

import numpy as np
import pytest

import cbrkc
import orc
from trace_metrics import compare

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def normal_v0(seed, n):
    """-68 + N(0, 1) mV per cell from Philox4x32-10 (Box-Muller on two
    53-bit uniforms of counter (cell, 0, 0, 0), key (seed, 0x43454c4c))"""
    v0 = np.zeros(n)
    for c in range(n):
        r = cbrkc.philox4x32_10([c, 0, 0, 0], [seed & 0xffffffff, 0x43454C4C])
        u1 = ((r[0] >> 5) * 67108864.0 + (r[1] >> 6)) / 9007199254740992.0
        u2 = ((r[2] >> 5) * 67108864.0 + (r[3] >> 6)) / 9007199254740992.0
        v0[c] = -68.0 + math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)
    return v0


def oracle_run(morph, n, v0, steps, atol, rtol, stim):
    o = orc.Pop.pkj(morph, n, 1, v0=v0, atol=atol, rtol=rtol, libc_rand=False)
    o.stable_split()
    o.set_stim(*stim)
    o.threads(min(16, os.cpu_count() or 1))
    out = np.zeros((steps, n))
    for k in range(steps):
        o.set_current(k * 0.125)
        o.advance()
        out[k] = o.get(orc.VNEW)[1599::1600]
    return out


def test_plain_workload_9_cells_fp64_rkc_against_the_oracle(synth_dir):
    n, steps, seed = 9, 800, 101
    morph = str(synth_dir / "pkj_plain9.txt")
    cbrkc.synth_pkj_morph(morph, seed=seed, ncell=n)
    v0 = normal_v0(seed, n)
    stim = (1599, 20.0, 80.0, 0.5e-3)
    print("case: plain workload, 9 cells, global FP64 RKC rtol 1e-4 atol 1e-6, 0.5 nA [20, 80) ms, 100 ms", flush=True)
    p = cbrkc.Population(cbrkc.PKJ, 3, 3, engine=cbrkc.ENGINE_GLOBAL_FP64, morph=morph, v0=v0, atol=1e-6,
                         rtol=1e-4, stim=(stim[0], stim[1], stim[2], np.full(n, stim[3])), **cbrkc.STABLE)
    gpu = p.run(steps, 0.0, 1599)
    ref = oracle_run(morph, n, v0, steps, 1e-6, 1e-4, (stim[0], stim[1], stim[2], np.full(n, stim[3])))
    v1 = v0.copy()
    v1[0] = np.nextafter(v1[0], 0.0)
    floor = oracle_run(morph, n, v1, steps, 1e-6, 1e-4, (stim[0], stim[1], stim[2], np.full(n, stim[3])))
    assert np.all(np.isfinite(gpu))
    first = float(np.max(np.abs(gpu[0] - ref[0])))
    cg, dg, sg = compare(ref, gpu)
    cf, df, sf = compare(ref, floor)
    print(f"first macro step max |dV| {first:.3e} mV; GPU vs oracle: spikes {cg.sum(axis=0).tolist()}, "
          f"max spike-time diff {dg.max():.4f} ms, max subthreshold {sg.max():.4f} mV; oracle one-ulp floor: "
          f"max spike-time diff {df.max():.4f} ms, max subthreshold {sf.max():.4f} mV", flush=True)
    assert first <= 1e-9
    assert cg[:, 0].sum() > 0 and np.all(cg[:, 0] == cg[:, 1]), cg.tolist()
    assert dg.max() <= max(0.125, 2.0 * df.max())
    assert sg.max() <= max(0.1, 2.0 * sf.max())


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_scale_stress_8192_cells_properties(synth_dir):
    b = _bench()
    n, steps, seed = 8192, 400, 2024
    morph = str(synth_dir / "pkj_scale8192.txt")
    cbrkc.synth_pkj_morph(morph, seed=seed, ncell=n, dscale=b.philox_dscale(seed, n))
    v0, amp = b.philox_params(seed, n)
    print("case: scale-stress workload, 8192 cells x 1600 compartments, diameters x U(0.5, 2), 50 ms", flush=True)

    def run(engine, threads):
        p = cbrkc.Population(cbrkc.PKJ, n, 1, engine=engine, morph=morph, v0=v0, atol=1e-6, rtol=1e-6,
                             stim=(1599, 0.0, 1e9, amp), **cbrkc.STABLE)
        p.autotune(0, threads=threads)
        out = p.run(steps, 0.0, 1599)
        del p
        return out

    mixed = run(cbrkc.ENGINE_CELL_MIXED, 512)
    fp64 = run(cbrkc.ENGINE_CELL_FP64, 256)
    fp64b = run(cbrkc.ENGINE_CELL_FP64, 320)     # only the reduction order differs
    assert np.all(np.isfinite(mixed)) and np.all(np.abs(mixed) < 200.0)
    cm, dm, sm = compare(fp64, mixed)
    cf, df, sf = compare(fp64, fp64b)
    mis_m, mis_f = int(np.sum(cm[:, 0] != cm[:, 1])), int(np.sum(cf[:, 0] != cf[:, 1]))
    print(f"mixed vs FP64: {mis_m} cells with another spike count, spikes {cm.sum(axis=0).tolist()}, "
          f"median spike-time shift {np.median(dm):.4f} ms, median subthreshold {np.median(sm):.4f} mV; "
          f"FP64 reorder floor: {mis_f} cells, median shift {np.median(df):.4f} ms, "
          f"median subthreshold {np.median(sf):.4f} mV", flush=True)
    assert cm[:, 0].sum() > 1000
    assert mis_m <= max(4, 3 * mis_f)
    assert np.median(dm) <= max(0.005, 3.0 * np.median(df))
    assert np.median(sm) <= max(0.05, 3.0 * np.median(sf))


def _oracle_cells(files, v0, amp, steps, seed, gid0):
    """each cell as its own oracle population (the reference algorithm on a
    population of one), all cells in parallel host threads: ctypes releases
    the GIL inside the oracle's calls"""
    from concurrent.futures import ThreadPoolExecutor

    def one(c):
        o = orc.Pop.pkj(files[c], 1, 1, v0=v0[c:c + 1], atol=1e-6, rtol=1e-6, libc_rand=False)
        o.gid_offset(gid0[c] * 1600)
        o.stable_split()
        o.enable_poisson(rate_hz=5.0, g_mS=1e-6, tau_ms=2.0, e_mV=0.0, seed=seed)
        o.set_stim(1599, 0.0, 1e9, amp[c])
        o.threads(1)
        out = np.zeros(steps)
        for k in range(steps):
            o.set_current(k * 0.125)
            o.advance()
            out[k] = o.get(orc.VNEW)[1599]
        return out

    with ThreadPoolExecutor(max_workers=min(len(files), 16)) as ex:
        return np.stack(list(ex.map(one, range(len(files)))), axis=1)
