"""Parity of the plain FP64 population-global RKC engine (HIP) with the CPU
oracle on the same inputs.

Both restate solve_rkc.cu with the same expression order and the same
deterministic norms.  With the ion update switched off the RKC loop uses only
correctly rounded operations (+ - * / sqrt), so the two must agree bit for
bit.  The ion update calls exp/log, whose device implementation differs
from the host's by about one ulp; with it on, the comparison is by tolerance
(and the single update itself is compared at the ulp level)."""
import os

import numpy as np
import pytest

import cbrkc
import orc
from trace_metrics import compare

pytestmark = pytest.mark.gpu


def make_pair(morph, nx, ny, v0, stim, atol, rtol, ion_interval=None, refresh_fn=False, rho_floor=False):
    o = orc.Pop.pkj(morph, nx, ny, v0=v0, atol=atol, rtol=rtol, libc_rand=False)
    o.set_stim(*stim)
    if ion_interval:
        orc.lib().or_pop_set_integ(o.h, ion_interval)
    orc.lib().or_pop_set_refresh_fn(o.h, 1 if refresh_fn else 0)
    if rho_floor:
        orc.lib().or_pop_set_rho_floor(o.h, 1.2 * max(o.cable_rho()))
    g = cbrkc.Population(cbrkc.PKJ, nx, ny, morph=morph, v0=v0, atol=atol, rtol=rtol, stim=stim,
                         engine=cbrkc.ENGINE_GLOBAL_FP64, ion_interval=ion_interval, refresh_fn=refresh_fn,
                         rho_floor=rho_floor)
    return o, g


@pytest.mark.parametrize("case", ["ref1", "synth3"])
def test_rkc_loop_bit_exact_without_ion_updates(ref_data, synth_dir, case):
    # The first macro step with the in-step ion updates disabled runs the
    # initialisation, the power method and ~100 accepted/rejected steps
    # using only correctly rounded arithmetic: results must be identical.
    # (From the second macro step on, the unconditional ion update at the
    # macro-step start, solve_rkc.cu:1274-1278, brings in exp/log.)
    if case == "ref1":
        morph, n, v0, amp = os.path.join(ref_data, "Pkj.txt"), 1, np.array([-67.3]), 2e-3
    else:
        morph, n = str(synth_dir / "pkj3b.txt"), 3
        cbrkc.synth_pkj_morph(morph, seed=5, ncell=3)
        v0, amp = np.array([-68.5, -67.0, -68.9]), [5e-4, 1e-3, 2e-3]
    print(f"case: {case}, ion updates off, first macro step", flush=True)
    o, g = make_pair(morph, n, 1, v0, (1599, 0.0, 1e9, amp), 1e-6, 1e-6, ion_interval=1e30)
    oc, ohw, _, _ = o.controller()
    gc, ghw = g.controller()
    assert np.array_equal(oc, gc), (oc, gc)
    o.set_current(0.0)
    o.advance()
    g.set_current(0.0)
    g.advance()
    for fo, fg in ((orc.VNEW, cbrkc.F_VNEW), (orc.YN, cbrkc.F_YN), (orc.FN, cbrkc.F_FN), (orc.Y, cbrkc.F_Y)):
        assert np.array_equal(o.get(fo), g.read(fg)), fo
    oc, ohw, _, _ = o.controller()
    gc, ghw = g.controller()
    assert np.array_equal(oc, gc) and np.array_equal(ohw, ghw)
    so, sg = o.stats(), g.stats()
    print("oracle", so, "gpu", sg, flush=True)
    assert so["steps"] >= 5
    for k in ("steps", "rejects", "stages", "rho_iters"):
        assert so[k] == sg[k], k


def test_one_ion_update_matches_to_ulps(ref_data):
    morph = os.path.join(ref_data, "Pkj.txt")
    print("case: Pkj.txt 1x1, first macro step with ion updates", flush=True)
    o, g = make_pair(morph, 1, 1, np.array([-67.3]), (1599, 0.0, 1e9, 2e-3), 1e-6, 1e-6)
    o.set_current(0.0)
    o.advance()
    g.set_current(0.0)
    g.advance()
    for fo, fg in [(orc.CA, cbrkc.F_CA), (orc.REV, cbrkc.F_REV), (orc.B, cbrkc.F_B)] + \
                  [(orc.ION0 + k, cbrkc.F_GATE0 + k) for k in range(18)]:
        a, b = o.get(fo), g.read(fg)
        # b sums channel currents that nearly cancel, which turns ulp noise of the
        # device exp/log into ~1e-11 relative; 1e-9 still pins every term.
        assert np.allclose(a, b, rtol=1e-9, atol=1e-300), (fo, np.max(np.abs(a - b)))


def run_compare(o, g, steps, n=1):
    worst = 0.0
    t = 0.0
    to, tg = np.zeros((steps, n)), np.zeros((steps, n))
    for k in range(steps):
        o.set_current(t)
        o.advance()
        g.set_current(t)
        g.advance()
        t += cbrkc.H_MAX
        vo, vg = o.get(orc.VNEW), g.read(cbrkc.F_VNEW)
        worst = max(worst, float(np.max(np.abs(vo - vg))))
        to[k], tg[k] = vo[1599::1600], vg[1599::1600]
    return worst, compare(to, tg)


def test_reference_morphology_one_cell(ref_data):
    morph = os.path.join(ref_data, "Pkj.txt")
    print("case: Pkj.txt 1x1, 24 macro steps with ion updates", flush=True)
    o, g = make_pair(morph, 1, 1, np.array([-67.3]), (1599, 0.5, 1e9, 2e-3), 1e-6, 1e-6)
    worst, _ = run_compare(o, g, 24)
    print("oracle", o.stats(), "gpu", g.stats(), "max|dV|", worst, flush=True)
    assert worst < 0.1


def test_synthetic_trees_three_cells(synth_dir):
    morph = str(synth_dir / "pkj3.txt")
    cbrkc.synth_pkj_morph(morph, seed=11, ncell=3)
    print("case: synthetic 3 cells, 40 macro steps", flush=True)
    v0 = np.array([-68.5, -67.0, -68.9])
    # rtol 1e-6 (TOL_PKJ); the synthetic trees need the 2^-8 ms ion interval and
    # the fn refresh to stay stable under the split scheme
    o, g = make_pair(morph, 3, 1, v0, (1599, 1.0, 1e9, [5e-4, 1e-3, 2e-3]), 1e-6, 1e-6, **cbrkc.STABLE)
    worst, (counts, dt, sub) = run_compare(o, g, 80, n=3)
    print("oracle", o.stats(), "gpu", g.stats(), "max|dV|", worst, "counts", counts.tolist(), "dt", dt, "sub", sub,
          flush=True)
    assert np.all(counts[:, 0] == counts[:, 1]) and counts.sum() > 0
    assert np.all(dt <= 0.01) and np.all(sub <= 0.1)
