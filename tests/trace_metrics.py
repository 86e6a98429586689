"""Comparison metrics for voltage traces sampled once per macro step.

A spiking trace cannot be compared pointwise: a spike that moves by a few
microseconds changes the sampled upstroke by tens of mV.  Parity is therefore
stated as (1) identical spike counts, (2) spike-time differences, with the
threshold crossing located by linear interpolation between samples, and (3)
the voltage difference away from spikes (both traces below -40 mV and more
than 1 ms from any spike of either trace).
"""
import numpy as np

H_MAX = 0.125


def crossings(tr, thr=-10.0):
    """Upward threshold crossings of a [steps, cells] trace; sample k is at
    t = (k + 1) H_MAX and the sample before the first counts as 100 mV, as the
    reference's raster bookkeeping does (main.cu:162-169, pkj.cu:403-409)."""
    out = []
    for c in range(tr.shape[1]):
        v = tr[:, c]
        prev = np.concatenate([[100.0], v[:-1]])
        idx = np.where((v > thr) & (prev <= thr))[0]
        ts = []
        for k in idx:
            if k == 0:
                ts.append(H_MAX)
                continue
            f = (thr - v[k - 1]) / (v[k] - v[k - 1])
            ts.append(k * H_MAX + f * H_MAX)
        out.append(np.array(ts))
    return out


def compare(a, b, thr=-10.0):
    """Returns per-cell (count_a, count_b), max spike-time difference [ms] and
    max subthreshold |dV| [mV] between traces a and b ([steps, cells])."""
    sa, sb = crossings(a, thr), crossings(b, thr)
    counts = np.array([(len(x), len(y)) for x, y in zip(sa, sb)])
    dt = np.array([np.max(np.abs(x - y)) if len(x) == len(y) and len(x) else 0.0 for x, y in zip(sa, sb)])
    t = (np.arange(a.shape[0]) + 1) * H_MAX
    sub = np.zeros(a.shape[1])
    for c in range(a.shape[1]):
        m = (a[:, c] < -40.0) & (b[:, c] < -40.0)
        for s in np.concatenate([sa[c], sb[c]]):
            m &= np.abs(t - s) > 1.0
        if m.any():
            sub[c] = np.max(np.abs(a[m, c] - b[m, c]))
    return counts, dt, sub
