"""Pins of the fixed in-degree network (oracle/or_net.c) and the product's
generator against it (host only, no GPU).

* Generator: the product's list (cb_net_generate) equals the oracle's
  element for element, at a small size and at the full network size (2048
  granule, 512 Golgi, 512 Purkinje cells: exactly 780404 synapses); every
  postsynaptic cell has its kind's in-degree (one more granule input for the
  first total - sum Purkinje cells); presynaptic cells, target compartments
  and weights stay in their ranges and are spread as uniform draws should be
  (chi-square over cells and compartments, mean and spread of the weights);
  mossy fibres are numbered one per synapse; another seed changes the draws,
  the same seed repeats them.
* Delivery, through closed forms on the oracle: a single spike of one
  presynaptic cell gives its targets' slots exactly w exp(-dt / tau)^m, m
  intervals after its arrival, delay + 1 intervals after the spike; a cell
  that fires every interval drives a slot to w / (1 - exp(-dt / tau)); the
  mossy-fibre draws fire at rate x dt to binomial accuracy; a slot adds its
  synapses in generation order (the sum of two weights is the floating-point
  sum in that order).
"""
import ctypes

import numpy as np
import pytest

import cbrkc
import net_oracle as NO
import orc

N_FULL = (2048, 512, 512)
COMP = (4, 131, 1600)


def product_list(cfg, n, comp):
    return cbrkc.net_generate(cfg, n, comp)


def test_product_generator_equals_the_oracle_small():
    cfg = cbrkc.net_config(total=0, deg_mf_gr=3, deg_go_gr=2, deg_gr_go=7, deg_gr_pkj=11, seed=99)
    n = (37, 5, 3)
    a = product_list(cfg, n, COMP)
    b = NO.oracle_generate(NO.spec_of(cfg, n, COMP))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_full_size_network_list():
    cfg = cbrkc.net_config()
    kind, pre, post, w = product_list(cfg, N_FULL, COMP)
    assert len(kind) == 780404
    ok, op, oq, ow = NO.oracle_generate(NO.spec_of(cfg, N_FULL, COMP))
    assert np.array_equal(kind, ok) and np.array_equal(pre, op) and np.array_equal(post, oq)
    assert np.array_equal(w, ow)
    n_gr, n_go, n_pkj = N_FULL
    deg = [cfg.deg_mf_gr, cfg.deg_go_gr, cfg.deg_gr_go, cfg.deg_gr_pkj]
    extra = 780404 - sum(d * n for d, n in zip(deg, (n_gr, n_gr, n_go, n_pkj)))
    assert extra == 116
    post_pop = {0: 0, 1: 0, 2: 1, 3: 2}
    pre_n = {1: n_go, 2: n_gr, 3: n_gr}
    lo = {0: (cfg.gr_target_comp, cfg.gr_target_comp + 1), 1: (cfg.go_target_lo, cfg.go_target_hi),
          2: (cfg.pkj_target_lo, cfg.pkj_target_hi)}
    for k in range(4):
        m = kind == k
        p = post_pop[k]
        cells = post[m] // COMP[p]
        local = post[m] % COMP[p]
        counts = np.bincount(cells, minlength=N_FULL[p])
        want = np.full(N_FULL[p], deg[k])
        if k == 3:
            want[:extra] += 1
        assert np.array_equal(counts, want), k
        # generation order: cells ascending
        assert np.all(np.diff(cells) >= 0)
        assert np.all((local >= lo[p][0]) & (local < lo[p][1]))
        if k == 0:   # one mossy fibre per synapse, numbered in order
            assert np.array_equal(pre[m], np.arange(m.sum()))
        else:
            assert pre[m].min() >= 0 and pre[m].max() < pre_n[k]
            h = np.bincount(pre[m], minlength=pre_n[k]).astype(float)
            e = m.sum() / pre_n[k]
            chi2 = ((h - e) ** 2 / e).sum()
            dof = pre_n[k] - 1
            assert abs(chi2 - dof) < 6 * np.sqrt(2 * dof), (k, chi2, dof)
        if lo[p][1] - lo[p][0] > 1:
            span = lo[p][1] - lo[p][0]
            h = np.bincount(local - lo[p][0], minlength=span).astype(float)
            e = m.sum() / span
            chi2 = ((h - e) ** 2 / e).sum()
            assert abs(chi2 - (span - 1)) < 6 * np.sqrt(2 * (span - 1)), (k, chi2)
    assert w.min() >= cfg.w_min_nS and w.max() < cfg.w_max_nS
    assert abs(w.mean() - 1.0) < 5 * np.sqrt(1 / 12 / len(w))
    assert abs(w.std() - np.sqrt(1 / 12)) < 1e-3


def test_seed_changes_the_draws_and_repeats_them():
    n = (64, 8, 4)
    a = product_list(cbrkc.net_config(seed=1, total=0), n, COMP)
    b = product_list(cbrkc.net_config(seed=1, total=0), n, COMP)
    c = product_list(cbrkc.net_config(seed=2, total=0), n, COMP)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[1][a[0] != 0], c[1][c[0] != 0])
    assert not np.array_equal(a[3], c[3])


@pytest.mark.parametrize("bad", [dict(total=10), dict(delay_ms=0.1), dict(w_min_nS=2.0), dict(go_target_hi=200),
                                 dict(gr_spike_comp=4)])
def test_generator_rejects_bad_specs(bad):
    cfg = cbrkc.net_config(**bad)
    with pytest.raises(cbrkc.CbError):
        product_list(cfg, (16, 2, 2), COMP)


def test_total_beyond_one_extra_input_per_purkinje_cell_is_rejected():
    cfg = cbrkc.net_config(total=4 * 16 + 4 * 16 + 100 * 2 + 1392 * 2 + 3)
    with pytest.raises(cbrkc.CbError):
        product_list(cfg, (16, 2, 2), COMP)
    cfg = cbrkc.net_config(total=4 * 16 + 4 * 16 + 100 * 2 + 1392 * 2 + 2)
    assert len(product_list(cfg, (16, 2, 2), COMP)[0]) == cfg.total


@pytest.fixture(scope="module")
def small(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("net")
    files = NO.synth_files(tmp, 12, 2, 1)
    return files


def _net(files, cfg, v0=None):
    lib = NO.L()
    v0 = v0 or (np.full(12, -70.0), np.full(2, -60.0), np.full(1, -68.0))
    pops = NO.oracle_pops(files, v0)
    spec = NO.spec_of(cfg)
    h = lib.or_net_create(pops[0].h, pops[1].h, pops[2].h, ctypes.byref(spec))
    assert h
    return lib, pops, h


def _g(lib, h, pop, rec, nc):
    out = np.zeros(nc)
    lib.or_net_conductance(h, pop, rec, orc._dp(out))
    return out


def test_single_spike_decays_exponentially(small):
    cfg = cbrkc.net_config(total=0, deg_mf_gr=0, deg_go_gr=0, deg_gr_go=5, deg_gr_pkj=9, mf_rate_hz=0.0, seed=5)
    lib, pops, h = _net(small, cfg)
    kind, pre, post, w = NO.oracle_generate(NO.spec_of(cfg, (12, 2, 1), COMP))
    stride = 12 + 2 + 1
    delay = 8
    d_ampa = np.exp(-0.125 / cfg.tau_ampa_ms)
    spike = np.zeros(stride, np.uint8)
    spike[3] = 1                     # granule cell 3 fires in interval 0
    none = np.zeros(stride, np.uint8)
    P = ctypes.POINTER(ctypes.c_ubyte)
    hist = []
    for k in range(delay + 1 + 6):
        # interval k delivers the spikes of interval k - 1 - delay
        arr = spike if k - 1 - delay == 0 else none
        lib.or_net_deliver_given(h, arr.ctypes.data_as(P))
        hist.append(_g(lib, h, 1, 0, pops[1].nc))
    for k in range(delay + 1):
        assert not hist[k].any()
    m = (kind == 2) & (pre == 3)
    want = np.zeros(pops[1].nc)
    np.add.at(want, post[m], w[m] * 1e-6)
    assert want.any()
    for j in range(6):
        got = hist[delay + 1 + j]
        np.testing.assert_allclose(got, want * d_ampa ** j, rtol=1e-14, atol=0)
    lib.or_net_free(h)


def test_regular_firing_reaches_the_geometric_steady_state(small):
    cfg = cbrkc.net_config(total=0, deg_mf_gr=0, deg_go_gr=3, deg_gr_go=0, deg_gr_pkj=0, mf_rate_hz=0.0,
                           delay_ms=0.125, seed=8)
    lib, pops, h = _net(small, cfg)
    kind, pre, post, w = NO.oracle_generate(NO.spec_of(cfg, (12, 2, 1), COMP))
    stride = 15
    both = np.zeros(stride, np.uint8)
    both[12] = both[13] = 1          # both Golgi cells fire every interval
    P = ctypes.POINTER(ctypes.c_ubyte)
    for _ in range(2000):   # d^2000 = exp(-41.7): the series has converged
        lib.or_net_deliver_given(h, both.ctypes.data_as(P))
    d = np.exp(-0.125 / cfg.tau_gaba_ms)
    want = np.zeros(pops[0].nc)
    np.add.at(want, post[kind == 1], w[kind == 1] * 1e-6)
    np.testing.assert_allclose(_g(lib, h, 0, 1, pops[0].nc), want / (1.0 - d), rtol=1e-12)
    assert not _g(lib, h, 0, 0, pops[0].nc).any()
    lib.or_net_free(h)


def test_mossy_fibres_fire_at_their_rate(small):
    rate = 400.0
    cfg = cbrkc.net_config(total=0, deg_mf_gr=4, deg_go_gr=0, deg_gr_go=0, deg_gr_pkj=0, mf_rate_hz=rate,
                           w_min_nS=1.0, w_max_nS=1.0, seed=3)
    lib, pops, h = _net(small, cfg)
    d = np.exp(-0.125 / cfg.tau_ampa_ms)
    prev = np.zeros(pops[0].nc)
    hits = 0
    K = 2000
    for k in range(K):
        lib.or_net_deliver_given(h, None)
        g = _g(lib, h, 0, 0, pops[0].nc)
        hits += np.rint((g - prev * d) / 1e-6).sum()
        prev = g
    n = 4 * 12 * K
    p = rate * 0.125e-3
    assert abs(hits / n - p) < 5 * np.sqrt(p * (1 - p) / n), (hits / n, p)
    lib.or_net_free(h)


def test_a_slot_adds_in_generation_order(small):
    # granule cell 0's dendrite takes its 4 mossy fibres in one AMPA slot
    cfg = cbrkc.net_config(total=0, deg_mf_gr=4, deg_go_gr=0, deg_gr_go=0, deg_gr_pkj=0, mf_rate_hz=4000.0,
                           seed=21)
    lib, pops, h = _net(small, cfg)
    kind, pre, post, w = NO.oracle_generate(NO.spec_of(cfg, (12, 2, 1), COMP))
    lib.or_net_deliver_given(h, None)
    g = _g(lib, h, 0, 0, pops[0].nc)
    key = (cfg.seed & 0xffffffff) ^ 0x6d666472, cfg.seed >> 32
    for c in range(12):
        acc = 0.0
        for j in np.nonzero(post == c * 4)[0]:
            r = orc.philox([int(pre[j]), 0, 0, 0], key)
            if 4000.0 * 0.125e-3 > (r[0] + 1.0) / 4294967296.0:
                acc += w[j] * 1e-6
        assert g[c * 4] == acc   # G starts at 0: the first interval's sum itself
    lib.or_net_free(h)
