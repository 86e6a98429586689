"""Pins of the oracle's synapses (oracle/or_syn.c, restating syn.cu and the
synaptic terms of solve_rkc.cu / solve_cnm.cu / solve_bem.cu).

* Connectivity: every connection rule is checked as a set against the grid
  geometry (every pair within range is connected, no other), the contact
  compartments against their geometric meaning (the granule axon compartment
  at the Golgi / Purkinje height, the Purkinje compartment nearest in x), and
  each weight normalisation against its invariant: the weights onto one
  postsynaptic cell sum to a constant (Golgi: 64, granule: 160000, Purkinje:
  25 x 300000 x 10 / 512), whatever the cell's in-degree.
* Updates: without spikes every trace decays by exp(-0.125 / tau) per macro
  step, for each of the reference's time constants; a presynaptic upward
  crossing of 0 mV adds exactly one to every trace, a crossing needs the
  previous value below 0 and the present one above; the mossy-fibre draws
  fire at 5 Hz (60 Hz inside [900, 1000) ms) to binomial accuracy; the
  NMDA trace of the mossy-fibre synapse is fed from the updated AMPA trace.
* Coupling, through closed forms of a passive cable with a synaptic
  conductance g to E on some compartments: the BE step solves
  (C/DT + L + G) v' = C/DT v + G E + I and the CN step (from the second step
  on) (C/DT + L/2 + G/2) v' = (C/DT - L/2 - G/2) v + G E + I; on the RKC path
  the Golgi cell's synaptic current of a macro step acts exactly as an
  injected current -s fixed at the macro step's start, and the granule cell's
  synaptic term equals val (V - E) added to the frozen current.
"""
import ctypes
import os

import numpy as np
import pytest

import orc
from test_oracle_cn import generic, tree_laplacian, L as LCN

GR_COMP, GO_COMP, PKJ_COMP = 578, 131, 1600
TRACE0, TRACE1, TRACE2, WEIGHT, VAL, OLD_V = range(6)


def L():
    lib = LCN()
    if not getattr(lib, "_syn_bound", False):
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        for name, args in (("or_mfgr_create", [ctypes.c_int, ctypes.c_uint64]),
                           ("or_grgo_create", [vp, vp]), ("or_gogr_create", [vp, vp]),
                           ("or_grpkj_create", [vp, vp, dp, ctypes.c_uint64]),
                           ("or_mlipkj_create", [ctypes.c_int, ctypes.c_uint64])):
            getattr(lib, name).restype = vp
            getattr(lib, name).argtypes = args
        lib.or_syn_free.argtypes = [vp]
        lib.or_syn_update.argtypes = [vp, ctypes.c_double]
        lib.or_syn_n.argtypes = [vp]
        lib.or_syn_get.argtypes = [vp, ctypes.c_int, dp]
        lib.or_syn_set.argtypes = [vp, ctypes.c_int, dp]
        lib.or_syn_comp.argtypes = [vp, ip, ip]
        lib.or_pop_attach_syn.argtypes = [vp, vp]
        lib.or_be_init.argtypes = [vp]
        lib.or_be_step.argtypes = [vp]
        lib.or_cn_step.argtypes = [vp]
        lib._syn_bound = True
    return lib


class Syn:
    def __init__(self, h):
        assert h
        self.h = h
        self.n = L().or_syn_n(h)

    def __del__(self):
        if getattr(self, "h", None):
            L().or_syn_free(self.h)

    def get(self, f):
        out = np.zeros(self.n)
        L().or_syn_get(self.h, f, orc._dp(out))
        return out

    def set(self, f, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        L().or_syn_set(self.h, f, orc._dp(x))

    def comp(self):
        pre, post = np.zeros(self.n, dtype=np.int32), np.zeros(self.n, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        L().or_syn_comp(self.h, pre.ctypes.data_as(ip), post.ctypes.data_as(ip))
        return pre, post

    def update(self, t):
        L().or_syn_update(self.h, t)


def pkj_x(ref_data):
    a = np.loadtxt(os.path.join(ref_data, "Pkj_location_info.csv"), delimiter=",")
    assert a.shape == (1600, 5) and np.all(a[:, 0] == np.arange(1600))
    return np.ascontiguousarray(a[:, 1])


def gr_grid(ref_data, nx, ny):
    from test_oracle_gr import gr_pop
    return gr_pop(ref_data, nx, ny, v0=np.full(nx * ny, -70.0))


def go_grid(ref_data, nx, ny):
    from test_oracle_go import go_pop
    return go_pop(ref_data, nx, ny, v0=np.full(nx * ny, -60.0))


def pos_gr(i, nx):
    return (i % nx) * 16.0, (i // nx) * 16.0


def pos_32(i, nx):
    return (i % nx) * 32.0 + 8.0, (i // nx) * 32.0 + 8.0


def test_mossy_fibre_inputs():
    s = Syn(L().or_mfgr_create(3, 7))
    pre, post = s.comp()
    assert s.n == 12
    np.testing.assert_array_equal(post, np.concatenate([np.arange(4) + c * GR_COMP for c in range(3)]))
    np.testing.assert_array_equal(s.get(WEIGHT), 10000.0)


@pytest.mark.parametrize("gr_shape,go_shape", [((4, 30), (2, 3)), ((6, 20), (3, 2))])
def test_granule_to_golgi_connectivity(ref_data, gr_shape, go_shape):
    gr = gr_grid(ref_data, *gr_shape)
    go = go_grid(ref_data, *go_shape)
    s = Syn(L().or_grgo_create(gr.h, go.h))
    pre, post = s.comp()
    pairs = [(p // GR_COMP, q // GO_COMP) for p, q in zip(pre, post)]
    want = set()
    for i in range(gr.n):
        for j in range(go.n):
            (xg, yg), (xo, yo) = pos_gr(i, gr_shape[0]), pos_32(j, go_shape[0])
            if abs(xo - xg) < 16 and abs(yo - yg) < 220:
                want.add((i, j))
    assert set(pairs) == want and len(pairs) == len(want) > 0
    assert pairs == sorted(pairs)                      # granule-major, Golgi order
    np.testing.assert_array_equal(post % GO_COMP, 0)   # onto compartment 0
    for (i, j), p in zip(pairs, pre):
        # the axonal compartment 327 -/+ one per 10 um of height difference
        dy = (8.0 + (j // go_shape[0]) * 32.0) - (i // gr_shape[0]) * 16.0
        assert p % GR_COMP == 327 - int(dy / 10.0)
    w = s.get(WEIGHT)
    for j in range(go.n):
        on = [k for k, (_, jj) in enumerate(pairs) if jj == j]
        if on:
            assert abs(sum(w[on]) - 64.0) < 1e-12
    np.testing.assert_array_equal(s.get(OLD_V), 1000.0)


@pytest.mark.parametrize("go_shape,gr_shape", [((2, 2), (4, 4)), ((3, 2), (7, 5))])
def test_golgi_to_granule_connectivity(ref_data, go_shape, gr_shape):
    gr = gr_grid(ref_data, *gr_shape)
    go = go_grid(ref_data, *go_shape)
    s = Syn(L().or_gogr_create(go.h, gr.h))
    pre, post = s.comp()
    pairs = [(p // GO_COMP, q // GR_COMP) for p, q in zip(pre, post)]
    want = {(j, i) for j in range(go.n) for i in range(gr.n)
            if abs(pos_32(j, go_shape[0])[0] - pos_gr(i, gr_shape[0])[0]) < 32
            and abs(pos_32(j, go_shape[0])[1] - pos_gr(i, gr_shape[0])[1]) < 32}
    assert set(pairs) == want and len(pairs) == len(want) > 0
    assert pairs == sorted(pairs)
    np.testing.assert_array_equal(pre % GO_COMP, 30)
    np.testing.assert_array_equal(post % GR_COMP, 16)
    w = s.get(WEIGHT)
    for i in range(gr.n):
        on = [k for k, (_, ii) in enumerate(pairs) if ii == i]
        if on:
            assert abs(sum(w[on]) - 160000.0) < 1e-9


def test_granule_to_purkinje_connectivity(ref_data):
    gr = gr_grid(ref_data, 8, 4)
    pk = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 3, 1, v0=[-65.0] * 3, libc_rand=False)
    x = pkj_x(ref_data)
    s = Syn(L().or_grpkj_create(gr.h, pk.h, orc._dp(x), 11))
    pre, post = s.comp()
    cells = [(p // GR_COMP, q // PKJ_COMP) for p, q in zip(pre, post)]
    pairs = cells[::25]
    want = [(i, j) for i in range(gr.n) for j in range(pk.n)
            if abs(pos_32(j, 3)[0] - pos_gr(i, 8)[0]) < 120 and abs(pos_32(j, 3)[1] - pos_gr(i, 8)[1]) < 220]
    assert pairs == want and s.n == 25 * len(want)
    assert all(cells[k] == pairs[k // 25] for k in range(s.n))
    w = s.get(WEIGHT)
    for j in range(pk.n):
        on = [k for k, (_, jj) in enumerate(cells) if jj == j]
        assert abs(sum(w[on]) - 25 * 300000.0 * 10.0 / 512.0) < 1e-6
    # each contact lies within the jitter (5 um) of the compartment nearest the
    # granule cell's x offset; the running minimum never gets worse in a pair
    for k0 in range(0, s.n, 25):
        i, j = cells[k0]
        dx = (i % 8) * 16.0 - (8.0 + (j % 3) * 32.0)
        best = np.min(np.abs(x - dx))
        d = np.abs(x[post[k0:k0 + 25] % PKJ_COMP] - dx)
        assert np.all(d <= best + 10.0 + 1e-9)
        dy = (i // 8) * 16.0 - 8.0
        assert np.all(pre[k0:k0 + 25] % GR_COMP == 327 + int(dy / 10.0))


def test_inhibitory_purkinje_input_is_inert():
    s = Syn(L().or_mlipkj_create(4, 3))
    pre, post = s.comp()
    np.testing.assert_array_equal(post, 1599 + PKJ_COMP * np.arange(4))
    np.testing.assert_array_equal(s.get(WEIGHT), 0.0)


def test_trace_decay_and_spike_detection(ref_data):
    gr = gr_grid(ref_data, 2, 16)
    go = go_grid(ref_data, 1, 1)
    cases = [(Syn(L().or_grgo_create(gr.h, go.h)), gr, (1.5, 31.0, 170.0)),
             (Syn(L().or_gogr_create(go.h, gr.h)), go, (7.0, 59.0))]
    for s, src, taus in cases:
        n = s.n
        assert n > 0
        for f, _ in enumerate(taus):
            s.set(f, np.ones(n))
        s.set(OLD_V, np.full(n, -50.0))
        v = src.get(orc.V)
        v[:] = -60.0
        src.set(orc.V, v)
        for k in range(1, 4):
            s.update(0.0)
            for f, tau in enumerate(taus):
                np.testing.assert_allclose(s.get(f), np.exp(-0.125 / tau) ** k, rtol=1e-15)
        # an upward crossing of 0 mV adds exactly one
        pre, _ = s.comp()
        before = [s.get(f) for f in range(len(taus))]
        v[pre] = 5.0
        src.set(orc.V, v)
        s.update(0.0)
        for f, tau in enumerate(taus):
            np.testing.assert_allclose(s.get(f), before[f] * np.exp(-0.125 / tau) + 1.0, rtol=1e-15)
        np.testing.assert_array_equal(s.get(OLD_V), 5.0)
        # staying above 0 mV is no new spike
        s.update(0.0)
        for f, tau in enumerate(taus):
            np.testing.assert_allclose(s.get(f), (before[f] * np.exp(-0.125 / tau) + 1.0) * np.exp(-0.125 / tau),
                                       rtol=1e-15)
        assert np.all(s.get(VAL) > 0.0)


def test_mossy_fibre_rate_and_nmda_feed():
    n = 20000
    s = Syn(L().or_mfgr_create(n // 4, 12345))
    fired = np.zeros(2)
    for k, t in enumerate((100.0, 950.0)):
        for _ in range(20):
            a0 = s.get(TRACE0)
            s.update(t)
            a1 = s.get(TRACE0)
            spike = np.round(a1 - a0 * np.exp(-0.125 / 1.2)).astype(int)
            assert set(np.unique(spike)) <= {0, 1}
            fired[k] += spike.sum()
            # NMDA trace = updated AMPA trace x exp(-0.125 / 52) + spike
            np.testing.assert_allclose(s.get(TRACE1), a1 * np.exp(-0.125 / 52.0) + spike, rtol=1e-15, atol=1e-300)
    for k, rate in enumerate((5.0, 60.0)):
        p = rate * 0.125e-3
        mean, sd = 20 * n * p, np.sqrt(20 * n * p * (1 - p))
        assert abs(fired[k] - mean) < 5 * sd, (k, fired[k], mean)


def passive(n, rng, ncomp_syn=4):
    Lr = tree_laplacian(n, rng)
    cm = rng.uniform(0.5, 2.0, n) * 1e-6
    v0 = rng.uniform(-70.0, -60.0, n)
    p = generic(Lr, cm, v0)
    iext = np.zeros(n)
    iext[5] = 2e-5
    p.set(orc.IEXT, iext)
    s = Syn(L().or_mfgr_create(1, 0))       # four synapses onto compartments 0..3, E = 0
    g = rng.uniform(1e-6, 5e-6, 4)
    s.set(VAL, g)
    assert L().or_pop_attach_syn(p.h, s.h) == 0
    G = np.zeros(n)
    G[:4] = g
    return p, s, Lr, cm, v0, iext, G


def test_be_step_with_a_synaptic_conductance():
    rng = np.random.default_rng(21)
    p, s, Lr, cm, v0, iext, G = passive(10, rng)
    assert L().or_be_init(p.h) == 0
    dt = 0.025
    A = np.diag(cm / dt + G) - Lr
    L().or_be_step(p.h)
    np.testing.assert_allclose(p.get(orc.V), np.linalg.solve(A, cm / dt * v0 + G * 0.0 + iext), rtol=1e-12)
    assert np.max(np.abs(p.get(orc.V) - np.linalg.solve(np.diag(cm / dt) - Lr, cm / dt * v0 + iext))) > 1e-6


def test_cn_step_with_a_synaptic_conductance():
    rng = np.random.default_rng(22)
    p, s, Lr, cm, v0, iext, G = passive(10, rng)
    assert LCN().or_cn_init(p.h) == 0
    dt = 0.025
    L().or_cn_step(p.h)                              # first step: the initial half sums lack the synapse
    v1 = p.get(orc.V)
    L().or_cn_step(p.h)
    Lp = -Lr
    A = np.diag(cm / dt + G / 2) + Lp / 2
    rhs = (np.diag(cm / dt - G / 2) - Lp / 2) @ v1 + G * 0.0 + iext
    np.testing.assert_allclose(p.get(orc.V), np.linalg.solve(A, rhs), rtol=1e-12)


def test_rkc_golgi_synaptic_current_is_a_fixed_injection(ref_data):
    """One macro step with grgo input equals one with the synaptic current of
    the macro step's start injected instead (b includes -i_ext)."""
    gr = gr_grid(ref_data, 2, 16)
    a, b = go_grid(ref_data, 1, 1), go_grid(ref_data, 1, 1)
    s = Syn(L().or_grgo_create(gr.h, a.h))
    vals = np.linspace(1e-7, 3e-7, s.n)
    s.set(VAL, vals)
    assert L().or_pop_attach_syn(a.h, s.h) == 0
    for k in range(3):
        y = a.get(orc.Y)
        _, post = s.comp()
        inj = np.zeros(a.nc)
        np.add.at(inj, post, -(vals * (y[post] - 0.0)))
        b.set(orc.IEXT, inj)                 # b holds -i_ext: i_ext = -s gives the same b
        a.advance()
        b.advance()
        np.testing.assert_allclose(a.get(orc.VNEW), b.get(orc.VNEW), rtol=0, atol=1e-10)
    assert np.max(np.abs(a.get(orc.VNEW) - go_grid_step(ref_data))) > 1e-6


def go_grid_step(ref_data):
    c = go_grid(ref_data, 1, 1)
    for _ in range(3):
        c.advance()
    return c.get(orc.VNEW)


def test_rkc_granule_synaptic_term_in_the_frozen_current(ref_data):
    """An ion update of a granule population with synapses rebuilds the frozen
    current as the same population's without them plus the synaptic current
    val (y - E) at the contacts, y the voltage the update uses."""
    lib = orc.lib()
    lib.or_model_update_ion_and_b.argtypes = [ctypes.c_void_p, ctypes.c_double]
    a, b = gr_grid(ref_data, 2, 1), gr_grid(ref_data, 2, 1)
    mf = Syn(L().or_mfgr_create(2, 0))
    go = go_grid(ref_data, 1, 1)
    gg = Syn(L().or_gogr_create(go.h, a.h))
    mf.set(VAL, np.linspace(1e-6, 8e-6, mf.n))
    gg.set(VAL, np.linspace(2e-6, 3e-6, gg.n))
    assert L().or_pop_attach_syn(a.h, mf.h) == 0 and L().or_pop_attach_syn(a.h, gg.h) == 0
    y = a.get(orc.Y)
    y[::5] += 3.0
    a.set(orc.Y, y)
    b.set(orc.Y, y)
    lib.or_model_update_ion_and_b(a.h, 0.01)
    lib.or_model_update_ion_and_b(b.h, 0.01)
    want = b.get(orc.B)
    for s, e in ((mf, 0.0), (gg, -82.0)):
        _, post = s.comp()
        for k, c in enumerate(post):
            want[c] += s.get(VAL)[k] * (y[c] - e)
    np.testing.assert_allclose(a.get(orc.B), want, rtol=1e-14, atol=1e-20)
    assert np.max(np.abs(a.get(orc.B) - b.get(orc.B))) > 1e-6


def test_rkc_purkinje_synaptic_current_decays_over_the_ion_interval(ref_data):
    """Purkinje cells take the granule input of the macro step's start, s =
    sum val (y - E), into each ion update's frozen current scaled by
    exp(-dt / 0.6), dt the time since the last ion update (the 0.6 ms AMPA
    decay of solve_rkc.cu:1247)."""
    lib = orc.lib()
    lib.or_model_update_ion_and_b.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.or_syn_macro_start.argtypes = [ctypes.c_void_p]
    gr = gr_grid(ref_data, 2, 2)
    a = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=[-65.0], libc_rand=False)
    b = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=[-65.0], libc_rand=False)
    s = Syn(L().or_grpkj_create(gr.h, a.h, orc._dp(pkj_x(ref_data)), 5))
    vals = np.linspace(1e-6, 2e-6, s.n)
    s.set(VAL, vals)
    assert L().or_pop_attach_syn(a.h, s.h) == 0
    y = a.get(orc.Y)
    lib.or_syn_macro_start(a.h)
    _, post = s.comp()
    syn = np.zeros(a.nc)
    np.add.at(syn, post, vals * (y[post] - 0.0))
    for dt in (0.0078125, 0.05):
        lib.or_model_update_ion_and_b(a.h, dt)
        lib.or_model_update_ion_and_b(b.h, dt)
        np.testing.assert_allclose(a.get(orc.B) - b.get(orc.B), syn * np.exp(-dt / 0.6), rtol=1e-9, atol=1e-22)
