"""Pins of the oracle's backward-Euler (BE) path (oracle/or_cn.c and the BE
parts of the cell models, restating solve_bem.cu and the xx_update_ion /
go_KAHP_update / gr_Na_update kernels it calls).

* One BE step of a passive cable solves (C/DT + L) v' = (C/DT) v + I: the
  oracle's step against numpy's dense solve of that system, built here from
  the conductances (a halved matrix, a wrong sign or a missing capacitance
  term fails).
* Every cell type's BE diagonal addend and right-hand side against its frozen
  current (pinned for the RKC path): with the state at V_s,
  gsum V - esum = Sum_k g_k (V - E_k) - I_ext + Cm/DT (V - V_s) at any V, which
  a dropped or doubled term, or a missing capacitive part, breaks.
* The gate updates are exponential Euler: over a very long step every gate
  lands exactly on its steady state (the rate functions pinned for the RKC
  path, Ca-dependent gates at the step's starting Ca); two steps of DT equal
  one of 2 DT (a Crank-Nicolson or forward-Euler update is not a semigroup);
  for a short step the change agrees with the CN update's (pinned against the
  RKC path) to O((DT/tau)^2), which a wrong time constant breaks.
* The implicit Markov steps (Golgi KAHP, granule Na): occupancy is conserved;
  a short step moves the state as the CN step does (same rate matrix); a very
  long step reaches the scheme's stationary distribution, which the CN step
  then leaves in place.
* The inferior olive's gap term enters the dendrite's right-hand side as
  minus the gap current the RKC path adds to its frozen current.
"""
import ctypes
import os

import numpy as np
import pytest

import orc
from test_oracle_cn import generic, tree_laplacian, L as LCN


def L():
    lib = LCN()
    if not getattr(lib, "_be_bound", False):
        dp = ctypes.POINTER(ctypes.c_double)
        lib.or_be_init.argtypes = [ctypes.c_void_p]
        lib.or_be_init.restype = ctypes.c_int
        lib.or_be_step.argtypes = [ctypes.c_void_p]
        lib.or_be_advance.argtypes = [ctypes.c_void_p, ctypes.c_double]
        lib.or_be_assemble_test.argtypes = [ctypes.c_void_p, dp, dp]
        for k in ("pkj", "io", "go", "gr"):
            getattr(lib, f"or_{k}_be_ion").argtypes = [ctypes.c_void_p, ctypes.c_double]
            getattr(lib, f"or_{k}_be_ge").argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, dp, dp]
            getattr(lib, f"or_{k}_cn_ion").argtypes = [ctypes.c_void_p, ctypes.c_double]
        lib.or_go_kahp_be.argtypes = [ctypes.c_double, ctypes.c_double, dp]
        lib.or_go_kahp_be.restype = ctypes.c_int
        lib.or_go_kahp.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        lib.or_go_kahp.restype = ctypes.c_int
        lib.or_gr_na_be.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp]
        lib.or_gr_na_be.restype = ctypes.c_int
        lib.or_gr_na_step.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        lib.or_gr_na_step.restype = ctypes.c_int
        lib.or_io_update_b.argtypes = [ctypes.c_void_p, dp, ctypes.c_int]
        lib._be_bound = True
    return lib


DT = 0.025

# voltage-gated gates by oracle index: (index, rate-function id of the steady state)
PKJ_V = list(range(15)) + [16]                   # every gate but z_KC (15), z_K2 (17)
PKJ_CA = [15, 17]
IO_V = [0, 1, 2, 3, 4, 5, 6, 7, 9]               # all but s_KCa (8)
GO_V = list(range(18))
GO_V_NOCA = [g for g in GO_V if g != 8]          # c_KC (8) depends on Ca
GR_V = [0, 1, 2, 4, 5]                           # n_KV a_KA b_KA s_KM ir_KIR; c_KCa (3) at Ca
NGATE = {"pkj": 18, "io": 10, "go": 24, "gr": 20}


def make(ref_data, kind, v0, n=1):
    if kind == "pkj":
        return orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), n, 1, v0=[v0] * n, libc_rand=False)
    if kind == "io":
        from test_oracle_io import io_pop
        return io_pop(ref_data, n, 1, v0=[v0] * n)
    if kind == "go":
        from test_oracle_go import go_pop
        return go_pop(ref_data, n, 1, v0=[v0] * n)
    from test_oracle_gr import gr_pop
    return gr_pop(ref_data, n, 1, v0=[v0] * n)


def rate(kind, which, v, x):
    if kind == "pkj":
        return orc.lib().or_pkj_rate(which, v, x)
    if kind == "io":
        from test_oracle_io import L as LI
        return LI().or_io_rate(which, v, int(x))
    if kind == "go":
        from test_oracle_go import L as LG
        return LG().or_go_rate(which, v, x)
    from test_oracle_gr import L as LR
    return LR().or_gr_rate(which, v, x)


def scramble(p, gates, rng):
    """Move the listed gates off their steady state (uniform in (0.05, 0.95))."""
    for g in gates:
        p.set(orc.ION0 + g, rng.uniform(0.05, 0.95, p.nc))


def test_passive_cable_be_step_is_the_be_system():
    rng = np.random.default_rng(11)
    n = 9
    Lr = tree_laplacian(n, rng)
    cm = rng.uniform(0.5, 2.0, n) * 1e-6
    v0 = rng.uniform(-70.0, -60.0, n)
    p = generic(Lr, cm, v0)
    iext = np.zeros(n)
    iext[4] = 3e-5
    p.set(orc.IEXT, iext)
    assert L().or_be_init(p.h) == 0
    A = np.diag(cm / DT) - Lr                          # C/DT + conductance Laplacian
    L().or_be_step(p.h)
    v1 = p.get(orc.V)
    np.testing.assert_allclose(v1, np.linalg.solve(A, cm / DT * v0 + iext), rtol=1e-12)
    L().or_be_step(p.h)
    np.testing.assert_allclose(p.get(orc.V), np.linalg.solve(A, cm / DT * v1 + iext), rtol=1e-12)
    # the CN system differs (halved matrix), so the two steps must not agree
    q = generic(Lr, cm, v0)
    q.set(orc.IEXT, iext)
    assert LCN().or_cn_init(q.h) == 0
    LCN().or_cn_step(q.h)
    assert np.max(np.abs(q.get(orc.V) - v1)) > 1e-6


def be_sums(p, kind):
    fn = getattr(L(), f"or_{kind}_be_ge")
    g, e = ctypes.c_double(), ctypes.c_double()
    gs, es = np.zeros(p.nc), np.zeros(p.nc)
    for i in range(p.nc):
        fn(p.h, i, DT, ctypes.byref(g), ctypes.byref(e))
        gs[i], es[i] = g.value, e.value
    return gs, es


@pytest.mark.parametrize("kind", ["pkj", "io", "go", "gr"])
def test_be_sums_match_the_frozen_current(ref_data, kind):
    p = make(ref_data, kind, -60.0)
    rng = np.random.default_rng(12)
    iext = np.zeros(p.nc)
    iext[::5] = rng.uniform(-1e-4, 1e-4, len(iext[::5]))
    p.set(orc.IEXT, iext)
    scramble(p, {"pkj": PKJ_V, "io": IO_V, "go": GO_V, "gr": GR_V}[kind], rng)
    upd = {"pkj": lambda vp: orc.lib().or_pkj_update_b(p.h, vp),
           "io": lambda vp: L().or_io_update_b(p.h, vp, 0),
           "go": lambda vp: L().or_go_update_b(p.h, vp),
           "gr": lambda vp: L().or_gr_update_b(p.h, vp)}[kind]
    if kind == "go":
        from test_oracle_go import L as LG
        LG()
    if kind == "gr":
        from test_oracle_gr import L as LR
        LR()
    cm = p.get(orc.CM)
    for v in (-80.0, -55.0, 20.0):
        # state at v_s, currents frozen at v: gsum v - esum = b(v) + Cm/DT (v - v_s)
        vs = v - 3.0
        p.set(orc.V, np.full(p.nc, vs))
        gs, es = be_sums(p, kind)
        vv = np.full(p.nc, v)
        upd(orc._dp(vv))
        b = p.get(orc.B)
        want = b + cm / DT * (v - vs)
        scale = np.abs(gs * v) + np.abs(es) + 1e-30
        assert np.max(np.abs((gs * v - es) - want) / scale) < 1e-13, kind


@pytest.mark.parametrize("kind", ["pkj", "io", "go", "gr"])
def test_long_step_lands_on_the_steady_state(ref_data, kind):
    v = -52.0
    p = make(ref_data, kind, v)
    rng = np.random.default_rng(13)
    gates = {"pkj": PKJ_V + PKJ_CA, "io": IO_V + [8], "go": GO_V, "gr": GR_V + [3]}[kind]
    scramble(p, gates, rng)
    ca0 = p.get(orc.CA)
    getattr(L(), f"or_{kind}_be_ion")(p.h, 1.0e8)
    for i in range(min(p.nc, 40)):
        for g in gates:
            got = p.get(orc.ION0 + g)[i]
            if kind == "pkj" and g in PKJ_CA:       # KC / K2 see the quantised starting Ca
                caq = np.floor((ca0[i] * 1000 - 0.04) / 0.099986667) * 0.099986667 + 0.04
                want = rate(kind, 2 * g, v, caq)
            elif kind == "io":                       # s_KCa is a function of Ca; h_Na of the compartment
                want = rate(kind, 2 * g, ca0[i] if g == 8 else v, i % 3)
            else:                                    # GO c_KC and GR c_KCa at the starting Ca
                want = rate(kind, 2 * g, v, ca0[i])
            assert got == want, (kind, i, g, got, want)


@pytest.mark.parametrize("kind", ["pkj", "io", "go", "gr"])
def test_gate_update_is_a_semigroup(ref_data, kind):
    """exp(-2 DT / tau) = exp(-DT / tau)^2: two steps of DT equal one of 2 DT."""
    gates = {"pkj": PKJ_V, "io": IO_V, "go": GO_V_NOCA, "gr": GR_V}[kind]
    rng = np.random.default_rng(14)
    a, b = make(ref_data, kind, -47.0), make(ref_data, kind, -47.0)
    x0 = {g: rng.uniform(0.05, 0.95, a.nc) for g in gates}
    for q in (a, b):
        for g in gates:
            q.set(orc.ION0 + g, x0[g])
    fn = getattr(L(), f"or_{kind}_be_ion")
    fn(a.h, 0.3)
    fn(a.h, 0.3)
    fn(b.h, 0.6)
    for g in gates:
        xa, xb = a.get(orc.ION0 + g), b.get(orc.ION0 + g)
        np.testing.assert_allclose(xa, xb, rtol=0, atol=2e-14, err_msg=f"{kind} gate {g}")
        assert np.max(np.abs(xa - x0[g])) > 1e-6, (kind, g)


@pytest.mark.parametrize("kind", ["pkj", "io", "go", "gr"])
def test_short_step_moves_gates_as_the_cn_update(ref_data, kind):
    gates = {"pkj": PKJ_V + PKJ_CA, "io": IO_V + [8], "go": GO_V, "gr": GR_V + [3]}[kind]
    gates = [g for g in gates if not (kind == "io" and g == 2)]      # m_Na is instantaneous
    rng = np.random.default_rng(15)
    dt = 1e-4
    a, b = make(ref_data, kind, -49.0), make(ref_data, kind, -49.0)
    x0 = {g: rng.uniform(0.05, 0.95, a.nc) for g in gates}
    for q in (a, b):
        for g in gates:
            q.set(orc.ION0 + g, x0[g])
    assert LCN().or_cn_init(b.h) == 0
    getattr(L(), f"or_{kind}_be_ion")(a.h, dt)
    getattr(L(), f"or_{kind}_cn_ion")(b.h, dt)
    for g in gates:
        da, db = a.get(orc.ION0 + g) - x0[g], b.get(orc.ION0 + g) - x0[g]
        moving = np.abs(db) > 1e-12
        assert moving.any(), (kind, g)
        rel = np.abs(da[moving] - db[moving]) / np.abs(db[moving])
        assert np.max(rel) < 2e-2, (kind, g, np.max(rel))


def test_kahp_implicit_step(ref_data):
    lib = L()
    rng = np.random.default_rng(16)
    for ca in (5e-5, 2e-4, 3e-3):
        x0 = rng.uniform(0.0, 1.0, 6)
        x0 /= x0.sum()
        x = x0.copy()
        assert lib.or_go_kahp_be(ca, DT, orc._dp(x)) == 0
        assert abs(x.sum() - 1.0) < 1e-14
        # short step: the same motion as the CN step with the same Ca
        dt = 1e-4
        xb, xc = x0.copy(), x0.copy()
        lib.or_go_kahp_be(ca, dt, orc._dp(xb))
        lib.or_go_kahp(ca, ca, dt, orc._dp(xc))
        d = np.max(np.abs(xc - x0))
        np.testing.assert_allclose(xb - x0, xc - x0, rtol=0, atol=2e-3 * d)
        # many long steps: the stationary distribution, left in place by a CN step
        xs = x0.copy()
        for _ in range(60):
            lib.or_go_kahp_be(ca, 1e3, orc._dp(xs))
        assert abs(xs.sum() - 1.0) < 1e-12 and xs.min() >= 0.0
        xcn = xs.copy()
        lib.or_go_kahp(ca, ca, DT, orc._dp(xcn))
        np.testing.assert_allclose(xcn, xs, rtol=0, atol=1e-14)
        assert np.max(np.abs(xs - x0)) > 1e-3


def test_na_implicit_step(ref_data):
    lib = L()
    rng = np.random.default_rng(17)
    for t in (5, 7):
        for v in (-70.0, -40.0, 10.0):
            x0 = rng.uniform(0.0, 1.0, 12)
            x0 /= x0.sum()
            x = x0.copy()
            assert lib.or_gr_na_be(t, v, DT, orc._dp(x)) == 0
            assert abs(x.sum() - 1.0) < 1e-13
            dt = 1e-7                      # dt times the fastest rate (~1e3 / ms at +10 mV) stays ~1e-4
            xb, xc = x0.copy(), x0.copy()
            lib.or_gr_na_be(t, v, dt, orc._dp(xb))
            lib.or_gr_na_step(t, v, v, dt, orc._dp(xc))
            # agree to second order in dt: relative to the step's largest move
            d = np.max(np.abs(xc - x0))
            np.testing.assert_allclose(xb - x0, xc - x0, rtol=0, atol=2e-3 * d)
            xs = x0.copy()
            for _ in range(400):
                lib.or_gr_na_be(t, v, 10.0, orc._dp(xs))
            assert abs(xs.sum() - 1.0) < 1e-10 and xs.min() >= -1e-15
            xcn = xs.copy()
            lib.or_gr_na_step(t, v, v, DT, orc._dp(xcn))
            np.testing.assert_allclose(xcn, xs, rtol=0, atol=1e-14)
    assert lib.or_gr_na_be(2, -60.0, DT, orc._dp(np.zeros(12))) == 2


def test_io_gap_term_on_the_be_right_hand_side(ref_data):
    from test_oracle_io import io_pop
    p = io_pop(ref_data, 2, 1, v0=[-62.0, -55.0])
    assert L().or_be_init(p.h) == 0
    diag, b = np.zeros(p.nc), np.zeros(p.nc)
    L().or_be_assemble_test(p.h, orc._dp(diag), orc._dp(b))
    gs, es = be_sums(p, "io")
    row, col, val = p.csr()
    dpos = [next(j for j in range(row[i], row[i + 1]) if col[j] == i) for i in range(p.nc)]
    np.testing.assert_array_equal(diag, -val[dpos] + gs)           # unhalved matrix diagonal + sums
    vv = p.get(orc.V)
    L().or_io_update_b(p.h, orc._dp(vv), 1)
    with_gap = p.get(orc.B)
    L().or_io_update_b(p.h, orc._dp(vv), 0)
    gap = with_gap - p.get(orc.B)                                 # what the RKC path adds to b
    assert abs(gap[0]) > 1e-6 and abs(gap[3]) > 1e-6
    assert gap[0] < 0 < gap[3]                                    # current flows from cell 1 into cell 0
    np.testing.assert_allclose(b - es, -gap, rtol=1e-12, atol=1e-18)
    np.testing.assert_array_equal((b - es)[[1, 2, 4, 5]], 0.0)


@pytest.mark.parametrize("kind", ["pkj", "io", "go", "gr"])
def test_be_interval_runs_five_steps(ref_data, kind):
    p = make(ref_data, kind, -60.0)
    assert L().or_be_init(p.h) == 0
    v0 = p.get(orc.V)
    L().or_be_advance(p.h, 0.0)
    v1 = p.get(orc.V)
    assert np.all(np.isfinite(v1))
    assert np.max(np.abs(v1 - v0)) > 0.0
    assert p.stats()["steps"] == 5


@pytest.mark.parametrize("kind", ["pkj", "io", "go", "gr"])
def test_ca_update_against_the_cn_path(ref_data, kind):
    """At constant V the BE Ca step is the CN path's Ca step (pinned against the
    RKC path) with the reference's BE differences made explicit: the Purkinje
    shell volume multiplied by the area once more, the olive's Ca current taken
    with the already-updated r_CaH gate; the Golgi pools and reversal are the
    same; granule Ca does not move."""
    v, dt = -20.0, 0.025
    rng = np.random.default_rng(18)
    a, b = make(ref_data, kind, v), make(ref_data, kind, v)
    gates = {"pkj": PKJ_V, "io": IO_V, "go": GO_V, "gr": GR_V}[kind]
    x0 = {g: rng.uniform(0.2, 0.9, a.nc) for g in gates}
    for q in (a, b):
        for g in gates:
            q.set(orc.ION0 + g, x0[g])
    ca0 = a.get(orc.CA)
    getattr(L(), f"or_{kind}_be_ion")(a.h, dt)
    if kind == "pkj":
        b.set(orc.SHELL, b.get(orc.SHELL) * b.get(orc.AREA))
    if kind == "io":
        b.set(orc.ION0 + 7, a.get(orc.ION0 + 7))          # r_CaH after the BE gate update
    assert LCN().or_cn_init(b.h) == 0
    getattr(L(), f"or_{kind}_cn_ion")(b.h, dt)
    ca_a, ca_b = a.get(orc.CA), b.get(orc.CA)
    if kind == "gr":
        np.testing.assert_array_equal(ca_a, ca0)
        return
    assert np.max(np.abs(ca_a - ca0)) > 0.0
    np.testing.assert_array_equal(ca_a, ca_b)
    if kind == "go":
        np.testing.assert_array_equal(a.get(108), b.get(108))          # second pool
        np.testing.assert_array_equal(a.get(orc.REV), b.get(orc.REV))
    if kind == "pkj":
        # reversal from the starting Ca; the fixed value on the soma type
        rev = a.get(orc.REV)
        c = LCN()
        z = make(ref_data, "pkj", v)
        z.set(orc.CA, ca0)
        assert c.or_cn_init(z.h) == 0
        L().or_pkj_cn_ion(z.h, 0.0)                             # DT = 0: Ca unchanged, mid-step Ca = start
        zr = z.get(orc.REV)
        np.testing.assert_array_equal(rev[1:], zr[1:])
        # compartment 1599 is the soma (fixed reversal on both paths); the CN /
        # RKC paths also fix compartment 0 (index multiple of 1599), BE does not
        fixed = 12.5 * np.log(2.4 / 4e-5)
        assert rev[1599] == zr[1599] == fixed
        assert zr[0] == fixed and rev[0] != fixed


def test_truncated_cg_noise_floor(ref_data):
    """The oracle's own spread on the BE workloads whose CG stops at the
    100-iteration cap far from convergence (Purkinje, granule cells): a
    1e-12 mV change of the initial voltage moves the result by 1e-7 .. 2e-4
    mV (Purkinje, 16 output intervals) and 1e-5 .. 5e-3 mV (granule cells,
    first output interval).  The GPU parity bar for Purkinje cells (1e-3 mV
    every interval) sits a factor of 5 above the first; granule cells are
    compared by spike times after the first interval."""
    from test_oracle_gr import gr_pop
    pk = []
    for eps in (0.0, 1e-12):
        o = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=[-65.0 + eps], libc_rand=False)
        assert L().or_be_init(o.h) == 0
        o.set_stim(1599, 0.5, 1e9, 2e-3)
        tr = []
        for k in range(16):
            L().or_be_advance(o.h, k * 0.125)
            tr.append(o.get(orc.V))
        pk.append(np.array(tr))
        assert o.stats()["stages"] == 100 * o.stats()["steps"]
    d = np.max(np.abs(pk[0] - pk[1]))
    assert 1e-7 < d < 2e-4, d
    gr = []
    for eps in (0.0, 1e-12):
        o = gr_pop(ref_data, 2, 1, v0=np.array([-70.0, -72.0]) + eps)
        assert L().or_be_init(o.h) == 0
        L().or_be_advance(o.h, 0.0)
        gr.append(o.get(orc.V))
        assert o.stats()["stages"] == 500
    d = np.max(np.abs(gr[0] - gr[1]))
    assert 1e-5 < d < 5e-3, d
