"""Pins of the oracle's Golgi-cell model (oracle/or_go.c, restating go.cu,
go_solve.cu, go_ion.cu and the GO parts of solve_rkc.cu).

The reference holds no Golgi-cell output to compare with, so each piece is
pinned by what the mathematics fixes, chosen so that a wrong sign, a swapped
operand or a transposed rate fails:

* rate functions: bounds, the limits at their removable 0/0 points (a sign
  error in an exponent moves the limit), and the half-activation voltages of
  the Boltzmann steady states;
* gates: the Crank-Nicolson update at constant voltage is the (1,1) Pade form
  of exp(-dt/tau);
* Ca pools: without Ca current both RK4 updates are the Taylor-4 decay to the
  resting level, and a pool at the fixed point of a constant current stays
  there; the second pool's reversal is the Nernst potential of a divalent ion
  at 23 C from physical constants (to 0.05 %; 37 C would be off by 4.7 %);
* KAHP: one step equals the Crank-Nicolson step of the generator matrix
  written independently from the scheme's state diagram, total occupancy is
  conserved, and many steps at fixed Ca converge to the generator's
  stationary distribution;
* cable matrix: series half-segment conductances from Go_131.txt geometry,
  zero row sums, symmetry, columns ascending; gap junctions: one entry of
  2.5e-6 mS per ordered pair of grid neighbours' compartments 0;
* initial state: gates at their steady state, KAHP closed, V = -60 + 2 (U - 0.5)
  with U the cell's libc rand() draw in order.

Whole-cell trajectories have no reference fixture: parity unpinned beyond
these parts (the shared RKC core is pinned in test_oracle_rkc.py).
"""
import ctypes
import math
import os

import numpy as np
import pytest

import orc

GO_GATES = ["m_NaT", "h_NaT", "r_NaR", "s_NaR", "p_NaP", "n_KV", "a_KA", "b_KA", "c_KC", "sl_Kslow",
            "ch_CaHVA", "ci_CaHVA", "cl_CaLVA", "cm_CaLVA", "hf_HCN1", "hf_HCN2", "hs_HCN1", "hs_HCN2",
            "o1_KAHP", "o2_KAHP", "c1_KAHP", "c2_KAHP", "c3_KAHP", "c4_KAHP"]
G = {n: i for i, n in enumerate(GO_GATES)}
VOLTAGE_GATES = GO_GATES[:18]
CA2, CA_OLD = 108, 109
COMP = 131


def L():
    lib = orc.lib()
    if not getattr(lib, "_go_bound", False):
        dp = ctypes.POINTER(ctypes.c_double)
        lib.or_go_create.restype = ctypes.c_void_p
        lib.or_go_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int]
        lib.or_go_rate.restype = ctypes.c_double
        lib.or_go_rate.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        lib.or_go_update_ion.argtypes = [ctypes.c_void_p, dp, ctypes.c_double]
        lib.or_go_update_b.argtypes = [ctypes.c_void_p, dp]
        lib.or_go_kahp.restype = ctypes.c_int
        lib.or_go_kahp.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        lib._go_bound = True
    return lib


def go_pop(ref_data, nx, ny, v0=None, seed=0):
    lib = L()
    lib.srand(seed)
    vp = None
    if v0 is not None:
        v0 = np.ascontiguousarray(v0, dtype=np.float64)
        vp = orc._dp(v0)
    h = lib.or_go_create(os.path.join(ref_data, "Go_131.txt").encode(), nx, ny, vp, 1e-6, 1e-6,
                         0 if v0 is not None else 1)
    assert h
    return orc.Pop(h, nx * ny, COMP)


def rate(g, v, ca=1e-4):
    return L().or_go_rate(2 * G[g], v, ca), L().or_go_rate(2 * G[g] + 1, v, ca)


def ab(g, v, ca=1e-4):
    return L().or_go_rate(1000 + 2 * G[g], v, ca), L().or_go_rate(1000 + 2 * G[g] + 1, v, ca)


def test_rates_bounded_and_positive():
    vs = np.linspace(-110.0, 50.0, 161) + 0.0137   # off the 0/0 points
    for g in VOLTAGE_GATES:
        for v in vs:
            for ca in (5e-5, 1e-3):
                x, t = rate(g, v, ca)
                # the HCN1 fast fraction 0.0021 v + 0.97 is not clamped
                # (go_ion.cu:71): above 14.3 mV it exceeds 1 and the slow
                # steady state turns (slightly) negative
                lo = (1.0 - (0.0021 * v + 0.97)) if g == "hs_HCN1" and v > 14.28 else 0.0
                assert lo <= x <= 1.0 and t > 0.0 and math.isfinite(t), (g, v, ca, x, t)


@pytest.mark.parametrize("gate,ab_index,v0,limit", [
    ("m_NaT", 0, -25.0, 0.417 * 10.0),           # 0.417 x / (1 - e^{-x/10})
    ("p_NaP", 0, -40.0, 0.421 * 5.0),            # 0.421 x / (1 - e^{-x/5})
    ("p_NaP", 1, -40.0, 0.287 * 5.0),            # -0.287 x / (1 - e^{x/5})
    ("n_KV", 0, -26.0, 0.062 * 10.0),            # 0.062 x / (1 - e^{-x/10})
    ("r_NaR", 0, 4.48, (1.11 + 68.5 * 6.8) * 1e-4),
    ("r_NaR", 1, -44.0, (66.0 + 21.7 * 0.11) * 1e-3),
])
def test_removable_singularities(gate, ab_index, v0, limit):
    for dv in (1e-6, -1e-6, 1e-4, -1e-4):
        got = ab(gate, v0 + dv)[ab_index]
        assert got == pytest.approx(limit, rel=2e-4), (gate, dv, got)


@pytest.mark.parametrize("gate,vhalf,activating", [
    ("p_NaP", -43.0, True), ("cl_CaLVA", -52.0, True), ("cm_CaLVA", -80.0, False),
    ("a_KA", -38.0, True), ("b_KA", -78.8, False), ("sl_Kslow", -35.0, True),
])
def test_boltzmann_half_activation(gate, vhalf, activating):
    assert rate(gate, vhalf)[0] == pytest.approx(0.5, abs=1e-15)
    lo, hi = rate(gate, vhalf - 10.0)[0], rate(gate, vhalf + 10.0)[0]
    assert (hi > 0.5 > lo) if activating else (lo > 0.5 > hi)


def test_hcn_components_split_one_boltzmann():
    # fast + slow steady states of each HCN channel add up to one Boltzmann
    # curve (half-activation -72.5 mV for HCN1, -81.9 mV for HCN2); the HCN2
    # fast fraction is clamped to [0, 1] outside [-108.7, -64.7] mV
    for v in np.linspace(-120.0, -40.0, 81) + 0.01:
        s1 = rate("hf_HCN1", v)[0] + rate("hs_HCN1", v)[0]
        s2 = rate("hf_HCN2", v)[0] + rate("hs_HCN2", v)[0]
        assert s1 == pytest.approx(1.0 / (1.0 + math.exp((v + 72.5) * 0.11)), rel=1e-14)
        assert s2 == pytest.approx(1.0 / (1.0 + math.exp((v + 81.9) * 0.16)), rel=1e-14)
    assert ab("hf_HCN2", -60.0)[0] == 0.0 and ab("hf_HCN2", -120.0)[0] == 1.0
    assert ab("hf_HCN2", -80.0)[0] == pytest.approx(0.0227 * 80.0 - 1.47, rel=1e-14)


def pade11(z):
    return (1.0 + 0.5 * z) / (1.0 - 0.5 * z)


def kinetic_generator(ca):
    """Generator Q (dx/dt = Q x) of the KAHP scheme from its state diagram:
    c1 <-> c2 <-> c3 <-> c4, c3 <-> o1, c4 <-> o2, Ca-dependent forward rates
    in the closed chain."""
    idx = {"c1": 0, "c2": 1, "c3": 2, "c4": 3, "o1": 4, "o2": 5}
    edges = [("c1", "c2", 200.0 * ca / 3.0), ("c2", "c1", 0.08),
             ("c2", "c3", 160.0 * ca / 3.0), ("c3", "c2", 0.08),
             ("c3", "c4", 80.0 * ca / 3.0), ("c4", "c3", 0.2),
             ("c3", "o1", 0.16), ("o1", "c3", 1.0),
             ("c4", "o2", 1.2), ("o2", "c4", 0.1)]
    Q = np.zeros((6, 6))
    for a, b, r in edges:
        Q[idx[b], idx[a]] += r
        Q[idx[a], idx[a]] -= r
    return Q


@pytest.mark.parametrize("ca_old,ca,dt", [(5e-5, 5e-5, 2.0 ** -6), (5e-5, 3e-3, 2.0 ** -6), (2e-3, 4e-4, 0.05)])
def test_kahp_step_is_crank_nicolson_of_the_scheme(ca_old, ca, dt):
    rng = np.random.default_rng(3)
    x = rng.uniform(0.0, 1.0, 6)
    x /= x.sum()
    st = x.copy()
    assert L().or_go_kahp(ca_old, ca, dt, orc._dp(st)) == 0
    I = np.eye(6)
    want = np.linalg.solve(I - 0.5 * dt * kinetic_generator(ca), (I + 0.5 * dt * kinetic_generator(ca_old)) @ x)
    np.testing.assert_allclose(st, want, rtol=1e-11, atol=1e-15)
    assert st.sum() == pytest.approx(1.0, abs=1e-14)


def test_kahp_converges_to_stationary_distribution():
    ca, dt = 1e-3, 2.0 ** -6
    st = np.array([1.0, 0, 0, 0, 0, 0])
    for _ in range(200000):
        L().or_go_kahp(ca, ca, dt, orc._dp(st))
    w, V = np.linalg.eig(kinetic_generator(ca))
    pi = np.real(V[:, np.argmin(np.abs(w))])
    pi /= pi.sum()
    np.testing.assert_allclose(st, pi, rtol=1e-9, atol=1e-14)


@pytest.mark.parametrize("dt", [2.0 ** -6, 0.05])
def test_cn_gates_and_calcium_decay(ref_data, dt):
    p = go_pop(ref_data, 1, 1, v0=[-55.0])
    rng = np.random.default_rng(11)
    x = {g: rng.uniform(0.05, 0.95, p.nc) for g in VOLTAGE_GATES}
    x["ch_CaHVA"] = np.zeros(p.nc)      # no Ca current into either pool
    x["cl_CaLVA"] = np.zeros(p.nc)
    for g in VOLTAGE_GATES:
        p.set(orc.ION0 + G[g], x[g])
    ca0 = rng.uniform(1e-4, 2e-3, p.nc)
    ca20 = rng.uniform(1e-4, 2e-3, p.nc)
    p.set(orc.CA, ca0)
    p.set(CA2, ca20)
    v = np.full(p.nc, -55.0)
    p.set(orc.V, v)
    L().or_go_update_ion(p.h, orc._dp(v), dt)
    z = -1.3 * dt
    t4 = 1.0 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    ca_rest = 5e-5
    ca1 = ca_rest + (ca0 - ca_rest) * t4
    np.testing.assert_allclose(p.get(orc.CA), ca1, rtol=1e-12)
    np.testing.assert_allclose(p.get(CA2), ca_rest + (ca20 - ca_rest) * t4, rtol=1e-12)
    # Taylor-4 is within the RK4 truncation bound of the exact decay
    assert np.all(np.abs(t4 - math.exp(z)) <= abs(z) ** 5 / 120 * 1.01)
    for g in VOLTAGE_GATES:
        new = p.get(orc.ION0 + G[g])
        for i in range(0, p.nc, 7):
            caa = 0.5 * (ca0[i] + ca1[i])
            xi, ti = rate(g, -55.0, caa)
            assert new[i] == pytest.approx(xi + pade11(-dt / ti) * (x[g][i] - xi), rel=1e-12, abs=1e-15), (g, i)


def test_reversal_is_divalent_nernst_at_23C(ref_data):
    p = go_pop(ref_data, 1, 1, v0=[-60.0])
    ca2 = np.geomspace(1e-5, 1e-2, p.nc)
    p.set(CA2, ca2)
    p.set(orc.ION0 + G["cl_CaLVA"], np.zeros(p.nc))
    v = np.full(p.nc, -60.0)
    p.set(orc.V, v)
    L().or_go_update_ion(p.h, orc._dp(v), 2.0 ** -6)
    ca2n = p.get(CA2)
    rev = p.get(orc.REV)
    k = rev / np.log(2.0 / ca2n)
    # R T / (2 F) at 23 C from CODATA constants, in mV
    nernst = 8.314462618 * (23.0 + 273.15) / (2.0 * 96485.33212) * 1e3
    np.testing.assert_allclose(k, nernst, rtol=5e-4)
    assert abs(k[0] / (8.314462618 * 310.15 / (2 * 96485.33212) * 1e3) - 1.0) > 0.04


def test_calcium_fixed_point_of_constant_current(ref_data):
    p = go_pop(ref_data, 1, 1, v0=[-60.0])
    soma = 30
    # I_Ca1 in mA/cm^2 from the soma's CaHVA density and gates
    area = p.get(orc.AREA)[soma]
    g_cahva = p.get(orc.COND0 + 8)[soma] / area
    ch, ci = 0.3, 0.6
    i_ca = 1e-3 * g_cahva * ch * ch * ci * (-60.0 - 120.0)
    ca_star = 5e-5 - i_ca / (2.0 * 9.6485e4 * 0.2e-4 * 1.3)
    ca = p.get(orc.CA)
    ca[soma] = ca_star
    p.set(orc.CA, ca)
    for g, x in (("ch_CaHVA", ch), ("ci_CaHVA", ci)):
        a = p.get(orc.ION0 + G[g])
        a[soma] = x
        p.set(orc.ION0 + G[g], a)
    v = np.full(p.nc, -60.0)
    p.set(orc.V, v)
    L().or_go_update_ion(p.h, orc._dp(v), 0.05)
    assert p.get(orc.CA)[soma] == pytest.approx(ca_star, rel=1e-13)


def read_go(ref_data):
    rows = np.loadtxt(os.path.join(ref_data, "Go_131.txt"))
    assert rows.shape == (COMP, 20)
    return rows


def test_cable_matrix_from_geometry(ref_data):
    rows = read_go(ref_data)
    p = go_pop(ref_data, 1, 1, v0=[-60.0])
    row, col, val = p.csr()
    r = 0.5 * rows[:, 2] * 1e-4
    ln = rows[:, 3] * 1e-4
    ra = rows[:, 4] * 1e-3
    R = ra * ln / (math.pi * r * r)          # axial resistance of each cylinder [kOhm]
    G_ = {}
    for i in range(COMP):
        d = int(rows[i, 1])
        if d >= 0:
            G_[(i, d)] = G_[(d, i)] = 1.0 / (0.5 * R[i] + 0.5 * R[d])
    seen = set()
    for i in range(COMP):
        cols = col[row[i]:row[i + 1]]
        assert list(cols) == sorted(cols)
        s = 0.0
        for k in range(row[i], row[i + 1]):
            if col[k] != i:
                assert val[k] == pytest.approx(G_[(i, col[k])], rel=1e-13)
                seen.add((i, int(col[k])))
            s += val[k]
        assert abs(s) <= 1e-15 * max(abs(v) for v in val[row[i]:row[i + 1]])
    assert seen == set(G_)
    np.testing.assert_allclose(p.get(orc.CM), rows[:, 5] * 2.0 * math.pi * r * ln, rtol=1e-14)
    np.testing.assert_allclose(p.get(orc.AREA), 2.0 * math.pi * r * ln, rtol=1e-14)
    for q in range(13):
        np.testing.assert_allclose(p.get(orc.COND0 + q), rows[:, 7 + q] * 2.0 * math.pi * r * ln, rtol=1e-14)


@pytest.mark.parametrize("nx,ny", [(2, 1), (2, 3), (3, 3)])
def test_grid_gap_junctions(ref_data, nx, ny):
    p = go_pop(ref_data, nx, ny, v0=[-60.0] * (nx * ny))
    row, col, val = p.csr()
    n = nx * ny
    gj = {}
    for i in range(p.nc):
        for k in range(row[i], row[i + 1]):
            if col[k] // COMP != i // COMP:
                gj[(i, int(col[k]))] = val[k]
    assert len(gj) == 2 * ((nx - 1) * ny + (ny - 1) * nx)
    for (a, b), g in gj.items():
        assert a % COMP == 0 and b % COMP == 0 and g == 2.5e-6 and gj[(b, a)] == g
        ca, cb = a // COMP, b // COMP
        assert abs(ca % nx - cb % nx) + abs(ca // nx - cb // nx) == 1
    for c in range(n):
        i = c * COMP
        assert abs(sum(val[row[i]:row[i + 1]])) < 1e-15


def test_initial_state(ref_data):
    lib = L()
    p = go_pop(ref_data, 3, 1, seed=0)
    lib.srand(0)
    draws = [lib.rand() for _ in range(3)]
    v = p.get(orc.V)
    for c in range(3):
        want = -60.0 + 2.0 * (draws[c] / 2147483647.0 - 0.5)
        np.testing.assert_array_equal(v[c * COMP:(c + 1) * COMP], want)
    for g in VOLTAGE_GATES:
        x = p.get(orc.ION0 + G[g])
        for c in range(3):
            assert x[c * COMP] == rate(g, v[c * COMP], 5e-5)[0]
    assert np.all(p.get(orc.ION0 + G["c1_KAHP"]) == 1.0)
    for g in ("o1_KAHP", "o2_KAHP", "c2_KAHP", "c3_KAHP", "c4_KAHP"):
        assert np.all(p.get(orc.ION0 + G[g]) == 0.0)
    assert np.all(p.get(orc.REV) == 120.0) and np.all(p.get(CA2) == 5e-5) and np.all(p.get(CA_OLD) == 5e-5)


def test_single_cell_fires_and_current_raises_rate(ref_data):
    """Property (no fixture): the Golgi cell is a pacemaker, and the 0.2 nA
    somatic step current raises its rate."""
    p = go_pop(ref_data, 1, 1, v0=[-60.0])
    prev, spikes = 100.0, []
    for k in range(int(500 / 0.125)):
        t = k * 0.125
        p.set_current(t)
        p.advance()
        v = p.get(orc.VNEW)[30]
        if v > 10.0 and prev <= 10.0:
            spikes.append(t + 0.125)
        prev = v
    assert len(spikes) >= 2
    p2 = go_pop(ref_data, 1, 1, v0=[-60.0])
    p2.set_stim(30, 0.0, 1e9, 0.2e-3)
    prev, s2 = 100.0, []
    for k in range(int(500 / 0.125)):
        t = k * 0.125
        p2.set_current(t)
        p2.advance()
        v = p2.get(orc.VNEW)[30]
        if v > 10.0 and prev <= 10.0:
            s2.append(t + 0.125)
        prev = v
    assert len(s2) > len(spikes)


def grid_spikes(ref_data, v0, steps=2300, stim=(30, 150.0, 300.0, 0.2e-3)):
    o = go_pop(ref_data, 2, 2, v0=v0)
    o.set_stim(*stim)
    prev, out = np.full(4, 100.0), []
    for k in range(steps):
        o.set_current(k * 0.125)
        o.advance()
        v = o.get(orc.VNEW)[30::COMP]
        for c in np.nonzero((v > 10.0) & (prev <= 10.0))[0]:
            out.append((0.125 * (k + 1), int(c)))
        prev = v
    return out


def test_noise_floor_of_the_grid_workload(ref_data):
    """The oracle's own spread on the 2 x 2 grid workload of the GPU parity
    test: initial voltages changed by one part in 1e15 keep every spike count
    but move spikes during the step current by up to two macro steps.  This
    is the bar the GPU comparison uses."""
    v0 = np.linspace(-61.0, -59.0, 4)
    a, b = grid_spikes(ref_data, v0), grid_spikes(ref_data, v0 * (1 + 1e-15))
    shift = 0.0
    for c in range(4):
        x = [t for t, cc in a if cc == c]
        y = [t for t, cc in b if cc == c]
        assert len(x) == len(y) >= 4
        assert [t for t in x if t < 150.0] == [t for t in y if t < 150.0]
        shift = max(shift, float(np.max(np.abs(np.array(x) - np.array(y)))))
    assert shift == 0.25
