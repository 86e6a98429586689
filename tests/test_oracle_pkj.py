"""Pins of the oracle's Purkinje model (oracle/or_pkj.c, restating pkj.cu,
pkj_solve.cu and pkj_ion.cu) against what the physics and the mathematics
fix independently of the reference text:

* rate functions (pkj_ion.cu:4-91): steady states in [0, 1] (0.8 / 0.2 for
  the two Kh components), positive time constants, the direction of every
  activation / inactivation curve, and continuity through the removable
  singularities the reference special-cases (pkj_ion.cu:53-54, 79-80);
* the CN gate update (pkj_ion.cu:337-356) at constant V is the (1,1) Pade
  approximant of exp(-dt / tau) around the steady state;
* the Ca shell RK4 (pkj_ion.cu:319-325) with the Ca channels closed is the
  4th-order Taylor polynomial of the exact exponential decay; an inward
  current only raises Ca and an outward one is clamped to zero;
* the Ca quantisation (pkj_ion.cu:335) lands on the 0.099986667 grid below
  the mid-step Ca, and the reversal potential is the Nernst potential at
  37 C from physical constants (pkj_ion.cu:329-331);
* the cable matrix (pkj_solve.cu:68-74) equals the axial conductance of two
  half-segments in series computed here from Pkj.txt's geometry, it is
  symmetric and a uniform voltage carries no axial current; capacitances
  follow the sphere / cylinder + spine areas of pkj.cu:60-90.
"""
import math
import os

import numpy as np
import pytest

import orc

GATES = ["m_NaF", "h_NaF", "m_NaP", "m_CaP", "h_CaP", "m_CaT", "h_CaT", "m_KA", "h_KA", "m_KC", "m_Kh1",
         "m_Kh2", "m_Kdr", "h_Kdr", "m_KM", "z_KC", "m_K2", "z_K2"]
G = {n: i for i, n in enumerate(GATES)}
CA0 = 4.0e-5


def inf(g, v, ca=1e-4):
    return orc.lib().or_pkj_rate(2 * G[g], v, ca)


def tau(g, v, ca=1e-4):
    return orc.lib().or_pkj_rate(2 * G[g] + 1, v, ca)


@pytest.fixture(scope="module")
def pkj(ref_data):
    return orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=[-65.0], libc_rand=False)


def test_rates_bounded_and_positive():
    for g in GATES:
        top = 0.8 if g == "m_Kh1" else 0.2 if g == "m_Kh2" else 1.0
        for v in np.linspace(-100.0, 60.0, 321):
            for ca in (CA0, 1e-3, 0.05):
                x, t = inf(g, v, ca), tau(g, v, ca)
                assert math.isfinite(x) and math.isfinite(t), (g, v)
                assert -1e-15 <= x <= top + 1e-15, (g, v, x)
                assert t > 0.0, (g, v, t)


@pytest.mark.parametrize("g", ["m_NaF", "m_NaP", "m_CaP", "m_CaT", "m_KA", "m_Kdr", "m_KM"])
def test_activation_rises_with_voltage(g):
    x = [inf(g, v) for v in np.linspace(-90.0, 40.0, 131)]
    assert np.all(np.diff(x) >= -1e-15) and x[-1] > x[0]


@pytest.mark.parametrize("g", ["h_NaF", "h_CaP", "h_CaT", "h_KA", "h_Kdr", "m_Kh1", "m_Kh2"])
def test_inactivation_falls_with_voltage(g):
    x = [inf(g, v) for v in np.linspace(-90.0, 40.0, 131)]
    assert np.all(np.diff(x) <= 1e-15) and x[-1] < x[0]


@pytest.mark.parametrize("g", ["z_KC", "z_K2"])
def test_calcium_gates_rise_with_calcium(g):
    x = [inf(g, -60.0, ca) for ca in np.geomspace(1e-5, 1.0, 60)]
    assert np.all(np.diff(x) > 0.0)


@pytest.mark.parametrize("which,v0", [(2 * 12 + 1, -12.0), (2 * 12, 8.0)])
def test_removable_singularities_are_continuous(which, v0):
    f = orc.lib().or_pkj_rate
    at = f(which, v0, 1e-4)
    for d in (1e-4, 1e-3):
        mid = 0.5 * (f(which, v0 - d, 1e-4) + f(which, v0 + d, 1e-4))
        assert at == pytest.approx(mid, rel=1e-6)


def pade11(z):
    return (1.0 + 0.5 * z) / (1.0 - 0.5 * z)


def fresh_state(p, v, ca, seed):
    rng = np.random.default_rng(seed)
    nc = p.nc
    p.set(orc.V, np.full(nc, v))
    p.set(orc.CA, np.full(nc, ca))
    x = {}
    for g in GATES:
        x[g] = rng.uniform(0.05, 0.95, nc)
        p.set(orc.ION0 + G[g], x[g])
    return x


@pytest.mark.parametrize("dt", [2.0 ** -8, 2.0 ** -6, 0.05])
def test_cn_gate_update_is_pade_at_constant_voltage(pkj, dt):
    v = -58.0
    x = fresh_state(pkj, v, 2e-4, seed=int(1 / dt))
    ca_old = pkj.get(orc.CA).copy()
    pkj.update_ion(np.full(pkj.nc, v), dt)
    for g in GATES:
        new = pkj.get(orc.ION0 + G[g])
        if g.startswith("z_"):
            continue
        xi, ti = inf(g, v), tau(g, v)
        np.testing.assert_allclose(new, xi + pade11(-dt / ti) * (x[g] - xi), rtol=1e-12, atol=1e-15)
    # z gates relax toward the steady state at the quantised mid-step Ca
    ca_mid = 0.5 * (ca_old + pkj.get(orc.CA))
    for g in ("z_KC", "z_K2"):
        new = pkj.get(orc.ION0 + G[g])
        R = pade11(-dt / tau(g, v))
        xinf = (new - R * x[g]) / (1.0 - R)
        k = 4.0 if g == "z_KC" else 0.2
        ca_q = k * xinf / (1.0 - xinf)       # inverts 1 / (1 + k / Ca)
        steps = (ca_q - 0.04) / 0.099986667
        assert np.allclose(steps, np.round(steps), atol=1e-6)
        below = ca_mid * 1000.0 - ca_q
        assert np.all(below >= -1e-9) and np.all(below < 0.099986667 + 1e-9)


@pytest.mark.parametrize("dt", [2.0 ** -8, 2.0 ** -6, 0.1])
def test_calcium_rk4_decay_is_taylor4(pkj, dt):
    fresh_state(pkj, -60.0, 3e-4, seed=3)
    nc = pkj.nc
    pkj.set(orc.ION0 + G["m_CaP"], np.zeros(nc))   # channels closed: no Ca current
    pkj.set(orc.ION0 + G["m_CaT"], np.zeros(nc))
    pkj.update_ion(np.full(nc, -60.0), dt)
    z = 10.0 * dt                                  # B_Ca = 10 / ms
    taylor = 1.0 - z + z * z / 2.0 - z ** 3 / 6.0 + z ** 4 / 24.0
    np.testing.assert_allclose(pkj.get(orc.CA), CA0 + (3e-4 - CA0) * taylor, rtol=1e-12)


def test_calcium_current_direction_and_clamp(pkj):
    nc, dt = pkj.nc, 2.0 ** -6
    fresh_state(pkj, -60.0, 3e-4, seed=4)
    pkj.update_ion(np.full(nc, -60.0), dt)
    open_ = pkj.get(orc.CA)
    z = 10.0 * dt
    closed = CA0 + (3e-4 - CA0) * (1.0 - z + z * z / 2.0 - z ** 3 / 6.0 + z ** 4 / 24.0)
    has_ca = pkj.get(orc.COND0 + 3) + pkj.get(orc.COND0 + 4) > 0
    assert np.all(open_[has_ca] > closed)          # inward current below the reversal potential
    fresh_state(pkj, -60.0, 3e-4, seed=4)
    pkj.set(orc.REV, np.full(nc, -100.0))          # driving force outward: clamped to no current
    pkj.update_ion(np.full(nc, -60.0), dt)
    np.testing.assert_allclose(pkj.get(orc.CA), np.full(nc, closed), rtol=1e-12)


def test_reversal_is_nernst(pkj):
    nc = pkj.nc
    fresh_state(pkj, -60.0, 2.5e-4, seed=5)
    pkj.update_ion(np.full(nc, -60.0), 2.0 ** -6)
    ca_mid = 0.5 * (2.5e-4 + pkj.get(orc.CA))
    rt_2f = 8.314462618 * 310.15 / (2.0 * 96485.33212) * 1e3    # mV at 37 C
    rev = pkj.get(orc.REV)
    idx = np.arange(nc)
    fixed = idx % 1599 == 0          # the reference's index-keyed somatic constant
    np.testing.assert_allclose(rev[~fixed], rt_2f * np.log(2.4 / ca_mid[~fixed]), rtol=2e-4)
    # the index-keyed compartments keep one constant, independent of their Ca
    assert np.all(rev[fixed] == rev[0]) and rev[0] == pytest.approx(12.5 * math.log(2.4 / CA0), rel=1e-15)


def read_geometry(path):
    rows = np.loadtxt(path)
    idx, parent = rows[:, 0].astype(int), rows[:, 1].astype(int)
    vec, d, typ = rows[:, 2:5], rows[:, 5], rows[:, 6].astype(int)
    r = 0.5 * d * 1e-4
    length = np.where(typ == 0, d * 1e-4, np.linalg.norm(vec, axis=1) * 1e-4)
    out = dict(parent=np.empty(len(idx), int), r=np.empty(len(idx)), len=np.empty(len(idx)),
               typ=np.empty(len(idx), int))
    out["parent"][idx] = parent
    out["r"][idx] = r
    out["len"][idx] = length
    out["typ"][idx] = typ
    return out


def test_cable_conductances_from_geometry(pkj, ref_data):
    geo = read_geometry(os.path.join(ref_data, "Pkj.txt"))
    row, col, val = pkj.csr()
    Ra = 0.25   # kOhm cm
    half = 0.5 * Ra * geo["len"] / (math.pi * geo["r"] ** 2)   # kOhm, half-segment axial resistance
    A = {}
    for i in range(pkj.nc):
        for k in range(row[i], row[i + 1]):
            if col[k] != i:
                A[(i, col[k])] = val[k]
    nedges = 0
    for i, p in enumerate(geo["parent"]):
        if p < 0:
            continue
        g = 1.0 / (half[i] + half[p])
        assert A[(i, p)] == pytest.approx(g, rel=1e-12)
        assert A[(p, i)] == A[(i, p)]
        nedges += 1
    assert len(A) == 2 * nedges
    # a uniform voltage carries no axial current
    b = pkj.get(orc.B).copy()
    pkj.set(orc.B, np.zeros(pkj.nc))
    assert np.max(np.abs(pkj.dydt(np.full(pkj.nc, -65.0)))) < 1e-9
    pkj.set(orc.B, b)


def test_capacitance_from_areas(pkj, ref_data):
    geo = read_geometry(os.path.join(ref_data, "Pkj.txt"))
    r, L, typ = geo["r"], geo["len"], geo["typ"]
    area = np.where(typ == 0, 4.0 * math.pi * r * r, 2.0 * math.pi * r * L)
    spines = np.where(typ == 3, L * 1.33e-8 * 13.0e4, 0.0)
    np.testing.assert_allclose(pkj.get(orc.CM), 1.64 * (area + spines), rtol=1e-12)
