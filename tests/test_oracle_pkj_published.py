"""The oracle's Purkinje channel model (oracle/or_pkj.c, restating
pkj_ion.cu:4-91 and the density table pkj.cuh:28-62) against the published
model it implements: De Schutter & Bower (1994), J. Neurophysiol. 71:375-400,
Appendix tables, in the form the authors distributed with the GENESIS
implementation (`setupalpha` parameters in SI units: rates in 1/s, voltages in
V, densities in S/m^2).  The table below is written in those units and
converted here, so it is not a retyping of the reference's ms / mV literals.

For a two-rate gate the oracle exposes x_inf and tau; alpha = x_inf / tau and
beta = (1 - x_inf) / tau recover the rates, which are compared with

    rate(V) = A / (B + exp((V + C) / D))        (GENESIS: (A + 0 V) / (B + exp((V + C) / D)))

at five voltages, at their defining points (V = -C gives A / (B + 1)) and, for
B = 0, through the log-slope d ln(rate) / dV = -1 / D.  The Ca-activated
gates are checked at their half-activation Ca and their fixed time constant,
the anomalous rectifier at its half-activation voltage and its two time
constants and fractions.  Two channels are not restated from the paper here:
the delayed rectifier (Kdr) and the muscarinic current (KM) rate forms; they
stay pinned by the closed-form properties of test_oracle_pkj.py only.

Densities: every channel conductance of every compartment of Pkj.txt divided
by its membrane area equals the published density of its compartment class
(soma, main dendrite, smooth / spiny dendrite), and the leak equals the area
(with the spines on spiny compartments: 13 spines per um of 1.33 um^2) over
the published membrane resistivity (10 / 30 kOhm cm^2).

The last test plants single-constant typos into a copy of the oracle and
shows that each one fails these checks.
"""
import ctypes
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import orc

GATES = ["m_NaF", "h_NaF", "m_NaP", "m_CaP", "h_CaP", "m_CaT", "h_CaT", "m_KA", "h_KA", "m_KC", "m_Kh1",
         "m_Kh2", "m_Kdr", "h_Kdr", "m_KM", "z_KC", "m_K2", "z_K2"]
G = {n: i for i, n in enumerate(GATES)}

# (A [1/s], B, C [V], D [V]) of alpha and beta; None: a constant rate, given
# alone in 1/s
PUBLISHED_SI = {
    "m_NaF": ((35.0e3, 0.0, 0.005, -0.010), (7.0e3, 0.0, 0.065, 0.020)),
    "h_NaF": ((0.225e3, 1.0, 0.080, 0.010), (7.5e3, 0.0, -0.003, -0.018)),
    "m_NaP": ((200.0e3, 1.0, -0.018, -0.016), (25.0e3, 1.0, 0.058, 0.008)),
    "m_CaP": ((8.5e3, 1.0, -0.008, -0.0125), (35.0e3, 1.0, 0.074, 0.0145)),
    "h_CaP": ((1.5, 1.0, 0.029, 0.008), (5.5, 1.0, 0.023, -0.008)),
    "m_CaT": ((2.6e3, 1.0, 0.021, -0.008), (0.18e3, 1.0, 0.040, 0.004)),
    "h_CaT": ((2.5, 1.0, 0.040, 0.008), (0.19e3, 1.0, 0.050, -0.010)),
    "m_KA": ((1.4e3, 1.0, 0.027, -0.012), (0.49e3, 1.0, 0.030, 0.004)),
    "h_KA": ((0.0175e3, 1.0, 0.050, 0.008), (1.3e3, 1.0, 0.013, -0.010)),
    "m_KC": (7.5e3, (0.11e3, 0.0, -0.035, 0.0149)),
    "m_K2": (25.0e3, (0.075e3, 0.0, 0.025, 0.006)),
}
# Ca-activated gates: z_inf = 1 / (1 + K / [Ca]) with K in mM, tau_z in s
PUBLISHED_Z = {"z_KC": (4.0, 0.010), "z_K2": (0.2, 0.010)}
# anomalous rectifier: m_inf = 1 / (1 + exp((V + C) / D)), two components
# holding fractions f with time constants tau [s]
PUBLISHED_KH = dict(C=0.082, D=0.007, parts={"m_Kh1": (0.8, 0.0076), "m_Kh2": (0.2, 0.0368)})

# densities [S/m^2] per class, and membrane resistivity [Ohm m^2]
CHANNELS = ["NaF", "NaP", "CaP", "CaT", "Kh", "Kdr", "KM", "KA", "KC", "K2"]   # conductance slots 1..10
PUBLISHED_DENSITY_SI = {
    "soma": dict(NaF=75000.0, NaP=10.0, CaT=5.0, Kh=3.0, Kdr=6000.0, KM=0.4, KA=150.0),
    "main": dict(CaP=45.0, CaT=5.0, Kdr=600.0, KM=0.1, KA=20.0, KC=800.0, K2=3.9),
    "dend": dict(CaP=45.0, CaT=5.0, KM=0.13, KC=800.0, K2=3.9),
}
RM_SI = {"soma": 1.0, "dend": 3.0}
SPINES_PER_M, SPINE_AREA_M2 = 13.0e6, 1.33e-12


def to_ms_mV(p):
    """A [1/s], C and D [V] -> A [1/ms], C and D [mV]"""
    A, B, C, D = p
    return A * 1e-3, B, C * 1e3, D * 1e3


def rate(p, v):
    A, B, C, D = to_ms_mV(p)
    return A / (B + math.exp((v + C) / D))


def oracle_rates(L, g, v, ca=1e-4):
    x = L.or_pkj_rate(2 * G[g], v, ca)
    t = L.or_pkj_rate(2 * G[g] + 1, v, ca)
    return x / t, (1.0 - x) / t


VOLTS = (-90.0, -62.5, -35.0, -7.5, 20.0)


def check_rates(L):
    for g, (pa, pb) in PUBLISHED_SI.items():
        for v in VOLTS:
            a, b = oracle_rates(L, g, v)
            want_a = pa * 1e-3 if not isinstance(pa, tuple) else rate(pa, v)
            want_b = rate(pb, v)
            assert a == pytest.approx(want_a, rel=1e-12), (g, "alpha", v, a, want_a)
            assert b == pytest.approx(want_b, rel=1e-12), (g, "beta", v, b, want_b)
        for k, p in enumerate((pa, pb)):
            if not isinstance(p, tuple):
                continue
            A, B, C, D = to_ms_mV(p)
            # defining point: the exponential is 1 at V = -C
            got = oracle_rates(L, g, -C)[k]
            assert got == pytest.approx(A / (B + 1.0), rel=1e-12), (g, k, "at V = -C")
            if B == 0.0:
                # a pure exponential: fixed log-slope -1/D
                r1, r2 = oracle_rates(L, g, -50.0)[k], oracle_rates(L, g, -40.0)[k]
                assert (math.log(r2) - math.log(r1)) / 10.0 == pytest.approx(-1.0 / D, rel=1e-9), (g, k, "slope")


def check_calcium_gates(L):
    for g, (K, tau_s) in PUBLISHED_Z.items():
        assert L.or_pkj_rate(2 * G[g], -60.0, K) == pytest.approx(0.5, rel=1e-14), g
        for ca in (1e-4, 0.1 * K, 10.0 * K):
            assert L.or_pkj_rate(2 * G[g], 0.0, ca) == pytest.approx(1.0 / (1.0 + K / ca), rel=1e-14), (g, ca)
        assert L.or_pkj_rate(2 * G[g] + 1, -30.0, 1e-3) == pytest.approx(tau_s * 1e3, rel=1e-14), g


def check_kh(L):
    C, D = PUBLISHED_KH["C"] * 1e3, PUBLISHED_KH["D"] * 1e3
    for g, (f, tau_s) in PUBLISHED_KH["parts"].items():
        assert L.or_pkj_rate(2 * G[g], -C, 1e-4) == pytest.approx(0.5 * f, rel=1e-14), g
        for v in VOLTS:
            assert L.or_pkj_rate(2 * G[g], v, 1e-4) == pytest.approx(f / (1.0 + math.exp((v + C) / D)), rel=1e-13)
            assert L.or_pkj_rate(2 * G[g] + 1, v, 1e-4) == pytest.approx(tau_s * 1e3, rel=1e-14), g


def pkj_types(path):
    return np.loadtxt(path, usecols=(6,), dtype=int)


def check_densities(L, morph):
    L.or_pkj_create.restype = ctypes.c_void_p
    L.or_pkj_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.c_double, ctypes.c_int]
    L.or_pop_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.or_pop_free.argtypes = [ctypes.c_void_p]
    v0 = np.array([-65.0])
    h = L.or_pkj_create(morph.encode(), None, 1, 1, v0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        1e-6, 1e-6, 0)
    assert h
    try:
        def get(what):
            a = np.empty(1600)
            L.or_pop_get(h, what, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            return a
        area = get(107)                     # cm^2, without spines
        types = pkj_types(morph)
        rows = np.loadtxt(morph)
        cls = np.where(types == 0, 0, np.where(types == 1, 1, 2))
        names = ["soma", "main", "dend"]
        for k, ch in enumerate(CHANNELS):
            cond = get(300 + 1 + k)         # mS
            for c in range(3):
                m = cls == c
                dens_mS_cm2 = PUBLISHED_DENSITY_SI[names[c]].get(ch, 0.0) * 0.1   # 1 S/m^2 = 0.1 mS/cm^2
                np.testing.assert_allclose(cond[m] / area[m], dens_mS_cm2, rtol=1e-13, atol=0.0,
                                           err_msg=f"{ch} density, {names[c]}")
        # leak: area (plus spines on spiny compartments) / Rm; Rm [Ohm m^2] -> [kOhm cm^2] is x 10
        leak = get(300)
        # Pkj.txt's x y z columns hold each segment's vector [um]
        length_cm = np.linalg.norm(rows[:, 2:5], axis=1) * 1e-4
        spine_cm2 = np.where(types == 3, length_cm * SPINES_PER_M * 1e-2 * SPINE_AREA_M2 * 1e4, 0.0)
        rm = np.where(types == 0, RM_SI["soma"], RM_SI["dend"]) * 10.0
        np.testing.assert_allclose(leak, (area + spine_cm2) / rm, rtol=1e-9, err_msg="leak")
    finally:
        L.or_pop_free(h)


def prepare(L):
    L.or_pkj_rate.restype = ctypes.c_double
    L.or_pkj_rate.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    return L


def test_rates_match_the_published_tables():
    check_rates(orc.lib())


def test_calcium_activated_gates_match_the_published_forms():
    check_calcium_gates(orc.lib())


def test_anomalous_rectifier_matches_the_published_form():
    check_kh(orc.lib())


def test_densities_and_leak_match_the_published_table(ref_data):
    check_densities(orc.lib(), os.path.join(ref_data, "Pkj.txt"))


ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
PLANTS = [
    # (label, text in or_pkj.c, replacement, check that must catch it)
    ("beta_h_CaP 0.0055 -> 0.055", "lf(v, 0.0055, 1.0, 23.0, -8.0)", "lf(v, 0.055, 1.0, 23.0, -8.0)", "rates"),
    ("alpha_m_NaF half-point sign", "lf(v, 35.0, 0.0, 5.0, -10.0)", "lf(v, 35.0, 0.0, -5.0, -10.0)", "rates"),
    ("beta_m_KA slope 4 -> 5", "lf(v, 0.490, 1.0, 30.0, 4.0)", "lf(v, 0.490, 1.0, 30.0, 5.0)", "rates"),
    ("z_K2 half-activation 0.20 -> 0.02", "1.0 / (1.0 + 0.20 / Ca)", "1.0 / (1.0 + 0.02 / Ca)", "calcium"),
    ("tau_m_Kh2 36.8 -> 38.6", "return 36.8;", "return 38.6;", "kh"),
    ("main-dendrite KC density 80 -> 8", "{0.0, 0.0, 4.5, 0.5, 0.0, 60.0, 0.01, 2.0, 80.0, 0.39}",
     "{0.0, 0.0, 4.5, 0.5, 0.0, 60.0, 0.01, 2.0, 8.0, 0.39}", "densities"),
]


@pytest.mark.parametrize("label,old,new,which", PLANTS, ids=[p[0] for p in PLANTS])
def test_a_planted_typo_fails(tmp_path, ref_data, label, old, new, which):
    """A copy of the oracle with one constant changed is caught by the check
    that covers that constant."""
    for f in os.listdir(ORACLE_DIR):
        if f.endswith((".c", ".h")):
            shutil.copy(os.path.join(ORACLE_DIR, f), tmp_path / f)
    src = (tmp_path / "or_pkj.c").read_text()
    assert src.count(old) == 1, label
    (tmp_path / "or_pkj.c").write_text(src.replace(old, new))
    srcs = sorted(str(p) for p in tmp_path.glob("*.c"))
    so = str(tmp_path / "liboracle_planted.so")
    subprocess.run(["gcc", "-O1", "-std=gnu99", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", "-o", so]
                   + srcs + ["-lm"], check=True, capture_output=True)
    L = prepare(ctypes.CDLL(so))
    checks = {"rates": lambda: check_rates(L), "calcium": lambda: check_calcium_gates(L), "kh": lambda: check_kh(L),
              "densities": lambda: check_densities(L, os.path.join(ref_data, "Pkj.txt"))}
    with pytest.raises(AssertionError):
        checks[which]()
