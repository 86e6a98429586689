"""Pins of the oracle's inferior-olive model (oracle/or_io.c, restating io.cu,
io_solve.cu, io_ion.cu, gap.cu and the IO parts of solve_rkc.cu).

* The reference's own output file tests/golden/ref_data/io_raster_by_RKC.csv
  (an 8 x 8 grid, libc rand() seeded with 0 as main.cu:37 does, the
  io_set_current pulse at 800-820 ms) is reproduced exactly: all 64 spikes,
  same times, same order.  That run exercises the RKC core, the ion update,
  the frozen current, the gap junctions and the stimulus together.
* The pieces are pinned separately by what the mathematics fixes: the CN
  gate update is the (1,1) Pade form of exp(-dt/tau) at constant voltage, the
  Ca RK4 with no Ca current is the Taylor-4 decay, the coupling matrix splits
  one internal conductance by the compartments' area fractions (Schweighofer,
  Doya & Kawato 1999: g_int / p and g_int / (1 - p)), the gap current is odd
  in the voltage difference with conductance G_C at zero and 0.2 G_C far from
  it, and the grid neighbour relation is symmetric on square grids.
"""
import ctypes
import math
import os

import numpy as np
import pytest

import orc

IO_GATES = ["k_CaL", "l_CaL", "m_Na", "h_Na", "n_Kdr", "p_Kdr", "x_K", "r_CaH", "s_KCa", "q_H"]
G = {n: i for i, n in enumerate(IO_GATES)}


def L():
    lib = orc.lib()
    if not getattr(lib, "_io_bound", False):
        dp = ctypes.POINTER(ctypes.c_double)
        lib.or_io_create.restype = ctypes.c_void_p
        lib.or_io_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int]
        lib.or_io_rate.restype = ctypes.c_double
        lib.or_io_rate.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int]
        lib.or_io_update_ion.argtypes = [ctypes.c_void_p, dp, ctypes.c_double]
        lib.or_io_neighbours.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        lib._io_bound = True
    return lib


def io_pop(ref_data, nx, ny, v0=None, seed=0):
    lib = L()
    lib.srand(seed)
    vp = None
    if v0 is not None:
        v0 = np.ascontiguousarray(v0, dtype=np.float64)
        vp = orc._dp(v0)
    h = lib.or_io_create(os.path.join(ref_data, "IO.txt").encode(), nx, ny, vp, 1e-6, 1e-6, 0 if v0 is not None else 1)
    assert h
    return orc.Pop(h, nx * ny, 3)


def read_raster(path):
    out = []
    for line in open(path):
        t, c = line.strip().split(",")
        out.append((float(t), int(c)))
    return out


def test_reproduces_reference_io_raster(ref_data):
    p = io_pop(ref_data, 8, 8)
    prev = np.full(64, 100.0)
    spikes = []
    for k in range(6480):          # to 810 ms, past the 800-820 ms pulse
        t = k * 0.125
        p.set_current(t)
        p.advance()
        v = p.get(orc.VNEW)[2::3]  # io_output_file: compartment 2, upward through +10 mV
        for c in np.nonzero((v > 10.0) & (prev <= 10.0))[0]:
            spikes.append((t + 0.125, int(c)))
        prev = v
    assert spikes == read_raster(os.path.join(ref_data, "io_raster_by_RKC.csv"))


def rate(g, v, comp=1):
    return L().or_io_rate(2 * G[g], v, comp), L().or_io_rate(2 * G[g] + 1, v, comp)


def test_rates_bounded():
    for g in IO_GATES:
        if g == "m_Na":
            continue
        # s_KCa is a function of Ca (>= 0), the others of V
        xs = np.geomspace(1e-3, 1e4, 50) if g == "s_KCa" else np.linspace(-100.0, 50.0, 151) + 0.013
        for v in xs:   # V off the reference's 0/0 points
            for comp in (0, 1, 2):
                x, t = rate(g, v, comp)
                assert 0.0 <= x <= 1.0 and t > 0.0, (g, v, comp, x, t)
    # the dendrite has no Na inactivation kinetics: h_inf = tau_h = 1
    assert rate("h_Na", -50.0, 0) == (1.0, 1.0)


def pade11(z):
    return (1.0 + 0.5 * z) / (1.0 - 0.5 * z)


@pytest.mark.parametrize("dt", [2.0 ** -6, 0.05])
def test_cn_update_and_calcium_decay(ref_data, dt):
    p = io_pop(ref_data, 2, 2, v0=[-55.0] * 4)
    rng = np.random.default_rng(7)
    x = {g: rng.uniform(0.05, 0.95, p.nc) for g in IO_GATES}
    for g in IO_GATES:
        p.set(orc.ION0 + G[g], x[g])
    p.set(orc.ION0 + G["r_CaH"], np.zeros(p.nc))      # no Ca current
    x["r_CaH"] = np.zeros(p.nc)
    ca0 = rng.uniform(0.5, 2.0, p.nc)
    p.set(orc.CA, ca0)
    v = np.full(p.nc, -55.0)
    L().or_io_update_ion(p.h, orc._dp(v), dt)
    z = 0.075 * dt
    ca1 = ca0 * (1.0 - z + z * z / 2 - z ** 3 / 6 + z ** 4 / 24)
    np.testing.assert_allclose(p.get(orc.CA), ca1, rtol=1e-13)
    comp = np.tile([0, 1, 2], 4)
    for g in IO_GATES:
        new = p.get(orc.ION0 + G[g])
        if g == "m_Na":
            np.testing.assert_allclose(new, rate(g, -55.0)[0], rtol=1e-14)
            continue
        for i in range(p.nc):
            arg = 0.5 * (ca0[i] + ca1[i]) if g == "s_KCa" else -55.0
            xi, ti = rate(g, arg, int(comp[i]))
            assert new[i] == pytest.approx(xi + pade11(-dt / ti) * (x[g][i] - xi), rel=1e-12, abs=1e-15), (g, i)


def test_coupling_splits_internal_conductance_by_area_fraction(ref_data):
    p = io_pop(ref_data, 1, 1, v0=[-60.0])
    row, col, val = p.csr()
    g = {}
    for i in range(3):
        for k in range(row[i], row[i + 1]):
            if col[k] != i:
                g[(i, col[k])] = val[k]
    g_int, p_ds, p_sa = 0.13, 0.2, 0.869
    assert g[(0, 1)] * (1.0 - p_ds) == pytest.approx(g_int, rel=1e-15)   # dendrite <- soma
    assert g[(1, 0)] * p_ds == pytest.approx(g_int, rel=1e-15)           # soma <- dendrite
    assert g[(1, 2)] * p_sa == pytest.approx(g_int, rel=1e-15)           # soma <- axon
    assert g[(2, 1)] * (1.0 - p_sa) == pytest.approx(g_int, rel=1e-15)   # axon <- soma
    assert sorted(g) == [(0, 1), (1, 0), (1, 2), (2, 1)]


def gap(diff):
    # the oracle's gap current, read back through a 2-cell population's b
    return 0.04 * (0.8 * math.exp(-0.01 * diff * diff) + 0.2) * diff


def test_gap_current_shape(ref_data):
    # drive two coupled cells to known dendrite voltages and read the current
    # the update adds to each dendrite's b
    p = io_pop(ref_data, 2, 1, v0=[-60.0, -60.0])
    for d in (0.0, 0.5, 3.0, 10.0, 40.0):
        vnew = np.array([-60.0 + d, -60.0, -60.0, -60.0, -60.0, -60.0])
        base = io_pop(ref_data, 2, 1, v0=[-60.0, -60.0])
        p.set(orc.V, np.full(6, -60.0))
        base.set(orc.V, np.full(6, -60.0))
        for q in (p, base):
            for g in IO_GATES:
                q.set(orc.ION0 + G[g], np.full(6, 0.3))
            q.set(orc.CA, np.full(6, 1.0))
        L().or_io_update_ion(p.h, orc._dp(vnew), 0.01)
        orc.lib().or_io_update_b.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        orc.lib().or_io_update_b(p.h, orc._dp(vnew), 1)
        b_gap = p.get(orc.B)[0]
        orc.lib().or_io_update_b(p.h, orc._dp(vnew), 0)
        b_no = p.get(orc.B)[0]
        i = b_gap - b_no
        assert i == pytest.approx(gap(d), rel=1e-12, abs=1e-15)
        if d > 0:
            assert 0.2 * 0.04 * d <= i <= 0.04 * d
    assert abs(gap(40.0) / (0.2 * 0.04 * 40.0) - 1.0) < 1e-6   # 0.2 G_C far from zero


@pytest.mark.parametrize("n", [2, 3, 8])
def test_square_grid_neighbours_symmetric(n):
    nb = {}
    out = (ctypes.c_int * 4)()
    for c in range(n * n):
        L().or_io_neighbours(c, n, n, out)
        nb[c] = [x for x in out if x >= 0]
    for c, lst in nb.items():
        for d in lst:
            assert c in nb[d]
    counts = sorted(len(v) for v in nb.values())
    assert counts.count(2) == 4 and max(counts) == (4 if n > 2 else 2)
