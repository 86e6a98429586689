"""Pins of the oracle's Crank-Nicolson path (oracle/or_cn.c, restating
solve_cnm.cu and the CN parts of the cell models).

* One CN step of a passive cable (no ion channels, a constant injected
  current) solves (C/DT + L/2) v' = (C/DT - L/2) v + I: the oracle's step is
  compared with numpy's dense solve of that system, built here from the
  conductances (a wrong half, sign or missing capacitance term fails).
* The Krylov solvers against library routines: CG against numpy's dense solve
  on a small SPD system and against scipy.sparse.linalg.cg run from x = 0 for
  exactly 100 iterations on an ill-conditioned one (the truncation the
  reference's Purkinje runs live in); BiCG against scipy's bicg on an
  unsymmetric system, both converged and truncated.
* Every cell type's CN conductance and reversal sums against its own frozen
  current (pinned for the RKC path): Sum_k g_k (V - E_k) - I_ext must equal
  gs V - es at any V, which a dropped or doubled term breaks.
* The CN ion updates: with the voltage held constant they equal the RKC
  ones (same state, same DT), so the shared cores are exercised with the CN
  path's voltage arguments.
"""
import ctypes
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import orc

GAMMA1, OMEGA1, VOLD = 111, 112, 110


def L():
    lib = orc.lib()
    if not getattr(lib, "_cn_bound", False):
        dp = ctypes.POINTER(ctypes.c_double)
        lib.or_cn_init.argtypes = [ctypes.c_void_p]
        lib.or_cn_init.restype = ctypes.c_int
        lib.or_cn_step.argtypes = [ctypes.c_void_p]
        lib.or_cn_advance.argtypes = [ctypes.c_void_p, ctypes.c_double]
        lib.or_cn_solve_test.argtypes = [ctypes.c_void_p, dp, dp, dp, ctypes.c_int]
        lib.or_cn_solve_test.restype = ctypes.c_int
        lib._cn_bound = True
    return lib


def csr_with_diag(M):
    """CSR of a dense matrix keeping every diagonal entry (zero or not)."""
    n = M.shape[0]
    row, col, val = [0], [], []
    for i in range(n):
        for j in range(n):
            if M[i, j] != 0.0 or i == j:
                col.append(j)
                val.append(M[i, j])
        row.append(len(col))
    return (np.array(row, dtype=np.int32), np.array(col, dtype=np.int32), np.array(val, dtype=np.float64))


def generic(Lrkc, cm, v0):
    """A model-free population from an RKC-signed conductance matrix (off-diagonal
    +g, diagonal -sum g)."""
    lib = orc.lib()
    row, col, val = csr_with_diag(Lrkc)
    n = len(cm)
    cm = np.ascontiguousarray(cm, dtype=np.float64)
    v0 = np.ascontiguousarray(v0, dtype=np.float64)
    b = np.zeros(n)
    ip = ctypes.POINTER(ctypes.c_int)
    h = lib.or_generic_create(n, row.ctypes.data_as(ip), col.ctypes.data_as(ip), orc._dp(val), orc._dp(cm),
                              orc._dp(b), orc._dp(v0), 1e-6, 1e-6)
    assert h
    return orc.Pop(h, 1, n)


def tree_laplacian(n, rng):
    """RKC-signed conductance matrix of a random tree (parent of i is below i)."""
    Lm = np.zeros((n, n))
    for i in range(1, n):
        p = int(rng.integers(0, i))
        g = rng.uniform(0.5, 2.0) * 1e-3
        Lm[i, p] = Lm[p, i] = g
    np.fill_diagonal(Lm, -Lm.sum(axis=1))
    return Lm


def test_passive_cable_step_is_the_cn_system():
    rng = np.random.default_rng(1)
    n = 9
    Lr = tree_laplacian(n, rng)
    cm = rng.uniform(0.5, 2.0, n) * 1e-6
    v0 = rng.uniform(-70.0, -60.0, n)
    p = generic(Lr, cm, v0)
    iext = np.zeros(n)
    iext[3] = 2e-5
    p.set(orc.IEXT, iext)
    assert L().or_cn_init(p.h) == 0
    dt = 0.025
    Lpos = -Lr                                          # conductance Laplacian, positive diagonal
    A = np.diag(cm / dt) + Lpos / 2
    rhs = (np.diag(cm / dt) - Lpos / 2) @ v0 + iext
    L().or_cn_step(p.h)
    np.testing.assert_allclose(p.get(orc.V), np.linalg.solve(A, rhs), rtol=1e-12)
    np.testing.assert_array_equal(p.get(VOLD), v0)
    # a second step carries the half sums forward (omega1 = iext / 2)
    v1 = p.get(orc.V)
    L().or_cn_step(p.h)
    rhs2 = (np.diag(cm / dt) - Lpos / 2) @ v1 + iext
    np.testing.assert_allclose(p.get(orc.V), np.linalg.solve(A, rhs2), rtol=1e-12)


def solve(p, diag, b, sym):
    x = np.zeros(len(b))
    it = L().or_cn_solve_test(p.h, orc._dp(np.ascontiguousarray(diag)), orc._dp(np.ascontiguousarray(b)),
                              orc._dp(x), 1 if sym else 0)
    return x, it


def cn_matrix(p, diag):
    """The CN path's system: the population's halved cable matrix with the
    given diagonal."""
    row, col, val = p.csr()
    n = p.nc
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(row[i], row[i + 1]):
            A[i, col[j]] = diag[i] if col[j] == i else -val[j] / 2.0
    return A


def test_cg_against_dense_and_scipy():
    rng = np.random.default_rng(2)
    n = 12
    p = generic(tree_laplacian(n, rng), np.full(n, 1e-6), np.zeros(n))
    assert L().or_cn_init(p.h) == 0
    _, _, val = p.csr()
    A0 = cn_matrix(p, np.zeros(n))
    diag = -A0.sum(axis=1) + rng.uniform(1e-5, 1e-4, n)      # diagonally dominant, SPD
    A = cn_matrix(p, diag)
    b = rng.normal(size=n)
    x, it = solve(p, diag, b, True)
    assert it <= n + 2
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10)
    # ill-conditioned: a long chain, tiny capacitance -> 100 iterations, truncated
    m = 400
    Lc = np.zeros((m, m))
    for i in range(m - 1):
        Lc[i, i + 1] = Lc[i + 1, i] = 1e-3
    np.fill_diagonal(Lc, -Lc.sum(axis=1))
    q = generic(Lc, np.full(m, 1e-6), np.zeros(m))
    assert L().or_cn_init(q.h) == 0
    diag = 1e-3 + 1e-7 * np.ones(m)
    diag[0] = diag[-1] = 5e-4 + 1e-7
    Aq = cn_matrix(q, diag)
    bq = rng.normal(size=m)
    xq, itq = solve(q, diag, bq, True)
    assert itq == 100
    xs, _ = spla.cg(sp.csr_matrix(Aq), bq, x0=np.zeros(m), rtol=1e-300, atol=0.0, maxiter=100)
    assert np.linalg.norm(xq - np.linalg.solve(Aq, bq)) > 1e-3 * np.linalg.norm(np.linalg.solve(Aq, bq))
    np.testing.assert_allclose(xq, xs, rtol=1e-8, atol=1e-8 * np.max(np.abs(xs)))


def test_bicg_against_scipy():
    rng = np.random.default_rng(3)
    n = 15
    Lr = tree_laplacian(n, rng)
    # unsymmetric: scale each row's off-diagonals (IO's per-compartment split)
    s = rng.uniform(0.5, 2.0, n)
    Lu = Lr * s[:, None]
    np.fill_diagonal(Lu, 0.0)
    np.fill_diagonal(Lu, -Lu.sum(axis=1))
    p = generic(Lu, np.full(n, 1e-6), np.zeros(n))
    assert L().or_cn_init(p.h) == 0
    diag = -cn_matrix(p, np.zeros(n)).sum(axis=1) + rng.uniform(1e-4, 1e-3, n)
    A = cn_matrix(p, diag)
    assert not np.allclose(A, A.T)
    b = rng.normal(size=n)
    x, it = solve(p, diag, b, False)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9)
    xs, info = spla.bicg(sp.csr_matrix(A), b, x0=np.zeros(n), rtol=1e-14, atol=0.0, maxiter=200)
    assert info == 0
    np.testing.assert_allclose(x, xs, rtol=1e-9)
    assert 1 < it <= 100


@pytest.mark.parametrize("kind", ["pkj", "io", "go", "gr"])
def test_cn_sums_match_the_frozen_current(ref_data, kind):
    if kind == "pkj":
        p = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=[-60.0], libc_rand=False)
        upd = orc.lib().or_pkj_update_b
    elif kind == "io":
        from test_oracle_io import io_pop
        p = io_pop(ref_data, 2, 1, v0=[-60.0, -62.0])
        upd = lambda h, vp: orc.lib().or_io_update_b(h, vp, 0)
        orc.lib().or_io_update_b.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    elif kind == "go":
        from test_oracle_go import go_pop, L as LG
        p = go_pop(ref_data, 1, 1, v0=[-60.0])
        upd = LG().or_go_update_b
    else:
        from test_oracle_gr import gr_pop, L as LR
        p = gr_pop(ref_data, 1, 1, v0=[-70.0])
        upd = LR().or_gr_update_b
    rng = np.random.default_rng(4)
    iext = np.zeros(p.nc)
    iext[::7] = rng.uniform(-1e-4, 1e-4, len(iext[::7]))
    p.set(orc.IEXT, iext)
    assert L().or_cn_init(p.h) == 0
    gs = 2.0 * p.get(GAMMA1)
    es = 2.0 * p.get(OMEGA1)
    for v in (-80.0, -55.0, 20.0):
        vv = np.full(p.nc, v)
        upd(p.h, orc._dp(vv))
        b = p.get(orc.B)
        scale = np.maximum(np.abs(gs * v), np.abs(es)) + np.abs(b) + 1e-30
        assert np.max(np.abs(b - (gs * v - es)) / scale) < 1e-12, kind


@pytest.mark.parametrize("kind", ["pkj", "io", "go", "gr"])
def test_cn_ion_update_at_constant_voltage_equals_rkc_update(ref_data, kind):
    """At V = V_old = V_new both paths move Ca and the gates identically."""
    v0 = -58.0
    pops = []
    for _ in range(2):
        if kind == "pkj":
            q = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=[v0], libc_rand=False)
        elif kind == "io":
            from test_oracle_io import io_pop
            q = io_pop(ref_data, 1, 1, v0=[v0])
        elif kind == "go":
            from test_oracle_go import go_pop
            q = go_pop(ref_data, 1, 1, v0=[v0])
        else:
            from test_oracle_gr import gr_pop
            q = gr_pop(ref_data, 1, 1, v0=[v0])
        pops.append(q)
    a, b = pops
    assert L().or_cn_init(b.h) == 0
    fn = {"pkj": orc.lib().or_pkj_cn_ion, "io": orc.lib().or_io_cn_ion, "go": orc.lib().or_go_cn_ion,
          "gr": orc.lib().or_gr_cn_ion}[kind]
    fn.argtypes = [ctypes.c_void_p, ctypes.c_double]
    upd = {"pkj": orc.lib().or_pkj_update_ion, "io": orc.lib().or_io_update_ion, "go": orc.lib().or_go_update_ion,
           "gr": orc.lib().or_gr_update_ion}[kind]
    upd.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_double]
    vv = np.full(a.nc, v0)
    upd(a.h, orc._dp(vv), 0.025)
    fn(b.h, 0.025)
    for g in range(18 if kind == "pkj" else 10 if kind == "io" else 24 if kind == "go" else 20):
        np.testing.assert_array_equal(a.get(orc.ION0 + g), b.get(orc.ION0 + g))
    np.testing.assert_array_equal(a.get(orc.CA), b.get(orc.CA))
