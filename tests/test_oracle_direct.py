"""Pins of the oracle's exact CN / BE solve (or_cn_set_direct: tree elimination
in leaf order in place of the reference's truncated CG / BiCG).

* The solve against numpy's dense solve of the same system, on every cell
  type's own CN matrix (the Purkinje tree of Pkj.txt, the granule and Golgi
  trees, the inferior olive's unsymmetric three-compartment cells) and on
  random forests: a wrong elimination factor, a transposed off-diagonal or a
  wrong back-substitution neighbour fails at the 1e-11 bar.
* Graphs with a cycle (a tree plus one edge, Purkinje cells joined by gap
  junctions, a Golgi grid) are refused: elimination in leaf order would leave
  them unsolved.
* A passive cable's CN step with the exact solve equals the dense solve of
  (C/DT + L/2) v' = (C/DT - L/2) v + I to 1e-13, and where the reference's
  CG / BiCG converge (Golgi cell, olive grid) the two paths' trajectories
  agree to the CG tolerance; on the Purkinje tree, where the 100 CG
  iterations stop short, the exact solve's residual is at rounding level and
  the truncated iterate's is not.
"""
import ctypes
import os

import numpy as np
import pytest

import orc
from test_oracle_cn import L as LC, cn_matrix, generic, tree_laplacian
from test_oracle_io import io_pop
from test_oracle_go import go_pop
from test_oracle_gr import gr_pop


def L():
    lib = LC()
    if not getattr(lib, "_direct_bound", False):
        dp = ctypes.POINTER(ctypes.c_double)
        lib.or_cn_set_direct.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.or_cn_set_direct.restype = ctypes.c_int
        lib.or_cn_direct_test.argtypes = [ctypes.c_void_p, dp, dp, dp]
        lib.or_cn_direct_test.restype = ctypes.c_int
        lib.or_be_init.argtypes = [ctypes.c_void_p]
        lib.or_be_advance.argtypes = [ctypes.c_void_p, ctypes.c_double]
        lib._direct_bound = True
    return lib


def direct(p, diag, b):
    x = np.zeros(len(b))
    rc = L().or_cn_direct_test(p.h, orc._dp(np.ascontiguousarray(diag)), orc._dp(np.ascontiguousarray(b)),
                               orc._dp(x))
    return x, rc


def populations(ref_data):
    yield "pkj", orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=[-65.0], libc_rand=False)
    yield "gr", gr_pop(ref_data, 1, 2, v0=[-70.0, -69.0])
    yield "go", go_pop(ref_data, 1, 1, v0=[-60.0])
    yield "io", io_pop(ref_data, 2, 2, v0=[-60.0, -61.0, -62.0, -63.0])


def test_direct_solve_against_dense_on_every_cell_type(ref_data):
    rng = np.random.default_rng(3)
    for name, p in populations(ref_data):
        assert L().or_cn_init(p.h) == 0, name
        A0 = cn_matrix(p, np.zeros(p.nc))
        off = np.abs(A0).sum(axis=1)
        diag = off + rng.uniform(1e-6, 1e-4, p.nc)          # the CN diagonal's size: sum |g| + C/DT + ...
        b = rng.normal(size=p.nc)
        x, rc = direct(p, diag, b)
        assert rc == 0, name
        A = cn_matrix(p, diag)
        ref = np.linalg.solve(A, b)
        np.testing.assert_allclose(x, ref, rtol=1e-11, atol=1e-11 * np.max(np.abs(ref)), err_msg=name)
        if name == "io":
            assert not np.allclose(A, A.T)                     # the unsymmetric case is exercised


def test_direct_solve_on_random_forests():
    rng = np.random.default_rng(4)
    for trial in range(20):
        sizes = rng.integers(1, 40, size=rng.integers(1, 6))
        n = int(sizes.sum())
        Lr = np.zeros((n, n))
        o = 0
        for s in sizes:
            Lr[o:o + s, o:o + s] = tree_laplacian(int(s), rng)
            o += s
        perm = rng.permutation(n)                           # trees interleaved, no index order to rely on
        Lr = Lr[np.ix_(perm, perm)]
        p = generic(Lr, np.full(n, 1e-6), np.zeros(n))
        assert L().or_cn_init(p.h) == 0
        A0 = cn_matrix(p, np.zeros(n))
        diag = np.abs(A0).sum(axis=1) + rng.uniform(1e-5, 1e-3, n)
        A = cn_matrix(p, diag)
        A[np.triu_indices(n, 1)] *= rng.uniform(0.5, 1.5)   # unsymmetric values, same structure
        b = rng.normal(size=n)
        # the oracle reads the matrix from the population: rebuild it with the unsymmetric values
        Lu = -2.0 * A.copy()
        np.fill_diagonal(Lu, 0.0)
        np.fill_diagonal(Lu, -Lu.sum(axis=1))
        q = generic(Lu, np.full(n, 1e-6), np.zeros(n))
        assert L().or_cn_init(q.h) == 0
        x, rc = direct(q, diag, b)
        assert rc == 0
        np.testing.assert_allclose(x, np.linalg.solve(cn_matrix(q, diag), b), rtol=1e-11, atol=1e-12)


def test_cycles_are_refused(ref_data):
    rng = np.random.default_rng(5)
    n = 10
    Lr = tree_laplacian(n, rng)
    i, j = 2, 9
    if Lr[i, j] == 0.0:
        Lr[i, j] = Lr[j, i] = 1e-3
        np.fill_diagonal(Lr, 0.0)
        np.fill_diagonal(Lr, -Lr.sum(axis=1))
    p = generic(Lr, np.full(n, 1e-6), np.zeros(n))
    assert L().or_cn_init(p.h) == 0
    assert L().or_cn_set_direct(p.h, 1) == 1
    pk = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 2, 1, v0=[-65.0, -64.0], libc_rand=False,
                     gj=os.path.join(ref_data, "Pkj_GJ.csv"))
    assert L().or_cn_init(pk.h) == 0
    assert L().or_cn_set_direct(pk.h, 1) == 1
    go = go_pop(ref_data, 2, 2, v0=[-60.0] * 4)
    assert L().or_cn_init(go.h) == 0
    assert L().or_cn_set_direct(go.h, 1) == 1
    # a forest is accepted
    q = generic(tree_laplacian(n, rng), np.full(n, 1e-6), np.zeros(n))
    assert L().or_cn_init(q.h) == 0
    assert L().or_cn_set_direct(q.h, 1) == 0


def test_passive_cable_step_with_the_exact_solve():
    rng = np.random.default_rng(6)
    n = 30
    Lr = tree_laplacian(n, rng)
    cm = rng.uniform(0.5, 2.0, n) * 1e-6
    v0 = rng.uniform(-70.0, -60.0, n)
    p = generic(Lr, cm, v0)
    iext = np.zeros(n)
    iext[5] = 2e-5
    p.set(orc.IEXT, iext)
    assert L().or_cn_init(p.h) == 0
    assert L().or_cn_set_direct(p.h, 1) == 0
    dt = 0.025
    Lpos = -Lr
    A = np.diag(cm / dt) + Lpos / 2
    v = v0.copy()
    for _ in range(3):
        L().or_cn_step(p.h)
        v = np.linalg.solve(A, (np.diag(cm / dt) - Lpos / 2) @ v + iext)
        np.testing.assert_allclose(p.get(orc.V), v, rtol=1e-13)


@pytest.mark.parametrize("kind", ["go", "io"])
def test_exact_solve_equals_converged_krylov(ref_data, kind):
    """Golgi cell (CG) and 3 x 3 olive grid (BiCG): the reference's Krylov
    solves converge there, so both paths give the same trajectory."""
    mk = (lambda: go_pop(ref_data, 1, 1, v0=[-60.0])) if kind == "go" else \
        (lambda: io_pop(ref_data, 3, 3, v0=np.linspace(-62.0, -58.0, 9)))
    a, b = mk(), mk()
    for p in (a, b):
        assert L().or_cn_init(p.h) == 0
    assert L().or_cn_set_direct(b.h, 1) == 0
    for k in range(16):
        L().or_cn_advance(a.h, k * 0.125)
        L().or_cn_advance(b.h, k * 0.125)
    assert a.stats()["steps"] == 80
    assert a.stats()["stages"] < 80 * 100              # the Krylov solves converged before the cap
    np.testing.assert_allclose(b.get(orc.V), a.get(orc.V), rtol=0, atol=1e-9)


def test_exact_solve_residual_on_the_purkinje_tree(ref_data):
    p = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, v0=[-65.0], libc_rand=False)
    assert L().or_cn_init(p.h) == 0
    rng = np.random.default_rng(7)
    A0 = cn_matrix(p, np.zeros(p.nc))
    cm = p.get(orc.CM)
    diag = np.abs(A0).sum(axis=1) + cm / 0.025
    b = rng.normal(size=p.nc)
    x, rc = direct(p, diag, b)
    assert rc == 0
    A = cn_matrix(p, diag)
    assert np.linalg.norm(A @ x - b) <= 1e-13 * np.linalg.norm(b)
    from test_oracle_cn import solve as krylov
    y, it = krylov(p, diag, b, True)
    assert it == 100
    assert np.linalg.norm(A @ y - b) > 1e3 * np.linalg.norm(A @ x - b)


def test_be_path_with_the_exact_solve(ref_data):
    """BE: the exact solve on the granule pair gives the dense solution of the
    BE system the oracle assembles (pinned separately), step by step."""
    p = gr_pop(ref_data, 1, 2, v0=[-70.0, -69.0])
    assert L().or_be_init(p.h) == 0
    assert L().or_cn_set_direct(p.h, 1) == 0
    L().or_be_advance(p.h, 0.0)
    assert np.all(np.isfinite(p.get(orc.V)))
    assert p.stats()["stages"] == 0                    # no Krylov iterations


def test_step_setting_and_a_finer_passive_step():
    """or_cn_set_dt: a whole fraction of the 0.125 ms interval is accepted
    (the sub-step loop then runs 0.125 / dt steps), anything else refused; the
    CN step at dt = 2^-7 ms is the dense solve with that dt."""
    L().or_cn_set_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
    L().or_cn_set_dt.restype = ctypes.c_int
    rng = np.random.default_rng(8)
    n = 20
    Lr = tree_laplacian(n, rng)
    cm = rng.uniform(0.5, 2.0, n) * 1e-6
    v0 = rng.uniform(-70.0, -60.0, n)
    p = generic(Lr, cm, v0)
    iext = np.zeros(n)
    iext[2] = 2e-5
    p.set(orc.IEXT, iext)
    assert L().or_cn_init(p.h) == 0
    assert L().or_cn_set_direct(p.h, 1) == 0
    for bad in (0.03, 0.2, 0.125 / 3.5):
        assert L().or_cn_set_dt(p.h, bad) == 1
    dt = 2.0 ** -7
    assert L().or_cn_set_dt(p.h, dt) == 0
    Lpos = -Lr
    A = np.diag(cm / dt) + Lpos / 2
    L().or_cn_advance(p.h, 0.0)                        # one output interval = 16 steps of 2^-7 ms
    v = v0.copy()
    for _ in range(16):
        v = np.linalg.solve(A, (np.diag(cm / dt) - Lpos / 2) @ v + iext)
    np.testing.assert_allclose(p.get(orc.V), v, rtol=1e-12)
    assert p.stats()["steps"] == 16


def test_a_set_step_runs_a_whole_number_of_steps():
    """dt = 0.125 / 3 ms (not a binary fraction): each output interval runs
    exactly three steps at t, t + dt, t + 2 dt, where accumulating dt until
    0.125 is reached could run a fourth."""
    L().or_cn_set_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
    L().or_cn_set_dt.restype = ctypes.c_int
    rng = np.random.default_rng(9)
    n = 6
    p = generic(tree_laplacian(n, rng), np.full(n, 1e-6), np.full(n, -65.0))
    assert L().or_cn_init(p.h) == 0
    assert L().or_cn_set_dt(p.h, 0.125 / 3) == 0
    for k in range(4):
        L().or_cn_advance(p.h, k * 0.125)
    assert p.stats()["steps"] == 12
