"""ctypes bindings to the CPU FP64 oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker, never the product.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")
REF_DATA = os.path.join(ROOT, "tests", "golden", "ref_data")

# field ids understood by or_pop_get / or_pop_set
Y, YN, FN, VT1, VT2, EDE, VNEW = range(7)
V, CA, REV, B, CM, IEXT, SHELL, AREA = 100, 101, 102, 103, 104, 105, 106, 107
ION0, COND0 = 200, 300

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_pkj_create.restype = ctypes.c_void_p
        L.or_pkj_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                    dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.or_rkc_advance.argtypes = [ctypes.c_void_p]
        L.or_pop_free.argtypes = [ctypes.c_void_p]
        L.or_pop_nc.argtypes = [ctypes.c_void_p]
        L.or_pop_nnz.argtypes = [ctypes.c_void_p]
        L.or_pop_get.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.or_pop_set.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.or_pop_controller.argtypes = [ctypes.c_void_p, dp, dp, ctypes.POINTER(ctypes.c_int), dp]
        L.or_pop_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        L.or_pop_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_pop_csr.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), dp]
        L.or_pop_enable_syn.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.or_set_stim.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp]
        L.or_set_current.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.or_det_sum.restype = ctypes.c_double
        L.or_det_sum.argtypes = [dp, ctypes.c_long]
        L.or_pkj_rate.restype = ctypes.c_double
        L.or_pkj_rate.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.or_dydt.argtypes = [ctypes.c_void_p, dp, dp]
        L.or_pop_set_integ.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.or_pop_set_gid_offset.argtypes = [ctypes.c_void_p, ctypes.c_long]
        L.or_pop_set_refresh_fn.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_pop_set_rho_floor.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.or_pop_poisson_step.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double]
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.or_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.srand.argtypes = [ctypes.c_uint]
        ip = ctypes.POINTER(ctypes.c_int)
        L.or_generic_create.restype = ctypes.c_void_p
        L.or_generic_create.argtypes = [ctypes.c_int, ip, ip, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double]
        L.or_rkc_step_test.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        L.or_rkclow_test.argtypes = [ctypes.c_void_p]
        L.or_rkcrho_test.argtypes = [ctypes.c_void_p]
        L.or_rkcint_test.argtypes = [ctypes.c_void_p, ctypes.c_double, dp]
        L.or_pop_set_other.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.or_pop_set_hwork.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.or_pkj_update_ion.argtypes = [ctypes.c_void_p, dp, ctypes.c_double]
        L.or_pkj_update_b.argtypes = [ctypes.c_void_p, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def det_sum(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().or_det_sum(_dp(x), x.size)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib().or_philox4x32_10(c.ctypes.data_as(u32p), k.ctypes.data_as(u32p), o.ctypes.data_as(u32p))
    return o


class Pop:
    """One oracle population (a reference neuron_t + neuron_solve_t pair)."""

    def __init__(self, handle, n, comp):
        self.h = handle
        self.n = n
        self.comp = comp
        self.nc = lib().or_pop_nc(handle)

    @classmethod
    def pkj(cls, morph, nx=1, ny=1, v0=None, atol=1e-6, rtol=1e-6, libc_rand=True, gj=None, seed=0):
        L = lib()
        if libc_rand:
            L.srand(seed)
        vp = None
        if v0 is not None:
            v0 = np.ascontiguousarray(v0, dtype=np.float64)
            vp = _dp(v0)
        h = L.or_pkj_create(morph.encode(), gj.encode() if gj else None, nx, ny, vp, atol, rtol,
                            1 if libc_rand else 0)
        if not h:
            raise RuntimeError("oracle pkj_create failed")
        return cls(h, nx * ny, 1600)

    @classmethod
    def generic(cls, L, cm, b, v0, atol=1e-6, rtol=1e-6):
        """A population given by its cable matrix alone (dense L in the RKC
        sign convention: off-diagonal +g), no ion model."""
        L = np.asarray(L, dtype=np.float64)
        nc = L.shape[0]
        row, col, val = [0], [], []
        for i in range(nc):
            for j in range(nc):
                if j != i and L[i, j] != 0.0:
                    col.append(j)
                    val.append(L[i, j])
            row.append(len(col))
        row = np.array(row, dtype=np.int32)
        col = np.array(col if col else [0], dtype=np.int32)
        val = np.array(val if val else [0.0])
        cm = np.ascontiguousarray(cm, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        v0 = np.ascontiguousarray(v0, dtype=np.float64)
        ip = ctypes.POINTER(ctypes.c_int)
        h = lib().or_generic_create(nc, row.ctypes.data_as(ip), col.ctypes.data_as(ip), _dp(val), _dp(cm), _dp(b),
                                    _dp(v0), atol, rtol)
        if not h:
            raise RuntimeError("oracle generic_create failed")
        return cls(h, 1, nc)

    def rkc_step(self, h, m):
        lib().or_rkc_step_test(self.h, h, m)

    def rkclow(self):
        lib().or_rkclow_test(self.h)

    def rkcrho(self):
        lib().or_rkcrho_test(self.h)
        return self.controller()[0][2]

    def rkcint(self, t):
        out = np.empty(self.nc)
        lib().or_rkcint_test(self.h, t, _dp(out))
        return out

    def set_other(self, i, v):
        lib().or_pop_set_other(self.h, i, v)

    def set_hwork(self, i, v):
        lib().or_pop_set_hwork(self.h, i, v)

    def update_ion(self, vnew, dt):
        vnew = np.ascontiguousarray(vnew, dtype=np.float64)
        lib().or_pkj_update_ion(self.h, _dp(vnew), dt)

    def update_b(self, vv):
        vv = np.ascontiguousarray(vv, dtype=np.float64)
        lib().or_pkj_update_b(self.h, _dp(vv))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_pop_free(self.h)
            self.h = None

    def get(self, what):
        a = np.empty(self.nc)
        lib().or_pop_get(self.h, what, _dp(a))
        return a

    def set(self, what, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        lib().or_pop_set(self.h, what, _dp(a))

    def set_stim(self, comp, t0, t1, amp):
        amp = np.broadcast_to(np.asarray(amp, dtype=np.float64), (self.n,)).copy()
        lib().or_set_stim(self.h, comp, t0, t1, _dp(amp))

    def set_current(self, t):
        lib().or_set_current(self.h, t)

    def controller(self):
        o = np.zeros(14)
        hw = np.zeros(8)
        fl = (ctypes.c_int * 3)()
        sd = np.zeros(1)
        lib().or_pop_controller(self.h, _dp(o), _dp(hw), fl, _dp(sd))
        return o, hw, list(fl), sd[0]

    def stats(self):
        s = (ctypes.c_long * 5)()
        lib().or_pop_stats(self.h, s)
        return dict(steps=s[0], rejects=s[1], stages=s[2], rho_iters=s[3], ion_updates=s[4])

    def csr(self):
        nnz = lib().or_pop_nnz(self.h)
        row = np.zeros(self.nc + 1, dtype=np.int32)
        col = np.zeros(nnz, dtype=np.int32)
        val = np.zeros(nnz)
        ip = ctypes.POINTER(ctypes.c_int)
        lib().or_pop_csr(self.h, row.ctypes.data_as(ip), col.ctypes.data_as(ip), _dp(val))
        return row, col, val

    def dydt(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        dy = np.empty_like(y)
        lib().or_dydt(self.h, _dp(y), _dp(dy))
        return dy

    def threads(self, n):
        lib().or_pop_set_threads(self.h, n)

    def gid_offset(self, off):
        lib().or_pop_set_gid_offset(self.h, off)

    def stable_split(self, ion_interval=2.0 ** -8, rho_floor=True):
        """the safeguards used for the synthetic Purkinje workloads: ion interval,
        fn refresh and the spectral-radius floor 1.2 x the exact cable radius
        (computed here with scipy's ARPACK, independently of the product)"""
        lib().or_pop_set_integ(self.h, ion_interval)
        lib().or_pop_set_refresh_fn(self.h, 1)
        if rho_floor:
            lib().or_pop_set_rho_floor(self.h, 1.2 * max(self.cable_rho()))

    def cable_rho(self):
        """per-cell spectral radius of C^-1 L from the oracle's own CSR (scipy eigsh)"""
        import scipy.sparse as sp
        import scipy.sparse.linalg as sla
        row, col, val = self.csr()
        cm = self.get(CM)
        out = []
        for c in range(self.n):
            s, e = c * self.comp, (c + 1) * self.comp
            L = sp.csr_matrix((val, col, row), shape=(self.nc, self.nc))[s:e, s:e]
            d = np.asarray(L.sum(axis=1)).ravel() - L.diagonal()
            L = L - sp.diags(L.diagonal()) - sp.diags(d)   # cell-local diagonal (drops intercell terms)
            S = sp.diags(cm[s:e] ** -0.5) @ L @ sp.diags(cm[s:e] ** -0.5)
            out.append(float(sla.eigsh(-S, k=1, which="LA", return_eigenvectors=False, tol=1e-12)[0]))
        return out

    def enable_poisson(self, rate_hz=5.0, g_mS=1e-6, tau_ms=2.0, e_mV=0.0, seed=0):
        lib().or_pop_enable_syn(self.h, e_mV)
        self._poisson = (seed & 0xffffffff, seed >> 32, np.exp(-0.125 / tau_ms),
                         1.0 - np.exp(-rate_hz * 0.125 * 1e-3), g_mS)
        self._step = 0

    def advance(self):
        if getattr(self, "_poisson", None):
            k0, k1, decay, prob, g = self._poisson
            lib().or_pop_poisson_step(self.h, self._step, k0, k1, decay, prob, g)
            self._step += 1
        r = lib().or_rkc_advance(self.h)
        if r != 0:
            raise RuntimeError(f"oracle RKC failed: idid={r}")


def split_morph(path, ncell, comp, outdir):
    """Writes one single-cell morphology file per cell block of `path`."""
    lines = [l for l in open(path) if l.strip()]
    assert len(lines) == ncell * comp
    out = []
    for c in range(ncell):
        p = os.path.join(outdir, f"cell{c}.txt")
        with open(p, "w") as f:
            f.writelines(lines[c * comp:(c + 1) * comp])
        out.append(p)
    return out
