"""Python mirror of the reference's per-population interface over libcbrkc.so.

The reference drives each cell type through
    xx_initialize / xx_solve_initialize      (pkj.cu:240, pkj_solve.cu:286)
    xx_set_current                           (pkj.cu:333)
    xx_solve_update_v                        (pkj_solve.cu:395)
    xx_output_file                           (pkj.cu:370)
    xx_finalize / xx_solve_finalize          (pkj.cu:308, pkj_solve.cu:368)
This module exposes the same steps on a `Population` object (and as
reference-named functions for Purkinje cells).  Every call goes through the
C-ABI declared in include/cbrkc.h; there is no fallback: if the HIP library
is missing or no GPU is present, calls fail loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcbrkc.so")

CB_OK = 0
GR, GO, PKJ, IO = 0, 1, 2, 3
BE, CN, RKC = 0, 1, 2
ENGINE_GLOBAL_FP64, ENGINE_CELL_FP64, ENGINE_CELL_MIXED, ENGINE_TREE_FP64 = 0, 1, 2, 3
F_VNEW, F_V, F_CA, F_REV, F_B, F_IEXT, F_Y, F_YN, F_FN, F_CA2, F_CA_OLD, F_VOLD = range(12)
F_GATE0 = 100
REC_ALL = -1     # cb_pop_run: record every compartment
STIM_NONE = -2   # cb_pop_config.stim_comp: no injected current (network_sim on)
H_MAX = 0.125

# every symbol include/cbrkc.h declares
EXPORTS = [
    "cb_config_default", "cb_pop_create", "cb_pop_destroy", "cb_pop_set_current", "cb_pop_advance",
    "cb_pop_run", "cb_pop_run_noalloc", "cb_pop_set_raster", "cb_pop_spikes", "cb_pop_read", "cb_pop_size", "cb_pop_stats", "cb_pop_controller",
    "cb_synth_pkj_morph", "cb_philox4x32_10", "cb_last_error", "cb_pkj_initialize",
    "cb_pkj_solve_update_v", "cb_pkj_set_current", "cb_pkj_output_file", "cb_pkj_finalize",
    "cb_pop_kernel_time", "cb_pop_autotune", "cb_pop_mapping", "cb_pkj_cable_spectral_radius",
    "cb_io_initialize", "cb_io_solve_update_v", "cb_io_set_current", "cb_io_output_file", "cb_io_finalize",
    "cb_go_initialize", "cb_go_solve_update_v", "cb_go_set_current", "cb_go_output_file", "cb_go_finalize",
    "cb_gr_initialize", "cb_gr_solve_update_v", "cb_gr_set_current", "cb_gr_output_file", "cb_gr_finalize",
    "cb_mfgr_create", "cb_gogr_create", "cb_grgo_create", "cb_grpkj_create", "cb_mlipkj_create",
    "cb_syn_finalize", "cb_gr_synapse_update", "cb_go_synapse_update", "cb_pkj_synapse_update",
    "cb_syn_size", "cb_syn_comp", "cb_syn_read", "cb_syn_build",
    "cb_mix_config_default", "cb_mix_create", "cb_mix_run", "cb_mix_layout", "cb_mix_stats", "cb_mix_autotune",
    "cb_mix_kernel_time", "cb_mix_use_graph", "cb_mix_destroy", "cb_synth_gr_chains", "cb_synth_go_trees",
    "cb_synth_io_gj", "cb_net_config_default", "cb_net_generate", "cb_net_create", "cb_net_run", "cb_net_spikes",
    "cb_net_conductance", "cb_net_stats", "cb_net_time", "cb_net_destroy",
]

# synapse kinds and fields (cb_syn_build, cb_syn_read)
SYN_MFGR, SYN_GOGR, SYN_GRGO, SYN_GRPKJ, SYN_MLIPKJ = range(5)
SYN_TRACE0, SYN_TRACE1, SYN_TRACE2, SYN_WEIGHT, SYN_VAL, SYN_OLD_V = range(6)

# The synthetic Purkinje trees (10 um soma on a thin tapering trunk) need
# three safeguards the reference does not have to stay stable: with the
# reference's ion-update interval 2^-6 ms, 709 of 1024 such cells blow up
# within 200 ms (115 at 2^-7, none at 2^-8); the stale fn after an ion update
# and spectral radii underestimated by the power method (up to 2x) each
# produced blow-ups as well.
STABLE = dict(ion_interval=2.0 ** -8, refresh_fn=True, rho_floor=True)


class CbError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [
        ("cell_type", ctypes.c_int),
        ("solver", ctypes.c_int),
        ("engine", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("morph_path", ctypes.c_char_p),
        ("gj_path", ctypes.c_char_p),
        ("atol", ctypes.c_double),
        ("rtol", ctypes.c_double),
        ("v0", ctypes.POINTER(ctypes.c_double)),
        ("stim_comp", ctypes.c_int),
        ("stim_t0", ctypes.c_double),
        ("stim_t1", ctypes.c_double),
        ("stim_amp", ctypes.POINTER(ctypes.c_double)),
        ("poisson_on", ctypes.c_int),
        ("poisson_rate_hz", ctypes.c_double),
        ("poisson_g_mS", ctypes.c_double),
        ("poisson_tau_ms", ctypes.c_double),
        ("poisson_e_mV", ctypes.c_double),
        ("poisson_seed", ctypes.c_uint64),
        ("ion_interval_ms", ctypes.c_double),
        ("refresh_fn", ctypes.c_int),
        ("rho_floor", ctypes.c_int),
        ("device", ctypes.c_int),
        ("step_ms", ctypes.c_double),
        ("hmin_ms", ctypes.c_double),
        ("comp", ctypes.c_int),
    ]


class NetConfig(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("deg_mf_gr", ctypes.c_int),
        ("deg_go_gr", ctypes.c_int),
        ("deg_gr_go", ctypes.c_int),
        ("deg_gr_pkj", ctypes.c_int),
        ("total", ctypes.c_int64),
        ("w_min_nS", ctypes.c_double),
        ("w_max_nS", ctypes.c_double),
        ("tau_ampa_ms", ctypes.c_double),
        ("tau_gaba_ms", ctypes.c_double),
        ("e_ampa_mV", ctypes.c_double),
        ("e_gaba_mV", ctypes.c_double),
        ("delay_ms", ctypes.c_double),
        ("mf_rate_hz", ctypes.c_double),
        ("gr_target_comp", ctypes.c_int),
        ("go_target_lo", ctypes.c_int),
        ("go_target_hi", ctypes.c_int),
        ("pkj_target_lo", ctypes.c_int),
        ("pkj_target_hi", ctypes.c_int),
        ("gr_spike_comp", ctypes.c_int),
        ("go_spike_comp", ctypes.c_int),
        ("pkj_spike_comp", ctypes.c_int),
        ("gr_thr_mV", ctypes.c_double),
        ("go_thr_mV", ctypes.c_double),
        ("pkj_thr_mV", ctypes.c_double),
    ]


class MixPop(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("morph_path", ctypes.c_char_p),
        ("comp", ctypes.c_int),
        ("v0", ctypes.POINTER(ctypes.c_double)),
        ("amp", ctypes.POINTER(ctypes.c_double)),
        ("stim_comp", ctypes.c_int),
        ("rec_comp", ctypes.c_int),
        ("thr", ctypes.c_double),
    ]


class MixConfig(ctypes.Structure):
    _fields_ = [
        ("pop", MixPop * 4),
        ("io_gj_path", ctypes.c_char_p),
        ("interval_ms", ctypes.c_double),
        ("pkj_atol", ctypes.c_double),
        ("pkj_rtol", ctypes.c_double),
        ("pkj_stable", ctypes.c_int),
        ("device", ctypes.c_int),
    ]


_lib = None


def lib():
    """Loads libcbrkc.so (built in-tree); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CbError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.cb_last_error.restype = ctypes.c_char_p
        L.cb_config_default.argtypes = [ctypes.POINTER(Config), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.cb_pop_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
        L.cb_pop_destroy.argtypes = [vp]
        L.cb_pop_set_current.argtypes = [vp, ctypes.c_double]
        L.cb_pop_advance.argtypes = [vp]
        L.cb_pop_run.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp]
        L.cb_pop_run_noalloc.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp]
        L.cb_pop_set_raster.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.cb_pop_spikes.argtypes = [vp, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        ip = ctypes.POINTER(ctypes.c_int)
        L.cb_mix_config_default.argtypes = [ctypes.POINTER(MixConfig)]
        L.cb_mix_create.argtypes = [ctypes.POINTER(MixConfig), ctypes.POINTER(vp)]
        L.cb_mix_run.argtypes = [vp, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_uint32)]
        L.cb_mix_layout.argtypes = [vp, ip, ip, ip, ip]
        L.cb_mix_stats.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        L.cb_mix_autotune.argtypes = [vp, ctypes.c_int, ip]
        L.cb_mix_kernel_time.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int]
        L.cb_mix_use_graph.argtypes = [vp, ctypes.c_int]
        L.cb_mix_destroy.argtypes = [vp]
        L.cb_synth_gr_chains.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ip,
                                         ctypes.c_int]
        L.cb_synth_go_trees.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int]
        L.cb_synth_io_gj.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double]
        L.cb_pop_read.argtypes = [vp, ctypes.c_int, dp]
        L.cb_pop_size.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.cb_pop_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.cb_pop_controller.argtypes = [vp, dp, dp]
        L.cb_pop_kernel_time.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int]
        L.cb_pop_mapping.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        L.cb_pop_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.cb_pkj_cable_spectral_radius.argtypes = [ctypes.c_char_p, ctypes.c_int, dp]
        L.cb_synth_pkj_morph.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.cb_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.cb_philox4x32_10.restype = None
        L.cb_pkj_initialize.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.cb_pkj_solve_update_v.argtypes = [vp]
        L.cb_pkj_set_current.argtypes = [vp, ctypes.c_double]
        L.cb_pkj_output_file.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_char_p, ctypes.c_char_p]
        L.cb_pkj_finalize.argtypes = [vp]
        L.cb_io_initialize.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.cb_io_solve_update_v.argtypes = [vp]
        L.cb_io_set_current.argtypes = [vp, ctypes.c_double]
        L.cb_io_output_file.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_char_p, ctypes.c_char_p]
        L.cb_io_finalize.argtypes = [vp]
        L.cb_go_initialize.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.cb_go_solve_update_v.argtypes = [vp]
        L.cb_go_set_current.argtypes = [vp, ctypes.c_double]
        L.cb_go_output_file.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_char_p, ctypes.c_char_p]
        L.cb_go_finalize.argtypes = [vp]
        L.cb_gr_initialize.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.cb_gr_solve_update_v.argtypes = [vp]
        L.cb_gr_set_current.argtypes = [vp, ctypes.c_double]
        L.cb_gr_output_file.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_char_p, ctypes.c_char_p]
        L.cb_gr_finalize.argtypes = [vp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.cb_mfgr_create.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.cb_gogr_create.argtypes = [vp, vp, ctypes.POINTER(vp)]
        L.cb_grgo_create.argtypes = [vp, vp, ctypes.POINTER(vp)]
        L.cb_grpkj_create.argtypes = [vp, vp, ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.cb_mlipkj_create.argtypes = [vp, vp, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.cb_syn_finalize.argtypes = [vp]
        L.cb_gr_synapse_update.argtypes = [ctypes.c_double, vp, vp]
        L.cb_go_synapse_update.argtypes = [ctypes.c_double, vp]
        L.cb_pkj_synapse_update.argtypes = [ctypes.c_double, vp, vp]
        L.cb_syn_size.argtypes = [vp, ip]
        L.cb_syn_comp.argtypes = [vp, ip, ip]
        L.cb_syn_read.argtypes = [vp, ctypes.c_int, dp]
        L.cb_syn_build.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp,
                                   ctypes.c_uint64, ip, ip, ip, dp]
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.cb_net_config_default.argtypes = [ctypes.POINTER(NetConfig)]
        L.cb_net_generate.argtypes = [ctypes.POINTER(NetConfig), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_int8), ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32), dp, i64p]
        L.cb_net_create.argtypes = [vp, vp, vp, ctypes.POINTER(NetConfig), ctypes.POINTER(vp)]
        L.cb_net_run.argtypes = [vp, ctypes.c_int]
        L.cb_net_spikes.argtypes = [vp, ctypes.c_int, dp, ip, ip, ip]
        L.cb_net_conductance.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.cb_net_stats.argtypes = [vp, i64p]
        L.cb_net_time.argtypes = [vp, dp, dp, dp]
        L.cb_net_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _check(rc):
    if rc != CB_OK:
        raise CbError(f"cbrkc error {rc}: {lib().cb_last_error().decode()}")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib().cb_philox4x32_10(c.ctypes.data_as(u32p), k.ctypes.data_as(u32p), o.ctypes.data_as(u32p))
    return o


def synth_pkj_morph(path, seed, ncell, soma_r_um=10.0, lmin_um=5.0, lmax_um=20.0, d0_um=3.0, d1_um=0.5,
                    dscale=None):
    """Writes ncell seeded synthetic Purkinje trees (Pkj.txt format) to path."""
    ds = None
    if dscale is not None:
        dscale = np.ascontiguousarray(dscale, dtype=np.float64)
        ds = _dp(dscale)
    _check(lib().cb_synth_pkj_morph(path.encode(), seed, ncell, soma_r_um, lmin_um, lmax_um, d0_um, d1_um, ds))


def synth_gr_chains(path, template, seed, ncell, types=None):
    """granule chains in the Gr.txt format (default: dendrite, soma, axon
    initial segment, axon), geometry from the template's rows x U(0.8, 1.25)"""
    t = None
    if types is not None:
        arr = (ctypes.c_int * len(types))(*types)
        t = arr
    _check(lib().cb_synth_gr_chains(path.encode(), template.encode(), seed, ncell, t, len(types) if types else 0))


def synth_go_trees(path, template, seed, ncell):
    """Golgi trees with the template's topology, geometry x U(0.8, 1.25)"""
    _check(lib().cb_synth_go_trees(path.encode(), template.encode(), seed, ncell))


def synth_io_gj(path, seed, ncell, k=6, g_min_nS=0.5, g_max_nS=1.5):
    """olive gap junctions: every cell joined to k random partners"""
    _check(lib().cb_synth_io_gj(path.encode(), seed, ncell, k, g_min_nS, g_max_nS))


class Mix:
    """The batched engine over a mix of cell types (cb_mix_*).  pops maps a
    cell type (GR, GO, PKJ, IO) to a dict with n, morph, v0 and optionally
    amp, comp, stim_comp, rec_comp, thr."""

    def __init__(self, pops, io_gj=None, interval_ms=0.125, pkj_tol=1e-6, pkj_stable=True, device=0):
        L = lib()
        cfg = MixConfig()
        _check(L.cb_mix_config_default(ctypes.byref(cfg)))
        self._keep = []
        for t, spec in pops.items():
            p = cfg.pop[t]
            p.n = spec["n"]
            mp = spec["morph"].encode()
            self._keep.append(mp)
            p.morph_path = mp
            v0 = np.ascontiguousarray(spec["v0"], dtype=np.float64)
            self._keep.append(v0)
            p.v0 = _dp(v0)
            if spec.get("amp") is not None:
                amp = np.ascontiguousarray(np.broadcast_to(spec["amp"], (p.n,)), dtype=np.float64)
                self._keep.append(amp)
                p.amp = _dp(amp)
            for k in ("comp", "stim_comp", "rec_comp", "thr"):
                if k in spec:
                    setattr(p, k, spec[k])
        if io_gj:
            g = io_gj.encode()
            self._keep.append(g)
            cfg.io_gj_path = g
        cfg.interval_ms = interval_ms
        cfg.pkj_atol = cfg.pkj_rtol = pkj_tol
        cfg.pkj_stable = 1 if pkj_stable else 0
        cfg.device = device
        h = ctypes.c_void_p()
        _check(L.cb_mix_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        n4, w4, l4 = (ctypes.c_int * 4)(), (ctypes.c_int * 4)(), (ctypes.c_int * 4)()
        w = ctypes.c_int()
        _check(L.cb_mix_layout(h, n4, w4, ctypes.byref(w), l4))
        self.ncell = list(n4)
        self.word_off = list(w4)
        self.words = w.value
        self.offset = [int(sum(self.ncell[:t])) for t in range(4)]
        self.total = int(sum(self.ncell))

    def close(self):
        if getattr(self, "h", None):
            _check(lib().cb_mix_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nint, record=True):
        """(rec [nint, total cells], spikes [nint, words] uint32)"""
        rec = np.zeros((nint, self.total)) if record else None
        spk = np.zeros((nint, max(self.words, 1)), dtype=np.uint32) if record else None
        _check(lib().cb_mix_run(self.h, nint, _dp(rec) if record else None,
                                spk.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if record else None))
        return rec, spk

    def trace(self, rec, t):
        """the recorded voltages of type t, [nint, cells]"""
        return rec[:, self.offset[t]:self.offset[t] + self.ncell[t]]

    def spikes(self, spk, t):
        """[nint, cells] booleans: an upward threshold crossing in that interval"""
        n, w0 = self.ncell[t], self.word_off[t]
        c = np.arange(n)
        return ((spk[:, w0 + c // 32] >> (c % 32).astype(np.uint32)) & 1).astype(bool)

    def stats(self, t):
        s = (ctypes.c_int64 * 6)()
        _check(lib().cb_mix_stats(self.h, t, s))
        return dict(steps=s[0], rejects=s[1], stages=s[2], rho_iters=s[3], ion_updates=s[4], macro_steps=s[5])

    def lanes(self):
        n4, w4, l4 = (ctypes.c_int * 4)(), (ctypes.c_int * 4)(), (ctypes.c_int * 4)()
        w = ctypes.c_int()
        _check(lib().cb_mix_layout(self.h, n4, w4, ctypes.byref(w), l4))
        return list(l4)

    def autotune(self, trial=8):
        l4 = (ctypes.c_int * 4)()
        _check(lib().cb_mix_autotune(self.h, trial, l4))
        return list(l4)

    def kernel_time(self, reset=False):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        _check(lib().cb_mix_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n), 1 if reset else 0))
        return ms.value, n.value

    def use_graph(self, on):
        _check(lib().cb_mix_use_graph(self.h, 1 if on else 0))


def cable_spectral_radius(path, ncell):
    """Host-only: exact spectral radius of each cell's cable operator C^-1 L."""
    out = np.zeros(ncell)
    _check(lib().cb_pkj_cable_spectral_radius(path.encode(), ncell, _dp(out)))
    return out


class Population:
    """One cell population with its solver (the reference's neuron_t + neuron_solve_t)."""

    def __init__(self, cell_type=PKJ, nx=1, ny=1, solver=RKC, engine=ENGINE_GLOBAL_FP64, morph=None, gj="",
                 atol=None, rtol=None, v0=None, stim=None, poisson=None, ion_interval=None, refresh_fn=False,
                 rho_floor=False, device=0, step_ms=None, hmin_ms=None, comp=None):
        L = lib()
        cfg = Config()
        _check(L.cb_config_default(ctypes.byref(cfg), cell_type, solver, nx, ny))
        cfg.engine = engine
        self._keep = []
        if morph:
            cfg.morph_path = morph.encode()
        # Purkinje grids: gj names the gap-junction pair list; "" (the default
        # here) builds uncoupled cells, None reads ./Pkj_GJ.csv as the
        # reference does and fails when it is missing
        if gj is not None:
            cfg.gj_path = gj.encode()
        if atol:
            cfg.atol = atol
        if rtol:
            cfg.rtol = rtol
        if v0 is not None:
            v0 = np.ascontiguousarray(v0, dtype=np.float64)
            self._keep.append(v0)
            cfg.v0 = _dp(v0)
        if stim == "none":
            cfg.stim_comp = STIM_NONE
        elif stim is not None:
            stim_comp, t0, t1, amp = stim
            amp = np.ascontiguousarray(np.broadcast_to(np.asarray(amp, dtype=np.float64), (nx * ny,)))
            self._keep.append(amp)
            cfg.stim_comp, cfg.stim_t0, cfg.stim_t1 = stim_comp, t0, t1
            cfg.stim_amp = _dp(amp)
        if poisson is not None:
            cfg.poisson_on = 1
            cfg.poisson_rate_hz = poisson.get("rate_hz", 5.0)
            cfg.poisson_g_mS = poisson.get("g_mS", 1e-6)
            cfg.poisson_tau_ms = poisson.get("tau_ms", 2.0)
            cfg.poisson_e_mV = poisson.get("e_mV", 0.0)
            cfg.poisson_seed = poisson.get("seed", 0)
        if ion_interval:
            cfg.ion_interval_ms = ion_interval
        cfg.refresh_fn = 1 if refresh_fn else 0
        cfg.rho_floor = 1 if rho_floor else 0
        cfg.device = device
        if step_ms:
            cfg.step_ms = step_ms
        if hmin_ms:
            cfg.hmin_ms = hmin_ms
        if comp:
            cfg.comp = comp
        h = ctypes.c_void_p()
        _check(L.cb_pop_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        n = ctypes.c_int()
        nc = ctypes.c_int()
        _check(L.cb_pop_size(h, ctypes.byref(n), ctypes.byref(nc)))
        self.n, self.nc = n.value, nc.value
        self.comp = self.nc // max(self.n, 1)
        self.engine = engine

    def close(self):
        if getattr(self, "h", None):
            _check(lib().cb_pop_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_current(self, t):
        _check(lib().cb_pop_set_current(self.h, t))

    def advance(self):
        _check(lib().cb_pop_advance(self.h))

    def run(self, nsteps, t0=0.0, rec_comp=1599, record=True):
        out = (np.zeros((nsteps, self.n)) if rec_comp >= 0 else np.zeros((nsteps, self.nc))) if record else None
        _check(lib().cb_pop_run(self.h, nsteps, t0, rec_comp, _dp(out) if record else None))
        return out

    def set_raster(self, comp, thr):
        """device raster on compartment comp (upward crossings of thr mV); comp < 0 stops it"""
        _check(lib().cb_pop_set_raster(self.h, comp, thr))

    def spikes(self):
        """all spikes found so far, as (times [ms], cells), taken out of the list"""
        ts, cs = [], []
        while True:
            t = np.zeros(4096)
            c = np.zeros(4096, dtype=np.int32)
            n = ctypes.c_int()
            _check(lib().cb_pop_spikes(self.h, 4096, _dp(t), c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       ctypes.byref(n)))
            ts.append(t[:n.value])
            cs.append(c[:n.value])
            if n.value < 4096:
                break
        return np.concatenate(ts), np.concatenate(cs)

    def run_noalloc(self, nsteps, t0=0.0, rec_comp=1599, record=True):
        """run() through cb_pop_run_noalloc (no device allocation)"""
        out = np.zeros((nsteps, self.n)) if record else None
        _check(lib().cb_pop_run_noalloc(self.h, nsteps, t0, rec_comp, _dp(out) if record else None))
        return out

    def read(self, field=F_VNEW):
        a = np.empty(self.nc)
        _check(lib().cb_pop_read(self.h, field, _dp(a)))
        return a

    def stats(self):
        s = (ctypes.c_int64 * 6)()
        _check(lib().cb_pop_stats(self.h, s))
        return dict(steps=s[0], rejects=s[1], stages=s[2], rho_iters=s[3], ion_updates=s[4], macro_steps=s[5])

    def kernel_time(self, reset=False):
        """(milliseconds, launches) of the main kernel, HIP-event timed on the engine's stream"""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        _check(lib().cb_pop_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n), 1 if reset else 0))
        return ms.value, n.value

    def autotune(self, trial_steps=8, threads=0):
        """Fastest launch mapping on a trial; threads > 0 selects one instead."""
        t = ctypes.c_int(threads)
        _check(lib().cb_pop_autotune(self.h, trial_steps, ctypes.byref(t)))
        return t.value

    def mapping(self):
        """the launch mapping in use (and the last autotune trial), as text"""
        buf = ctypes.create_string_buffer(4096)
        _check(lib().cb_pop_mapping(self.h, buf, 4096))
        return buf.value.decode()

    def controller(self):
        o = np.zeros(14)
        hw = np.zeros(8)
        _check(lib().cb_pop_controller(self.h, _dp(o), _dp(hw)))
        return o, hw


# ---- reference-named entry points (Purkinje) ----

NET_MFGR, NET_GOGR, NET_GRGO, NET_GRPKJ = range(4)
NET_GR, NET_GO, NET_PKJ = range(3)
NET_AMPA, NET_GABA = range(2)


def net_config(**kw):
    """cb_net_config_default, then the given fields"""
    cfg = NetConfig()
    _check(lib().cb_net_config_default(ctypes.byref(cfg)))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise CbError(f"no network field {k}")
        setattr(cfg, k, v)
    return cfg


def net_generate(cfg, n, comp):
    """the synapse list on the host (cb_net_generate): kind, pre, post, w_nS"""
    L = lib()
    cnt = ctypes.c_int64()
    _check(L.cb_net_generate(ctypes.byref(cfg), *n, *comp, 0, None, None, None, None, ctypes.byref(cnt)))
    m = cnt.value
    kind = np.zeros(m, np.int8)
    pre = np.zeros(m, np.int32)
    post = np.zeros(m, np.int32)
    w = np.zeros(m, np.float64)
    _check(L.cb_net_generate(ctypes.byref(cfg), *n, *comp, m, kind.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                             pre.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                             post.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(w), ctypes.byref(cnt)))
    return kind, pre, post, w


class Network:
    """A fixed in-degree network over a granule, a Golgi and a Purkinje
    population (cb_net_*); the populations must outlive it."""

    def __init__(self, gr, go, pkj, cfg=None):
        self.pops = (gr, go, pkj)
        cfg = cfg if cfg is not None else net_config()
        h = ctypes.c_void_p()
        _check(lib().cb_net_create(gr.h, go.h, pkj.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def run(self, nint):
        _check(lib().cb_net_run(self.h, nint))

    def spikes(self):
        """every raster entry so far as arrays (t [ms], population, cell)"""
        L = lib()
        ts, ps, cs = [], [], []
        while True:
            cap = 4096
            t = np.zeros(cap)
            p = np.zeros(cap, np.int32)
            c = np.zeros(cap, np.int32)
            n = ctypes.c_int()
            _check(L.cb_net_spikes(self.h, cap, _dp(t), p.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(n)))
            ts.append(t[:n.value])
            ps.append(p[:n.value])
            cs.append(c[:n.value])
            if n.value < cap:
                break
        return np.concatenate(ts), np.concatenate(ps), np.concatenate(cs)

    def conductance(self, pop, receptor):
        out = np.zeros(self.pops[pop].nc)
        _check(lib().cb_net_conductance(self.h, pop, receptor, _dp(out)))
        return out

    def stats(self):
        out = (ctypes.c_int64 * 9)()
        _check(lib().cb_net_stats(self.h, out))
        v = list(out)
        return dict(synapses=v[0:4], slots=v[4:7], intervals=v[7], delay=v[8])

    def times(self):
        """host wall ms: per population, exchanges, total"""
        pop = np.zeros(3)
        ex = ctypes.c_double()
        tot = ctypes.c_double()
        _check(lib().cb_net_time(self.h, _dp(pop), ctypes.byref(ex), ctypes.byref(tot)))
        return dict(populations=pop.tolist(), exchange=ex.value, total=tot.value)

    def close(self):
        if getattr(self, "h", None):
            _check(lib().cb_net_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pkj_initialize(nx, ny, solver="RKC"):
    """pkj_initialize + pkj_solve_initialize (pkj.cu:240, pkj_solve.cu:286); reads ./Pkj.txt."""
    h = ctypes.c_void_p()
    _check(lib().cb_pkj_initialize(nx, ny, solver.encode(), ctypes.byref(h)))
    p = Population.__new__(Population)
    p.h = h
    n = ctypes.c_int()
    nc = ctypes.c_int()
    _check(lib().cb_pop_size(h, ctypes.byref(n), ctypes.byref(nc)))
    p.n, p.nc, p.comp, p.engine, p._keep = n.value, nc.value, nc.value // max(n.value, 1), ENGINE_GLOBAL_FP64, []
    return p


def pkj_set_current(p, t):
    _check(lib().cb_pkj_set_current(p.h, t))


def pkj_solve_update_v(p):
    _check(lib().cb_pkj_solve_update_v(p.h))


def pkj_output_file(p, memv, raster_prev, t, v_csv, raster_csv):
    memv = np.ascontiguousarray(memv, dtype=np.float64)
    _check(lib().cb_pkj_output_file(p.h, _dp(memv), _dp(raster_prev), t, v_csv.encode(), raster_csv.encode()))


def pkj_finalize(p):
    p.close()


# ---- reference-named entry points (inferior olive, Golgi cells) ----

def _wrap(h):
    p = Population.__new__(Population)
    p.h = h
    n = ctypes.c_int()
    nc = ctypes.c_int()
    _check(lib().cb_pop_size(h, ctypes.byref(n), ctypes.byref(nc)))
    p.n, p.nc, p.comp, p.engine, p._keep = n.value, nc.value, nc.value // max(n.value, 1), ENGINE_GLOBAL_FP64, []
    return p


def io_initialize(nx, ny, solver="RKC"):
    """io_initialize + io_solve_initialize + io_gap_create; reads ./IO.txt."""
    h = ctypes.c_void_p()
    _check(lib().cb_io_initialize(nx, ny, solver.encode(), ctypes.byref(h)))
    return _wrap(h)


def io_set_current(p, t):
    _check(lib().cb_io_set_current(p.h, t))


def io_solve_update_v(p):
    _check(lib().cb_io_solve_update_v(p.h))


def io_output_file(p, memv, raster_prev, t, v_csv, raster_csv):
    memv = np.ascontiguousarray(memv, dtype=np.float64)
    _check(lib().cb_io_output_file(p.h, _dp(memv), _dp(raster_prev), t, v_csv.encode(), raster_csv.encode()))


def io_finalize(p):
    p.close()


def go_initialize(nx, ny, solver="RKC"):
    """go_initialize + go_solve_initialize; reads ./Go_131.txt."""
    h = ctypes.c_void_p()
    _check(lib().cb_go_initialize(nx, ny, solver.encode(), ctypes.byref(h)))
    return _wrap(h)


def go_set_current(p, t):
    _check(lib().cb_go_set_current(p.h, t))


def go_solve_update_v(p):
    _check(lib().cb_go_solve_update_v(p.h))


def go_output_file(p, memv, raster_prev, t, v_csv, raster_csv):
    memv = np.ascontiguousarray(memv, dtype=np.float64)
    _check(lib().cb_go_output_file(p.h, _dp(memv), _dp(raster_prev), t, v_csv.encode(), raster_csv.encode()))


def go_finalize(p):
    p.close()


def gr_initialize(nx, ny, solver="RKC"):
    """gr_initialize + gr_solve_initialize; reads ./Gr.txt."""
    h = ctypes.c_void_p()
    _check(lib().cb_gr_initialize(nx, ny, solver.encode(), ctypes.byref(h)))
    return _wrap(h)


def gr_set_current(p, t):
    _check(lib().cb_gr_set_current(p.h, t))


def gr_solve_update_v(p):
    _check(lib().cb_gr_solve_update_v(p.h))


def gr_output_file(p, memv, raster_prev, t, v_csv, raster_csv):
    memv = np.ascontiguousarray(memv, dtype=np.float64)
    _check(lib().cb_gr_output_file(p.h, _dp(memv), _dp(raster_prev), t, v_csv.encode(), raster_csv.encode()))


def gr_finalize(p):
    p.close()


# ---- synapses (syn.cu) ----

class Synapses:
    """One synapse set (the reference's synapse_t); coupled into its
    postsynaptic population on creation, uncoupled by close()."""

    def __init__(self, h, keep=()):
        self.h = h
        self._keep = list(keep)
        n = ctypes.c_int()
        _check(lib().cb_syn_size(self.h, ctypes.byref(n)))
        self.n = n.value

    def comp(self):
        pre, post = np.zeros(self.n, dtype=np.int32), np.zeros(self.n, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        if self.n:
            _check(lib().cb_syn_comp(self.h, pre.ctypes.data_as(ip), post.ctypes.data_as(ip)))
        return pre, post

    def read(self, field):
        out = np.zeros(self.n)
        if self.n:
            _check(lib().cb_syn_read(self.h, field, _dp(out)))
        return out

    def close(self):
        if self.h:
            _check(lib().cb_syn_finalize(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ph(p):
    return p.h if p is not None else None


def _syn(fn, *args, keep=()):
    h = ctypes.c_void_p()
    _check(fn(*args, ctypes.byref(h)))
    return Synapses(h, keep)


def mfgr_create(gr, seed):
    """mfgr_create (syn.cu:104-128)."""
    return _syn(lib().cb_mfgr_create, _ph(gr), seed, keep=(gr,))


def gogr_create(go, gr):
    """gogr_create (syn.cu:339-410)."""
    return _syn(lib().cb_gogr_create, _ph(go), _ph(gr), keep=(go, gr))


def grgo_create(gr, go):
    """grgo_create (syn.cu:181-261)."""
    return _syn(lib().cb_grgo_create, _ph(gr), _ph(go), keep=(gr, go))


def grpkj_create(gr, pkj, location_csv, seed):
    """grpkj_create (syn.cu:519-644)."""
    return _syn(lib().cb_grpkj_create, _ph(gr), _ph(pkj), location_csv.encode() if location_csv else None, seed,
                keep=(gr, pkj))


def mlipkj_create(pkj, gr, seed):
    """mlipkj_create (syn.cu:690-738)."""
    return _syn(lib().cb_mlipkj_create, _ph(pkj), _ph(gr), seed, keep=(pkj, gr))


def gr_synapse_update(t, mfgr, gogr):
    _check(lib().cb_gr_synapse_update(t, _ph(mfgr), _ph(gogr)))


def go_synapse_update(t, grgo):
    _check(lib().cb_go_synapse_update(t, _ph(grgo)))


def pkj_synapse_update(t, grpkj, mlipkj=None):
    _check(lib().cb_pkj_synapse_update(t, _ph(grpkj), _ph(mlipkj)))


def syn_build(kind, nx_a, ny_a, nx_b=0, ny_b=0, x1600=None, seed=0):
    """The connection rules on the host (no device): (pre, post, weight)."""
    n = ctypes.c_int()
    xp = None
    if x1600 is not None:
        x1600 = np.ascontiguousarray(x1600, dtype=np.float64)
        xp = _dp(x1600)
    _check(lib().cb_syn_build(kind, nx_a, ny_a, nx_b, ny_b, xp, seed, ctypes.byref(n), None, None, None))
    pre, post = np.zeros(n.value, dtype=np.int32), np.zeros(n.value, dtype=np.int32)
    w = np.zeros(n.value)
    ip = ctypes.POINTER(ctypes.c_int)
    _check(lib().cb_syn_build(kind, nx_a, ny_a, nx_b, ny_b, xp, seed, ctypes.byref(n), pre.ctypes.data_as(ip),
                              post.ctypes.data_as(ip), _dp(w)))
    return pre, post, w
