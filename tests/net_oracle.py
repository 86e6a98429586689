"""ctypes bindings to the oracle's network (oracle/or_net.c) and the helpers
the network tests share: the oracle spec from a cb_net_config, synthetic
populations of the three types in the reference formats, and the oracle
populations over the same files.  Test infrastructure only."""
import ctypes
import os

import numpy as np

import cbrkc
import orc

REF_DATA = orc.REF_DATA


class OrNetSpec(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("deg", ctypes.c_int * 4),
        ("total", ctypes.c_longlong),
        ("w_min_nS", ctypes.c_double),
        ("w_max_nS", ctypes.c_double),
        ("tau_ampa_ms", ctypes.c_double),
        ("tau_gaba_ms", ctypes.c_double),
        ("e_ampa_mV", ctypes.c_double),
        ("e_gaba_mV", ctypes.c_double),
        ("delay", ctypes.c_int),
        ("mf_rate_hz", ctypes.c_double),
        ("dt_ms", ctypes.c_double),
        ("n", ctypes.c_int * 3),
        ("comp", ctypes.c_int * 3),
        ("tgt_lo", ctypes.c_int * 3),
        ("tgt_hi", ctypes.c_int * 3),
        ("spike_comp", ctypes.c_int * 3),
        ("thr_mV", ctypes.c_double * 3),
    ]


def L():
    lib = orc.lib()
    if not getattr(lib, "_net_bound", False):
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        sp = ctypes.POINTER(OrNetSpec)
        lib.or_net_size.restype = ctypes.c_longlong
        lib.or_net_size.argtypes = [sp]
        lib.or_net_generate.restype = ctypes.c_longlong
        lib.or_net_generate.argtypes = [sp, ctypes.POINTER(ctypes.c_int8), ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_int32), dp]
        lib.or_net_create.restype = vp
        lib.or_net_create.argtypes = [vp, vp, vp, sp]
        lib.or_net_run.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_ubyte)]
        lib.or_net_conductance.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        lib.or_net_deliver_given.argtypes = [vp, ctypes.POINTER(ctypes.c_ubyte)]
        lib.or_net_free.argtypes = [vp]
        lib.or_gr_create_comp.restype = vp
        lib.or_gr_create_comp.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int]
        lib.or_go_create.restype = vp
        lib.or_go_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int]
        lib.or_rkc_init.argtypes = [vp]
        lib._net_bound = True
    return lib


def spec_of(cfg, n=(0, 0, 0), comp=(0, 0, 0), dt_ms=0.125):
    """the oracle's spec from a cbrkc.NetConfig (fields named alike)"""
    s = OrNetSpec()
    s.seed = cfg.seed
    s.deg[:] = [cfg.deg_mf_gr, cfg.deg_go_gr, cfg.deg_gr_go, cfg.deg_gr_pkj]
    s.total = cfg.total
    s.w_min_nS, s.w_max_nS = cfg.w_min_nS, cfg.w_max_nS
    s.tau_ampa_ms, s.tau_gaba_ms = cfg.tau_ampa_ms, cfg.tau_gaba_ms
    s.e_ampa_mV, s.e_gaba_mV = cfg.e_ampa_mV, cfg.e_gaba_mV
    s.delay = int(round(cfg.delay_ms / dt_ms))
    s.mf_rate_hz, s.dt_ms = cfg.mf_rate_hz, dt_ms
    s.n[:] = list(n)
    s.comp[:] = list(comp)
    s.tgt_lo[:] = [cfg.gr_target_comp, cfg.go_target_lo, cfg.pkj_target_lo]
    s.tgt_hi[:] = [cfg.gr_target_comp + 1, cfg.go_target_hi, cfg.pkj_target_hi]
    s.spike_comp[:] = [cfg.gr_spike_comp, cfg.go_spike_comp, cfg.pkj_spike_comp]
    s.thr_mV[:] = [cfg.gr_thr_mV, cfg.go_thr_mV, cfg.pkj_thr_mV]
    return s


def oracle_generate(spec):
    lib = L()
    m = lib.or_net_size(ctypes.byref(spec))
    kind = np.zeros(m, np.int8)
    pre = np.zeros(m, np.int32)
    post = np.zeros(m, np.int32)
    w = np.zeros(m)
    got = lib.or_net_generate(ctypes.byref(spec), kind.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                              pre.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                              post.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), orc._dp(w))
    assert got == m
    return kind, pre, post, w


def synth_files(tmp, n_gr, n_go, n_pkj, seed=11, shared=False):
    """granule chains (4 compartments), Golgi trees and Purkinje trees, one
    block per cell; shared: one granule and one Golgi block for all cells
    (the reference's format, the only one the oracle's granule and Golgi
    readers take; the Purkinje reader takes a block per cell)"""
    gr = os.path.join(str(tmp), "gr_chains.txt")
    go = os.path.join(str(tmp), "go_trees.txt")
    pk = os.path.join(str(tmp), "pkj_trees.txt")
    cbrkc.synth_gr_chains(gr, os.path.join(REF_DATA, "Gr.txt"), seed, 1 if shared else n_gr)
    cbrkc.synth_go_trees(go, os.path.join(REF_DATA, "Go_131.txt"), seed + 1, 1 if shared else n_go)
    cbrkc.synth_pkj_morph(pk, seed=seed + 2, ncell=n_pkj)
    return gr, go, pk


def oracle_pops(files, v0, pkj_tol=1e-6):
    """the oracle's populations over the files: granule 1e-9 (TOL_GR),
    Golgi 1e-6, Purkinje pkj_tol, initial voltages v0 = (gr, go, pkj)"""
    lib = L()
    gr_f, go_f, pk_f = files
    g0, o0, p0 = [np.ascontiguousarray(v, dtype=np.float64) for v in v0]
    gr = orc.Pop(lib.or_gr_create_comp(gr_f.encode(), 4, len(g0), 1, orc._dp(g0), 1e-9, 1e-9, 0), len(g0), 4)
    go = orc.Pop(lib.or_go_create(go_f.encode(), len(o0), 1, orc._dp(o0), 1e-6, 1e-6, 0), len(o0), 131)
    pk = orc.Pop.pkj(pk_f, len(p0), 1, v0=p0, atol=pkj_tol, rtol=pkj_tol, libc_rand=False)
    for p in (gr, go):
        p.set_stim(0, 0.0, 0.0, np.zeros(p.n))
    pk.set_stim(1599, 0.0, 0.0, np.zeros(pk.n))
    return gr, go, pk
