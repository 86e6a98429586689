"""The C-ABI library loads and exports every symbol include/cbrkc.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import numpy as np

import cbrkc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "cbrkc.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(cb_\w+)\s*\(", txt, re.M)))


def test_header_and_mirror_agree():
    assert declared_symbols() == sorted(cbrkc.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(cbrkc.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_config_defaults_and_error_path():
    L = cbrkc.lib()
    cfg = cbrkc.Config()
    assert L.cb_config_default(ctypes.byref(cfg), cbrkc.PKJ, cbrkc.RKC, 3, 3) == 0
    assert (cfg.nx, cfg.ny, cfg.stim_comp, cfg.engine) == (3, 3, -1, cbrkc.ENGINE_GLOBAL_FP64)
    h = ctypes.c_void_p()
    cfg.morph_path = b"/nonexistent/Pkj.txt"
    rc = L.cb_pop_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc != 0 and not h.value
    assert cbrkc.lib().cb_last_error().decode() != ""


def test_philox_known_answers_through_abi():
    # Random123 known-answer vectors for philox4x32-10
    assert list(cbrkc.philox4x32_10([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(cbrkc.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2)) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(cbrkc.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0])) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_oracle_philox_known_answers():
    """The oracle's own Philox4x32-10 (oracle/or_philox.c, the counter RNG of
    the synthetic inputs and the synapse draws) on the same published
    known-answer vectors, independently of the library's."""
    import orc
    assert [int(x) for x in orc.philox([0, 0, 0, 0], [0, 0])] == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert [int(x) for x in orc.philox([0xffffffff] * 4, [0xffffffff] * 2)] == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert [int(x) for x in orc.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                       [0xa4093822, 0x299f31d0])] == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def _create(tmp_path, monkeypatch, **kw):
    """cb_pop_create with cb_config_default plus the given fields, in tmp_path
    as working directory; returns (status, message)."""
    L = cbrkc.lib()
    cfg = cbrkc.Config()
    assert L.cb_config_default(ctypes.byref(cfg), kw.pop("cell_type", cbrkc.PKJ), kw.pop("solver", cbrkc.RKC),
                               kw.pop("nx", 1), kw.pop("ny", 1)) == 0
    keep = []
    for k, v in kw.items():
        if isinstance(v, str):
            v = v.encode()
            keep.append(v)
        setattr(cfg, k, v)
    monkeypatch.chdir(tmp_path)
    h = ctypes.c_void_p()
    rc = L.cb_pop_create(ctypes.byref(cfg), ctypes.byref(h))
    assert not h.value
    return rc, L.cb_last_error().decode()


CB_E_ARG, CB_E_IO, CB_E_UNSUP = -1, -2, -5


def test_failure_paths_before_the_device(tmp_path, monkeypatch, ref_data):
    """Argument and input-file failures come back as status codes with a
    message, without a GPU: the reference prints and exits on each of these
    (pkj.cu:28 missing morphology, pkj_solve.cu:86-87 missing Pkj_GJ.csv).
    The reference's idid = 3 (an error weight of zero, solve_rkc.cu:50-53)
    needs atol = 0; a non-positive atol selects the cell type's TOL here, so
    every weight is positive by construction."""
    import shutil
    shutil.copy(os.path.join(ref_data, "Pkj.txt"), tmp_path / "Pkj.txt")
    # a Purkinje grid needs ./Pkj_GJ.csv, as the reference does
    rc, msg = _create(tmp_path, monkeypatch, nx=2, ny=2)
    assert rc == CB_E_IO and "Pkj_GJ.csv" in msg, (rc, msg)
    rc, msg = _create(tmp_path, monkeypatch, morph_path="/nonexistent/Pkj.txt")
    assert rc == CB_E_IO, (rc, msg)
    rc, msg = _create(tmp_path, monkeypatch, cell_type=cbrkc.IO, morph_path="/nonexistent/IO.txt")
    assert rc == CB_E_IO, (rc, msg)
    # an explicit pair list: the per-cell engines refuse coupled cells
    shutil.copy(os.path.join(ref_data, "Pkj_GJ.csv"), tmp_path / "Pkj_GJ.csv")
    rc, msg = _create(tmp_path, monkeypatch, nx=2, ny=2, engine=cbrkc.ENGINE_CELL_MIXED)
    assert rc == CB_E_UNSUP and "gap junctions" in msg, (rc, msg)
    rc, msg = _create(tmp_path, monkeypatch, solver=7)
    assert rc == CB_E_ARG, (rc, msg)
    rc, msg = _create(tmp_path, monkeypatch, cell_type=9)
    assert rc == CB_E_ARG, (rc, msg)
    rc, msg = _create(tmp_path, monkeypatch, solver=cbrkc.CN, poisson_on=1)
    assert rc == CB_E_UNSUP, (rc, msg)
    rc, msg = _create(tmp_path, monkeypatch, step_ms=0.0625)
    assert rc == CB_E_ARG and "step_ms" in msg, (rc, msg)
    rc, msg = _create(tmp_path, monkeypatch, stim_comp=1600, stim_t0=0.0, stim_t1=1.0)
    assert rc == CB_E_ARG and "out of range" in msg, (rc, msg)
    # reference-named entry point: an unknown solver name (main.cu:60-102 prefix match)
    h = ctypes.c_void_p()
    assert cbrkc.lib().cb_pkj_initialize(1, 1, b"XYZ", ctypes.byref(h)) == CB_E_ARG


def test_granule_population_reads_the_reference_file_with_a_stimulus_compartment():
    """The mirror's stimulus tuple must not leak into cb_pop_config.comp (a
    stimulus on compartment 16 once set a 16-compartment chain and the
    reference's Gr.txt failed to read).  Files are read before any device
    call, so without a GPU the call stops at the HIP layer; with one it
    succeeds."""
    gr_txt = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_data", "Gr.txt")
    try:
        p = cbrkc.Population(cbrkc.GR, 2, 1, morph=gr_txt, v0=np.array([-70.0, -72.0]),
                             stim=(16, 2.0, 1e9, np.array([1e-5, 2e-5])))
        assert p.comp == 578
        p.close()
    except cbrkc.CbError as e:
        assert "READING_ERROR" not in str(e) and "error -2" not in str(e), str(e)
