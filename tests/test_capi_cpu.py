"""The C-ABI library loads and exports every symbol include/cbrkc.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

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
