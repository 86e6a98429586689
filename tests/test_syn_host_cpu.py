"""The library's synapse connection rules (cb_syn_build, host code, no
device) against the oracle's (oracle/or_syn.c): identical contacts and
weights, bit for bit, for several grids and jitter seeds."""
import os

import numpy as np
import pytest

import cbrkc
import orc
from test_oracle_syn import L, Syn, gr_grid, go_grid, pkj_x

KINDS = {"mfgr": cbrkc.SYN_MFGR, "gogr": cbrkc.SYN_GOGR, "grgo": cbrkc.SYN_GRGO, "grpkj": cbrkc.SYN_GRPKJ,
         "mlipkj": cbrkc.SYN_MLIPKJ}


def same(s, built):
    pre, post, w = built
    opre, opost = s.comp()
    assert s.n == len(post)
    np.testing.assert_array_equal(post, opost)
    if s.n:
        np.testing.assert_array_equal(w, s.get(3))
    return pre, opre


@pytest.mark.parametrize("gr_shape,go_shape", [((4, 8), (2, 2)), ((6, 20), (3, 2)), ((3, 5), (0, 0))])
def test_granule_golgi_rules(ref_data, gr_shape, go_shape):
    gr = gr_grid(ref_data, *gr_shape)
    built = cbrkc.syn_build(cbrkc.SYN_MFGR, *gr_shape, seed=3)
    pre, opre = same(Syn(L().or_mfgr_create(gr.n, 3)), built)
    np.testing.assert_array_equal(pre, opre)
    if go_shape[0] * go_shape[1] == 0:
        assert len(cbrkc.syn_build(cbrkc.SYN_GRGO, *gr_shape, *go_shape)[0]) == 0
        assert len(cbrkc.syn_build(cbrkc.SYN_GOGR, *go_shape, *gr_shape)[0]) == 0
        return
    go = go_grid(ref_data, *go_shape)
    pre, opre = same(Syn(L().or_grgo_create(gr.h, go.h)), cbrkc.syn_build(cbrkc.SYN_GRGO, *gr_shape, *go_shape))
    np.testing.assert_array_equal(pre, opre)
    pre, opre = same(Syn(L().or_gogr_create(go.h, gr.h)), cbrkc.syn_build(cbrkc.SYN_GOGR, *go_shape, *gr_shape))
    np.testing.assert_array_equal(pre, opre)


@pytest.mark.parametrize("seed", [0, 11, 2 ** 40 + 7])
def test_granule_purkinje_rule(ref_data, seed):
    gr = gr_grid(ref_data, 8, 4)
    pk = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 2, 2, v0=[-65.0] * 4, libc_rand=False)
    x = pkj_x(ref_data)
    built = cbrkc.syn_build(cbrkc.SYN_GRPKJ, 8, 4, 2, 2, x1600=x, seed=seed)
    pre, opre = same(Syn(L().or_grpkj_create(gr.h, pk.h, orc._dp(x), seed)), built)
    np.testing.assert_array_equal(pre, opre)
    built = cbrkc.syn_build(cbrkc.SYN_MLIPKJ, 2, 2, 8, 4, seed=seed)
    same(Syn(L().or_mlipkj_create(4, seed)), built)
    assert len(cbrkc.syn_build(cbrkc.SYN_MLIPKJ, 2, 2, 0, 0)[0]) == 0


def test_build_errors():
    with pytest.raises(cbrkc.CbError):
        cbrkc.syn_build(7, 1, 1)
    with pytest.raises(cbrkc.CbError):
        cbrkc.syn_build(cbrkc.SYN_GRPKJ, 2, 2, 1, 1)      # the location file's x column is needed
