"""Failure paths of the RKC engines through the C-ABI (GPU).

* A NaN in a cell's state: the reference's main stops on a NaN voltage
  (main.cu:200, 228, 257, 287, `isnan ( memv_test_xx [ 1 ] )`); here
  cb_pop_run / cb_pop_advance return CB_E_SOLVER with a message naming the
  state (and, for the per-cell engines, the cell) instead of stepping NaNs.
* The step-size floor: when a rejected step would go below hmin the reference
  sets idid = 4 and falls through to the accept branch (solve_rkc.cu:1003-1010),
  so the run goes on with the step it attempted.  With hmin raised to H_MAX
  (cb_pop_config.hmin_ms) every step is clamped to [hmin, hmax] = 0.125 ms:
  each macro step is exactly one accepted step per controller, whatever its
  error; rejections that fall below hmin are counted and then accepted, and
  the run ends without an error.
"""
import numpy as np
import pytest

import cbrkc

CB_E_SOLVER = -4


@pytest.fixture(scope="module")
def morph2(tmp_path_factory):
    d = tmp_path_factory.mktemp("fail")
    path = str(d / "pkj2.txt")
    cbrkc.synth_pkj_morph(path, seed=7, ncell=2)
    return path


@pytest.mark.gpu
@pytest.mark.parametrize("engine", [cbrkc.ENGINE_CELL_MIXED, cbrkc.ENGINE_CELL_FP64, cbrkc.ENGINE_GLOBAL_FP64],
                         ids=["mixed", "cell_fp64", "global_fp64"])
def test_nan_state_stops_the_run(morph2, engine):
    print(f"NaN initial voltage in cell 1, engine {engine}", flush=True)
    p = cbrkc.Population(cbrkc.PKJ, 2, 1, engine=engine, morph=morph2, v0=[-68.0, float("nan")],
                         stim=(1599, 0.0, 1e9, 0.0), **cbrkc.STABLE)
    with pytest.raises(cbrkc.CbError) as e:
        p.run(2, 0.0, 1599)
    msg = str(e.value)
    assert "non-finite" in msg, msg
    if engine != cbrkc.ENGINE_GLOBAL_FP64:
        assert "cell 1" in msg, msg


@pytest.mark.gpu
@pytest.mark.parametrize("engine", [cbrkc.ENGINE_CELL_MIXED, cbrkc.ENGINE_GLOBAL_FP64], ids=["mixed", "global_fp64"])
def test_steps_below_hmin_fall_through_as_accepted(morph2, engine):
    print(f"hmin = H_MAX, engine {engine}", flush=True)
    p = cbrkc.Population(cbrkc.PKJ, 2, 1, engine=engine, morph=morph2, v0=[-68.0, -67.0], hmin_ms=0.125,
                         stim=(1599, 0.0, 1e9, 0.5e-3), **cbrkc.STABLE)
    s0 = p.stats()
    out = p.run(6, 0.0, 1599)
    s1 = p.stats()
    assert np.all(np.isfinite(out))
    ncontrollers = 1 if engine == cbrkc.ENGINE_GLOBAL_FP64 else 2
    assert s1["steps"] - s0["steps"] == 6 * ncontrollers, (s0, s1)
    # some 0.125 ms steps miss the 1e-6 tolerance: rejected, then accepted at hmin
    assert s1["rejects"] - s0["rejects"] > 0, (s0, s1)
