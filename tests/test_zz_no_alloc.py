"""cb_pop_run_noalloc (the allocation-free variant of cb_pop_run) against
cb_pop_run, through the C-ABI: two populations built identically, one run
by each entry point, on two calls in a row (the second with fewer macro
steps than the first), for the mixed and the FP64 per-cell engines.  The
same kernels run in both, so recordings and final voltages are bit-equal.
A request whose recording exceeds the scratch array is refused with
CB_E_ARG before anything runs."""
import numpy as np
import pytest

import cbrkc


@pytest.fixture(scope="module")
def morph4(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("noalloc") / "pkj4.txt")
    cbrkc.synth_pkj_morph(path, seed=11, ncell=4)
    return path


@pytest.mark.gpu
@pytest.mark.parametrize("engine", [cbrkc.ENGINE_CELL_MIXED, cbrkc.ENGINE_CELL_FP64], ids=["mixed", "cell_fp64"])
def test_noalloc_run_equals_run(morph4, engine):
    kw = dict(engine=engine, morph=morph4, v0=[-68.0, -67.5, -68.5, -67.0],
              stim=(1599, 0.0, 1e9, np.array([0.3, 0.5, 0.7, 0.9]) * 1e-3),
              poisson=dict(rate_hz=5.0, g_mS=1e-6, tau_ms=2.0, e_mV=0.0, seed=3), **cbrkc.STABLE)
    a = cbrkc.Population(cbrkc.PKJ, 4, 1, **kw)
    b = cbrkc.Population(cbrkc.PKJ, 4, 1, **kw)
    t = 0.0
    for n in (24, 8):
        print(f"engine {engine}: {n} macro steps from t = {t} ms, cb_pop_run vs cb_pop_run_noalloc", flush=True)
        ra = a.run(n, t, 1599)
        rb = b.run_noalloc(n, t, 1599)
        assert ra.shape == rb.shape == (n, 4)
        assert np.all(np.isfinite(ra))
        np.testing.assert_array_equal(ra, rb)
        t += n * cbrkc.H_MAX
    np.testing.assert_array_equal(a.read(cbrkc.F_VNEW), b.read(cbrkc.F_VNEW))
    assert a.stats() == b.stats()
    # no recording: the run still advances the same way
    a.run(2, t, 1599, record=False)
    b.run_noalloc(2, t, 1599, record=False)
    np.testing.assert_array_equal(a.read(cbrkc.F_VNEW), b.read(cbrkc.F_VNEW))


@pytest.mark.gpu
def test_noalloc_refuses_a_recording_beyond_the_scratch(morph4):
    p = cbrkc.Population(cbrkc.PKJ, 4, 1, engine=cbrkc.ENGINE_CELL_MIXED, morph=morph4,
                         v0=[-68.0, -67.5, -68.5, -67.0], **cbrkc.STABLE)
    before = p.read(cbrkc.F_VNEW)
    with pytest.raises(cbrkc.CbError) as e:
        p.run_noalloc((8 << 20) // (8 * 4), 0.0, 1599)   # 8 MiB of recording: more than fits after the counter
    assert "scratch" in str(e.value)
    np.testing.assert_array_equal(p.read(cbrkc.F_VNEW), before)   # nothing ran


@pytest.mark.gpu
def test_noalloc_is_refused_by_the_global_engine(morph4):
    p = cbrkc.Population(cbrkc.PKJ, 4, 1, engine=cbrkc.ENGINE_GLOBAL_FP64, morph=morph4,
                         v0=[-68.0, -67.5, -68.5, -67.0], **cbrkc.STABLE)
    with pytest.raises(cbrkc.CbError) as e:
        p.run_noalloc(1, 0.0, 1599)
    assert "per-cell" in str(e.value)
