"""Non-square grids on the GPU against the oracle: the grid-neighbour rules
(the olive's gap-junction partners with the reference's id / ny row index,
the Golgi and Purkinje gap-junction pairs) on nx != ny, where a partner the
rule names outside the population is absent on both sides.

* olive 2 x 3 and 3 x 2, RKC: the first macro step within 1e-9 mV, then the
  same spikes per cell within one macro step over 30 ms;
* olive 3 x 2 with BE and CN (BiCG, gap currents on the right-hand side):
  the first output interval within 1e-12 / 1e-9 mV;
* Golgi 3 x 2 with CN (gap junctions inside the CG matrix): the first output
  interval within 1e-9 mV;
* Purkinje 3 x 2 with Pkj_GJ.csv, RKC: one macro step within 1e-9 mV.
"""
import os

import numpy as np
import pytest

import cbrkc
import orc
from test_gpu_io import crossings
from test_oracle_io import io_pop
from test_oracle_go import go_pop
from test_oracle_direct import L as LD

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(2, 3), (3, 2)])
def test_io_ragged_grid_rkc(ref_data, shape):
    nx, ny = shape
    n = nx * ny
    v0 = np.linspace(-64.0, -56.0, n)
    p = cbrkc.Population(cbrkc.IO, nx, ny, morph=os.path.join(ref_data, "IO.txt"), v0=v0, stim=(0, 2.0, 12.0, 5.5))
    o = io_pop(ref_data, nx, ny, v0=v0)
    orc.lib().or_set_stim(o.h, 0, 2.0, 12.0, orc._dp(np.full(n, 5.5)))
    print(f"IO {nx}x{ny}, RKC, 240 macro steps", flush=True)
    tg, to = [], []
    for k in range(240):
        p.set_current(k * 0.125)
        p.advance()
        o.set_current(k * 0.125)
        o.advance()
        vg, vo = p.read(cbrkc.F_VNEW), o.get(orc.VNEW)
        if k == 0:
            d = np.max(np.abs(vg - vo))
            print(f"first macro step: {d:.3e} mV", flush=True)
            assert d <= 1e-9
        tg.append(vg[2::3])
        to.append(vo[2::3])
    sg, so = crossings(np.array(tg)), crossings(np.array(to))
    print(f"spikes GPU {len(sg)} oracle {len(so)}", flush=True)
    assert len(so) > 0 and sorted(c for _, c in sg) == sorted(c for _, c in so)
    for c in range(n):
        a = [t for t, cc in sg if cc == c]
        b = [t for t, cc in so if cc == c]
        assert np.max(np.abs(np.array(a) - np.array(b)), initial=0.0) <= 0.125


def first_interval(p, o, adv):
    l_dt = 0.0
    while l_dt < 0.125:
        p.set_current(l_dt)
        p.advance()
        l_dt += 0.025
    adv(o.h, 0.0)
    return np.max(np.abs(p.read(cbrkc.F_V) - o.get(orc.V)))


@pytest.mark.parametrize("solver", [cbrkc.CN, cbrkc.BE])
def test_io_ragged_grid_cn_be(ref_data, solver):
    v0 = np.linspace(-64.0, -56.0, 6)
    p = cbrkc.Population(cbrkc.IO, 3, 2, solver=solver, morph=os.path.join(ref_data, "IO.txt"), v0=v0)
    o = io_pop(ref_data, 3, 2, v0=v0)
    assert (LD().or_cn_init(o.h) if solver == cbrkc.CN else LD().or_be_init(o.h)) == 0
    d = first_interval(p, o, LD().or_cn_advance if solver == cbrkc.CN else LD().or_be_advance)
    print(f"IO 3x2 {'CN' if solver == cbrkc.CN else 'BE'}: first interval {d:.3e} mV", flush=True)
    assert d <= (1e-9 if solver == cbrkc.CN else 1e-12)


def test_go_ragged_grid_cn(ref_data):
    v0 = np.linspace(-61.0, -59.0, 6)
    p = cbrkc.Population(cbrkc.GO, 3, 2, solver=cbrkc.CN, morph=os.path.join(ref_data, "Go_131.txt"), v0=v0)
    o = go_pop(ref_data, 3, 2, v0=v0)
    assert LD().or_cn_init(o.h) == 0
    d = first_interval(p, o, LD().or_cn_advance)
    print(f"GO 3x2 CN: first interval {d:.3e} mV", flush=True)
    assert d <= 1e-9


def test_pkj_ragged_grid_with_gap_junctions(ref_data):
    v0 = np.linspace(-66.0, -64.0, 6)
    p = cbrkc.Population(cbrkc.PKJ, 3, 2, morph=os.path.join(ref_data, "Pkj.txt"),
                         gj=os.path.join(ref_data, "Pkj_GJ.csv"), v0=v0)
    o = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 3, 2, v0=v0, libc_rand=False,
                    gj=os.path.join(ref_data, "Pkj_GJ.csv"))
    p.set_current(0.0)
    p.advance()
    o.set_current(0.0)
    o.advance()
    d = np.max(np.abs(p.read(cbrkc.F_VNEW) - o.get(orc.VNEW)))
    print(f"PKJ 3x2 with gap junctions, RKC: first macro step {d:.3e} mV", flush=True)
    assert d <= 1e-9
