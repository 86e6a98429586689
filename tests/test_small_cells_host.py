"""The batched engine's small-cell body (granule chains, Golgi trees, olive
cells on G lanes per cell) built for the host from the same source
(tests/host_emu: the device functions and the engine's host builders taken
verbatim from engine_cell.hip, one host thread per lane, AddressSanitizer and
UBSan on), no GPU:

* against the oracle cell by cell, the same case and bars as the GPU test of
  the mix (3 granule chains, 2 Golgi trees, 4 olive cells joined by a random
  junction list, 20 ms): every cell is the reference's RKC on a population
  of one (or_gr_create_comp / or_go_create / or_io_create), the first
  interval to rounding for granule and Golgi cells, the olive bar measured
  on the oracle's own one-ulp spread;
* every lane count the autotuner may pick runs in bounds on a larger mix.

The device build of this engine is switched off (cb_mix_create refuses)
while a GPU fault of its first device run is unexplained; these tests pin
the algorithm and the indexing of the code that would run there.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import cbrkc
import test_gpu_mix as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GR, GO, IO = cbrkc.GR, cbrkc.GO, cbrkc.IO


@pytest.fixture(scope="session")
def emu_bin(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("no host C++ compiler")
    out = tmp_path_factory.mktemp("host_emu")
    subprocess.run(["bash", os.path.join(ROOT, "tests", "host_emu", "build.sh"), str(out)], check=True,
                   capture_output=True, text=True)
    return str(out / "emu")


def run_emu(emu, tmp, pops, gj, nint, lanes=(0, 0, 0, 0)):
    """pops: {type: (morph, v0, amp)}; returns (rec [nint, cells in type
    order], spike words [nint, words])"""
    fmt = lambda a: " ".join(repr(float(x)) for x in a)  # noqa: E731
    cfg, out = os.path.join(tmp, "cfg.txt"), os.path.join(tmp, "out.bin")
    with open(cfg, "w") as f:
        for t, (morph, v0, amp) in pops.items():
            f.write(f"POP {t} {morph} {len(v0)} {fmt(v0)} {fmt(amp)}\n")
        if gj:
            f.write(f"GJ {gj}\n")
        f.write(f"NINT {nint}\nOUT {out}\nLANES {' '.join(str(x) for x in lanes)}\n")
    r = subprocess.run([emu, cfg], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    n = {t: len(p[1]) for t, p in pops.items()}
    tot = sum(n.values())
    words = sum((k + 31) // 32 for k in n.values())
    raw = open(out, "rb").read()
    rec = np.frombuffer(raw[:8 * nint * tot], dtype=np.float64).reshape(nint, tot)
    spk = np.frombuffer(raw[8 * nint * tot:], dtype=np.uint32).reshape(nint, words)
    return rec, spk


def test_small_cells_against_the_oracle(emu_bin, ref_data, tmp_path):
    lib = M._L()
    steps = 160
    grf = str(tmp_path / "gr_chains.txt")
    cbrkc.synth_gr_chains(grf, os.path.join(ref_data, "Gr.txt"), 7, 3)
    gr_v0, gr_amp = np.array([-70.0, -69.0, -71.0]), np.array([0.009e-3, 0.011e-3, 0.013e-3])
    gof = str(tmp_path / "go_trees.txt")
    cbrkc.synth_go_trees(gof, os.path.join(ref_data, "Go_131.txt"), 8, 2)
    go_v0, go_amp = np.array([-60.0, -61.0]), np.array([0.2e-3, 0.3e-3])
    gjf = str(tmp_path / "io_gj.csv")
    cbrkc.synth_io_gj(gjf, 9, 4, k=2, g_min_nS=0.5, g_max_nS=1.5)
    pairs = [(int(a), int(b), float(g)) for a, b, g in (l.strip().split(",") for l in open(gjf) if l.strip())]
    iof = os.path.join(ref_data, "IO.txt")
    io_v0, io_amp = np.array([-60.0, -58.0, -62.0, -59.0]), np.array([5.5, 4.5, 6.0, 5.0])
    rec, spk = run_emu(emu_bin, str(tmp_path), {GR: (grf, gr_v0, gr_amp), GO: (gof, go_v0, go_amp),
                                                IO: (iof, io_v0, io_amp)}, gjf, steps)
    import orc
    ref = np.zeros((steps, 3))
    for c, f in enumerate(M._blocks(grf, 4, 3, str(tmp_path), "gr")):
        o = orc.Pop(lib.or_gr_create_comp(f.encode(), 4, 1, 1, orc._dp(gr_v0[c:c + 1]), 1e-9, 1e-9, 0), 1, 4)
        o.set_stim(1, 0.0, 1e300, np.array([gr_amp[c]]))
        for k in range(steps):
            o.set_current(k * 0.125)
            o.advance()
            ref[k, c] = o.get(orc.VNEW)[3]
    M._check(ref, rec[:, 0:3], "granule")
    ref = np.zeros((steps, 2))
    for c, f in enumerate(M._blocks(gof, 131, 2, str(tmp_path), "go")):
        o = orc.Pop(lib.or_go_create(f.encode(), 1, 1, orc._dp(go_v0[c:c + 1]), 1e-6, 1e-6, 0), 1, 131)
        o.set_stim(30, 0.0, 1e300, np.array([go_amp[c]]))
        for k in range(steps):
            o.set_current(k * 0.125)
            o.advance()
            ref[k, c] = o.get(orc.VNEW)[30]
    M._check(ref, rec[:, 3:5], "Golgi")
    ref = M._olive_oracle(lib, iof, io_v0, io_amp, pairs, steps)
    nudged = M._olive_oracle(lib, iof, np.nextafter(io_v0, 0.0), io_amp, pairs, steps)
    M._check(ref, rec[:, 5:9], "olive", floor=M._floor(ref, nudged))
    # spike bits: upward crossings of 10 mV against the previous interval
    # (100 mV before the first), per type in its own words
    for lo, hi, w0 in ((0, 3, 0), (3, 5, 1), (5, 9, 2)):
        tr = rec[:, lo:hi]
        prev = np.vstack([np.full((1, hi - lo), 100.0), tr[:-1]])
        c = np.arange(hi - lo)
        bits = ((spk[:, w0 + c // 32] >> (c % 32).astype(np.uint32)) & 1).astype(bool)
        np.testing.assert_array_equal(bits, (tr > 10.0) & (prev <= 10.0))
    assert spk[:, 0].any() and spk[:, 1].any()


@pytest.mark.parametrize("lanes", [(2, 64, 0, 2), (4, 64, 0, 4)])
def test_every_lane_count_in_bounds(emu_bin, ref_data, tmp_path, lanes):
    grf = str(tmp_path / "gr.txt")
    cbrkc.synth_gr_chains(grf, os.path.join(ref_data, "Gr.txt"), 3, 40)
    gof = str(tmp_path / "go.txt")
    cbrkc.synth_go_trees(gof, os.path.join(ref_data, "Go_131.txt"), 4, 3)
    gjf = str(tmp_path / "gj.csv")
    cbrkc.synth_io_gj(gjf, 6, 20, k=6)
    pops = {GR: (grf, np.linspace(-72, -68, 40), np.linspace(0.008e-3, 0.016e-3, 40)),
            GO: (gof, np.linspace(-62, -58, 3), np.linspace(0.1e-3, 0.3e-3, 3)),
            IO: (os.path.join(ref_data, "IO.txt"), np.linspace(-62, -58, 20), np.linspace(3.0, 6.0, 20))}
    base, _ = run_emu(emu_bin, str(tmp_path), pops, gjf, 16)
    other, _ = run_emu(emu_bin, str(tmp_path), pops, gjf, 16, lanes)
    assert np.all(np.isfinite(other))
    # another lane count only reorders the norms' sums, which moves an
    # adaptive step now and then (the olive and Golgi cells at rtol 1e-6)
    np.testing.assert_allclose(other, base, atol=0.01, rtol=0)
