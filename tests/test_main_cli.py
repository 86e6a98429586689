"""The drop-in `main` (main.cu:35-355): argv handling and output files.

CPU tests cover the argument checks that return before any device work;
the GPU test runs one Purkinje cell of the reference morphology for 2 ms and
compares pkj_v_by_RKC.csv with the oracle (libc rand() seeded with 0, as
main.cu:37 does, gives both the same initial voltage).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAIN = os.path.join(ROOT, "cerebellum-neuron-model-simulation-using-rkc_amd", "main")


def run_main(args, cwd, timeout=120):
    return subprocess.run([MAIN] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(MAIN):
        pytest.fail(f"{MAIN} is not built (run __graft_entry__.build())")
    return MAIN


def test_unknown_solver_name(built, tmp_path):
    r = run_main(["XX", "0", "0", "RKC", "0", "0", "RKC", "1", "1", "RKC", "0", "0"], tmp_path)
    assert r.returncode == 0
    assert "Error!! -> set solver name in argv [ 1 ]" in r.stdout
    r = run_main(["RKC", "0", "0", "RKC", "0", "0", "RKC", "1", "1", "CQ", "0", "0"], tmp_path)
    assert "Error!! -> set solver name in argv [ 10 ]" in r.stdout


def test_prefix_match_and_output_files(built, tmp_path):
    # "RKCx" matches RKC and "BEx" BE by prefix; the files are named after the
    # matched solver; empty populations touch no device
    r = run_main(["RKCx", "0", "0", "BEx", "0", "0", "CN", "0", "0", "RKC", "0", "0", "--ts", "1"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    for name in ("gr_v_by_RKC.csv", "go_raster_by_BE.csv", "pkj_v_by_CN.csv", "io_raster_by_RKC.csv"):
        assert (tmp_path / name).exists(), name
    # comptime.csv echoes the argv solver names, as output.cu does
    assert (tmp_path / "comptime.csv").read_text().startswith("Gr,0,RKCx,Go,0,BEx,Pkj,0,CN,Io,0,RKC,CompTime,")


def test_usage_without_arguments(built, tmp_path):
    r = run_main([], tmp_path)
    assert r.returncode == 0 and "usage:" in r.stdout


@pytest.mark.gpu
def test_pkj_voltage_rows_match_oracle(built, tmp_path, ref_data):
    shutil.copy(os.path.join(ref_data, "Pkj.txt"), tmp_path / "Pkj.txt")
    print("main: 1 PKJ cell, RKC, 2 ms", flush=True)
    r = run_main(["RKC", "0", "0", "RKC", "0", "0", "RKC", "1", "1", "RKC", "0", "0", "--ts", "2", "--chunk", "5"],
                 tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [l.rstrip("\n").rstrip(",").split(",") for l in open(tmp_path / "pkj_v_by_RKC.csv")]
    got = np.array([[float(x) for x in row] for row in rows])
    assert got.shape == (16, 2)
    np.testing.assert_allclose(got[:, 0], 0.125 * np.arange(1, 17))
    o = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, libc_rand=True, seed=0)
    ref = []
    for k in range(16):
        o.set_current(k * 0.125)
        o.advance()
        ref.append(o.get(orc.VNEW)[1599])
    np.testing.assert_allclose(got[:, 1], ref, rtol=0, atol=2e-6)   # "%lf" keeps 6 decimals
    assert (tmp_path / "comptime.csv").read_text().startswith("Gr,0,RKC,Go,0,RKC,Pkj,1,RKC,Io,0,RKC,CompTime,")
    assert (tmp_path / "pkj_raster_by_RKC.csv").read_text() == ""


@pytest.mark.gpu
def test_io_raster_file_against_the_references(built, tmp_path, ref_data):
    """./main RKC 0 0 RKC 0 0 RKC 0 0 RKC 8 8 writes io_raster_by_RKC.csv in the
    reference's format with the reference's spikes: every cell once, within
    one macro step of the shipped file's time (run to 810 ms; the pulse is at
    800-820 ms).  Exact times are pinned by the oracle (test_oracle_io.py);
    the grid turns last-bit exp differences into one-step shifts."""
    shutil.copy(os.path.join(ref_data, "IO.txt"), tmp_path / "IO.txt")
    print("main: IO 8x8, RKC, 810 ms", flush=True)
    r = run_main(["RKC", "0", "0", "RKC", "0", "0", "RKC", "0", "0", "RKC", "8", "8", "--ts", "810"], tmp_path,
                 timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    lines = (tmp_path / "io_raster_by_RKC.csv").read_text().splitlines()
    for line in lines:
        t, c = line.split(",")
        assert t == "%f" % float(t) and c == str(int(c))
    got = {int(c): float(t) for t, c in (l.split(",") for l in lines)}
    ref = [l.strip().split(",") for l in open(os.path.join(ref_data, "io_raster_by_RKC.csv"))]
    assert len(lines) == len(ref) == 64 and len(got) == 64
    assert max(abs(got[int(c)] - float(t)) for t, c in ref) <= 0.125
    rows = (tmp_path / "io_v_by_RKC.csv").read_text().splitlines()
    assert len(rows) == 6480 and rows[0].count(",") == 1 + 64 * 3


@pytest.mark.gpu
def test_go_rows_and_raster_against_oracle(built, tmp_path, ref_data):
    """./main RKC 0 0 RKC 2 2 RKC 0 0 RKC 0 0: go_v_by_RKC.csv holds t and the
    soma (compartment 30) of every cell per macro step, go_raster_by_RKC.csv
    its upward crossings of +10 mV; libc rand() seeded 0 gives the oracle the
    same initial voltages.  Spike counts match the oracle, times within two
    macro steps (the oracle's own spread under a one-ulp change of its
    inputs, tests/test_oracle_go.py::test_noise_floor_of_the_grid_workload)."""
    shutil.copy(os.path.join(ref_data, "Go_131.txt"), tmp_path / "Go_131.txt")
    print("main: GO 2x2, RKC, 250 ms", flush=True)
    r = run_main(["RKC", "0", "0", "RKC", "2", "2", "RKC", "0", "0", "RKC", "0", "0", "--ts", "250"], tmp_path,
                 timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    rows = [l.rstrip("\n").rstrip(",").split(",") for l in open(tmp_path / "go_v_by_RKC.csv")]
    got = np.array([[float(x) for x in row] for row in rows])
    assert got.shape == (2000, 5)
    from test_oracle_go import go_pop
    o = go_pop(ref_data, 2, 2, seed=0)
    ref = []
    for k in range(2000):
        o.set_current(k * 0.125)
        o.advance()
        ref.append(o.get(orc.VNEW)[30::131])
    ref = np.array(ref)
    np.testing.assert_allclose(got[0, 1:], ref[0], rtol=0, atol=2e-6)
    lines = (tmp_path / "go_raster_by_RKC.csv").read_text().splitlines()
    for line in lines:
        t, c = line.split(",")
        assert t == "%f" % float(t) and c == str(int(c))
    spikes = sorted((int(c), float(t)) for t, c in (l.split(",") for l in lines))
    want = []
    prev = np.full(4, 100.0)
    for k in range(2000):
        for c in np.nonzero((ref[k] > 10.0) & (prev <= 10.0))[0]:
            want.append((int(c), 0.125 * (k + 1)))
        prev = ref[k]
    want.sort()
    assert len(spikes) == len(want) >= 4
    assert all(a[0] == b[0] and abs(a[1] - b[1]) <= 0.25 for a, b in zip(spikes, want))


@pytest.mark.gpu
def test_gr_rows_against_oracle(built, tmp_path, ref_data):
    """./main RKC 1 1 RKC 0 0 RKC 0 0 RKC 0 0: gr_v_by_RKC.csv holds t and
    compartment 16 per macro step; libc rand() seeded 0 gives the oracle the
    same initial voltage (first 2 ms, before the 50 ms pulse)."""
    shutil.copy(os.path.join(ref_data, "Gr.txt"), tmp_path / "Gr.txt")
    print("main: GR 1x1, RKC, 2 ms", flush=True)
    r = run_main(["RKC", "1", "1", "RKC", "0", "0", "RKC", "0", "0", "RKC", "0", "0", "--ts", "2"], tmp_path,
                 timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    rows = [l.rstrip("\n").rstrip(",").split(",") for l in open(tmp_path / "gr_v_by_RKC.csv")]
    got = np.array([[float(x) for x in row] for row in rows])
    assert got.shape == (16, 2)
    from test_oracle_gr import gr_pop
    o = gr_pop(ref_data, 1, 1, seed=0)
    ref = []
    for k in range(16):
        o.set_current(k * 0.125)
        o.advance()
        ref.append(o.get(orc.VNEW)[16])
    np.testing.assert_allclose(got[:, 1], ref, rtol=0, atol=2e-6)
    assert (tmp_path / "comptime.csv").read_text().startswith("Gr,1,RKC,Go,0,RKC,Pkj,0,RKC,Io,0,RKC,CompTime,")


@pytest.mark.gpu
def test_cn_pkj_rows_against_oracle(built, tmp_path, ref_data):
    """./main CN 0 0 CN 0 0 CN 1 1 CN 0 0 (pkj1.sh's CN runs): pkj_v_by_CN.csv
    from the CN engine, five 0.025 ms CN steps per row, against the oracle's CN
    path from the same libc rand() initial voltage."""
    shutil.copy(os.path.join(ref_data, "Pkj.txt"), tmp_path / "Pkj.txt")
    print("main: PKJ 1x1, CN, 2 ms", flush=True)
    r = run_main(["CN", "0", "0", "CN", "0", "0", "CN", "1", "1", "CN", "0", "0", "--ts", "2"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [l.rstrip("\n").rstrip(",").split(",") for l in open(tmp_path / "pkj_v_by_CN.csv")]
    got = np.array([[float(x) for x in row] for row in rows])
    assert got.shape == (16, 2)
    from test_oracle_cn import L as LC
    o = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, libc_rand=True, seed=0)
    assert LC().or_cn_init(o.h) == 0
    ref = []
    for k in range(16):
        LC().or_cn_advance(o.h, k * 0.125)
        ref.append(o.get(orc.V)[1599])
    np.testing.assert_allclose(got[:, 1], ref, rtol=0, atol=2e-6)
    assert (tmp_path / "comptime.csv").read_text().startswith("Gr,0,CN,Go,0,CN,Pkj,1,CN,Io,0,CN,CompTime,")


@pytest.mark.gpu
def test_be_io_rows_against_oracle(built, tmp_path, ref_data):
    """./main BE 0 0 BE 0 0 BE 0 0 BE 2 2: io_v_by_BE.csv from the BE engine
    (five 0.025 ms BE steps per row, BiCG with the gap currents on the
    right-hand side) against the oracle's BE path from the same libc rand()
    initial voltages, all three compartments of every cell, 10 ms."""
    shutil.copy(os.path.join(ref_data, "IO.txt"), tmp_path / "IO.txt")
    print("main: IO 2x2, BE, 10 ms", flush=True)
    r = run_main(["BE", "0", "0", "BE", "0", "0", "BE", "0", "0", "BE", "2", "2", "--ts", "10"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [l.rstrip("\n").rstrip(",").split(",") for l in open(tmp_path / "io_v_by_BE.csv")]
    got = np.array([[float(x) for x in row] for row in rows])
    assert got.shape == (80, 1 + 4 * 3)
    np.testing.assert_allclose(got[:, 0], 0.125 * np.arange(1, 81), rtol=0, atol=1e-9)
    from test_oracle_be import L as LB
    from test_oracle_io import io_pop
    o = io_pop(ref_data, 2, 2, seed=0)
    assert LB().or_be_init(o.h) == 0
    ref = []
    for k in range(80):
        LB().or_be_advance(o.h, k * 0.125)
        ref.append(o.get(orc.V))
    # the file holds "%lf" (6 decimals): 5e-7 of print rounding on top of the engine's 1e-9
    np.testing.assert_allclose(got[:, 1:], np.array(ref), rtol=0, atol=6e-7)
    assert (tmp_path / "comptime.csv").read_text().startswith("Gr,0,BE,Go,0,BE,Pkj,0,BE,Io,4,BE,CompTime,")


@pytest.mark.gpu
def test_network_rows_against_oracle(built, tmp_path, ref_data):
    """./main CN 4 4 CN 1 1 CN 1 1 CN 0 0 --network: the populations draw their
    initial voltages from libc rand() seeded 0 (GR, GO, PKJ), the mossy-fibre,
    granule-Purkinje and inhibitory sets take the next three draws as seeds
    (main.cu:139-145), and the synapses update after every macro step.
    Against the oracle network built the same way: the first granule row and
    every Golgi row at 6e-7 (%lf), every Purkinje row at 2e-6 (the bar of the
    single-Purkinje CN run above: that solve also stops at the 100-iteration
    cap, observed 9.9e-7 here); the granule rows within 0.1 mV,
    as the 16-cell granule CN solve stops at the 100-iteration CG cap and its
    truncated iterate drifts (the oracle's own spread on this network under a
    1e-12 mV change of the granule initial voltages reaches 1.7e-3 mV after
    one macro step and 3.6 mV around the mossy-fibre events).  MF_RASTER.csv
    and MLI_RASTER.csv are created empty."""
    import ctypes
    for f in ("Gr.txt", "Go_131.txt", "Pkj.txt", "Pkj_GJ.csv", "Pkj_location_info.csv"):
        shutil.copy(os.path.join(ref_data, f), tmp_path / f)
    print("main: GR 4x4, GO 1x1, PKJ 1x1, CN, --network, 20 ms", flush=True)
    r = run_main(["CN", "4", "4", "CN", "1", "1", "CN", "1", "1", "CN", "0", "0", "--ts", "20", "--network"],
                 tmp_path, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert "# of mfgr = 64" in r.stdout and "# of grpkj = 400" in r.stdout
    for f in ("MF_RASTER.csv", "MLI_RASTER.csv"):
        assert (tmp_path / f).read_text() == ""
    rows = lambda name: np.array([[float(x) for x in l.rstrip("\n").rstrip(",").split(",")]
                                  for l in open(tmp_path / name)])
    got_gr, got_go, got_pk = rows("gr_v_by_CN.csv"), rows("go_v_by_CN.csv"), rows("pkj_v_by_CN.csv")
    assert got_gr.shape == (160, 17) and got_go.shape == (160, 2) and got_pk.shape == (160, 2)

    from test_oracle_cn import L as LC
    from test_oracle_syn import L as LS, Syn, pkj_x
    from test_oracle_gr import L as LR
    from test_oracle_go import L as LG
    lib = orc.lib()
    lib.rand.restype = ctypes.c_int
    lib.srand(0)
    ogr = orc.Pop(LR().or_gr_create(os.path.join(ref_data, "Gr.txt").encode(), 4, 4, None, 1e-9, 1e-9, 1), 16, 578)
    ogo = orc.Pop(LG().or_go_create(os.path.join(ref_data, "Go_131.txt").encode(), 1, 1, None, 1e-6, 1e-6, 1), 1, 131)
    opk = orc.Pop(lib.or_pkj_create(os.path.join(ref_data, "Pkj.txt").encode(), None, 1, 1, None, 1e-6, 1e-6, 1),
                  1, 1600)
    seeds = [lib.rand(), lib.rand(), lib.rand()]
    mf = Syn(LS().or_mfgr_create(16, seeds[0]))
    gg = Syn(LS().or_gogr_create(ogo.h, ogr.h))
    rg = Syn(LS().or_grgo_create(ogr.h, ogo.h))
    rp = Syn(LS().or_grpkj_create(ogr.h, opk.h, orc._dp(pkj_x(ref_data)), seeds[1]))
    for o, s in ((ogr, mf), (ogr, gg), (ogo, rg), (opk, rp)):
        assert LS().or_pop_attach_syn(o.h, s.h) == 0
    for o in (ogr, ogo, opk):
        o.set_stim(0, 0.0, 0.0, np.zeros(o.n))    # network_sim: no xx_set_current, i_ext stays 0
        assert LC().or_cn_init(o.h) == 0
    lib.or_net_update.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_double]
    ref_gr, ref_go, ref_pk = [], [], []
    for k in range(160):
        for o in (ogr, ogo, opk):
            LC().or_cn_advance(o.h, k * 0.125)
        lib.or_net_update(mf.h, gg.h, rg.h, rp.h, k * 0.125)
        ref_gr.append(ogr.get(orc.V)[16::578])
        ref_go.append(ogo.get(orc.V)[30])
        ref_pk.append(opk.get(orc.V)[1599])
    assert np.max(mf.get(4)) > 0.0                   # mossy-fibre events happened (three in 20 ms)
    np.testing.assert_allclose(got_gr[0, 1:], ref_gr[0], rtol=0, atol=6e-7)
    print(f"granule rows: max |main - oracle| = {np.max(np.abs(got_gr[:, 1:] - np.array(ref_gr))):.3e} mV", flush=True)
    np.testing.assert_allclose(got_gr[:, 1:], np.array(ref_gr), rtol=0, atol=0.1)
    np.testing.assert_allclose(got_go[:, 1], ref_go, rtol=0, atol=6e-7)
    np.testing.assert_allclose(got_pk[:, 1], ref_pk, rtol=0, atol=2e-6)


@pytest.mark.gpu
def test_tree_option_rows_against_oracle(built, tmp_path, ref_data):
    """./main CN 0 0 CN 0 0 CN 1 1 CN 0 0 --tree: the Purkinje cell's CN steps
    solved exactly per cell, against the oracle's exact solve from the same
    libc rand() initial voltage (%lf rows, 2e-6); a Purkinje grid coupled by
    Pkj_GJ.csv is refused with the library's message."""
    import ctypes
    for f in ("Pkj.txt", "Pkj_GJ.csv"):
        shutil.copy(os.path.join(ref_data, f), tmp_path / f)
    print("main: PKJ 1x1, CN, --tree, 2 ms", flush=True)
    r = run_main(["CN", "0", "0", "CN", "0", "0", "CN", "1", "1", "CN", "0", "0", "--ts", "2", "--tree"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [l.rstrip("\n").rstrip(",").split(",") for l in open(tmp_path / "pkj_v_by_CN.csv")]
    got = np.array([[float(x) for x in row] for row in rows])
    assert got.shape == (16, 2)
    from test_oracle_direct import L as LD
    o = orc.Pop.pkj(os.path.join(ref_data, "Pkj.txt"), 1, 1, libc_rand=True, seed=0)
    assert LD().or_cn_init(o.h) == 0 and LD().or_cn_set_direct(o.h, 1) == 0
    ref = []
    for k in range(16):
        LD().or_cn_advance(o.h, k * 0.125)
        ref.append(o.get(orc.V)[1599])
    np.testing.assert_allclose(got[:, 1], ref, rtol=0, atol=2e-6)
    print("main: PKJ 2x2 with Pkj_GJ.csv, CN, --tree (refused)", flush=True)
    r = run_main(["CN", "0", "0", "CN", "0", "0", "CN", "2", "2", "CN", "0", "0", "--ts", "1", "--tree"], tmp_path)
    assert r.returncode != 0 and "gap junctions" in r.stderr, r.stdout[-500:] + r.stderr


@pytest.mark.gpu
def test_network_with_the_per_cell_purkinje_engine(built, tmp_path, ref_data):
    """./main CN 4 4 CN 1 1 RKC 1 1 CN 0 0 --network --engine cell against the
    same command on the global engine: the Purkinje cell (per-cell FP64 RKC
    engine, one macro step per call) receives the same granule -> Purkinje
    input, so its rows agree within the per-cell engines' synaptic bar
    (0.03 mV); the granule and Golgi rows (CN, the same engine in both runs)
    are identical.  With --network no step current is injected
    (network_sim), so the granule cells fire on mossy-fibre input alone, from
    about 15 ms on; 45 ms, because the reference's Purkinje RKC scheme itself
    turns non-finite on this network at 62.4 ms (the oracle run of the same
    network, without the synthetic-tree safeguards); a run without --network
    shows the input reaches the Purkinje soma."""
    outs = {}
    for eng, net in (("global", True), ("cell", True), ("global", False)):
        d = tmp_path / f"{eng}_{net}"
        d.mkdir()
        for f in ("Gr.txt", "Go_131.txt", "Pkj.txt", "Pkj_GJ.csv", "Pkj_location_info.csv"):
            shutil.copy(os.path.join(ref_data, f), d / f)
        print(f"main: GR 4x4 CN, GO 1x1 CN, PKJ 1x1 RKC --engine {eng}, network {net}, 45 ms", flush=True)
        r = run_main(["CN", "4", "4", "CN", "1", "1", "RKC", "1", "1", "CN", "0", "0", "--ts", "45", "--engine", eng]
                     + (["--network"] if net else []), d, timeout=600)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr
        rows = lambda name: np.array([[float(x) for x in l.rstrip("\n").rstrip(",").split(",")]
                                      for l in open(d / name)])
        outs[(eng, net)] = (rows("gr_v_by_CN.csv"), rows("go_v_by_CN.csv"), rows("pkj_v_by_RKC.csv"))
    (gg, og, pg), (gc, oc, pc) = outs[("global", True)], outs[("cell", True)]
    np.testing.assert_array_equal(gg, gc)
    np.testing.assert_array_equal(og, oc)
    d = np.max(np.abs(pg[:, 1] - pc[:, 1]))
    effect = np.max(np.abs(pg[:, 1] - outs[("global", False)][2][:, 1]))
    print(f"Purkinje rows: max |cell - global| = {d:.3e} mV; synaptic effect {effect:.3e} mV", flush=True)
    assert pg.shape == pc.shape == (360, 2) and d <= 0.03 and effect > 1e-3
