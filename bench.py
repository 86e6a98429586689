#!/usr/bin/env python3
"""Benchmark: compartment-steps/s of the batched mixed-precision RKC engine on
1024 synthetic Purkinje cells (1600 compartments each), with the voltage and
spike error against the per-cell FP64 engine on the same run, the plain FP64
population-global RKC beside it, and the CPU oracle on the 9-cell plain
workload.

One bench step = 80 macro steps (10 ms simulated) of all 1024 cells; the
default 20 timed steps are the 200 ms precision workload, run as one
persistent launch after 5 untimed steps (50 ms).  A compartment-step is one compartment advanced over one
macro step (H_MAX = 0.125 ms, the reference's output interval), independent
of the adaptive internal steps.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

`--workload scale` runs the scale-stress comparison instead: 8192 synthetic
Purkinje cells with per-cell diameter factors U(0.5, 2.0), 50 ms, the mixed
engine against the per-cell FP64 RKC engine, the plain FP64 global RKC and
the FP64 Crank-Nicolson engine with the exact per-cell tree solve at several
steps, each with its error against the FP64 RKC run (one JSON line, not the
headline metric).  `--workload rows` times every cell type and solver on the
reference's own morphology files with the oracle beside each;
`--workload network` runs the fixed in-degree network (2048 granule, 512
Golgi, 512 Purkinje cells, 780404 synapses, 1 ms delay, 5 Hz mossy fibres;
100 ms by default, `--net-intervals 8000` for 1 s) with the per-population
time split and the mixed-vs-FP64 Purkinje raster comparison.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "cerebellum-neuron-model-simulation-using-rkc_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

METRIC = "compartment-steps/s @1024 PCs RKC: mixed-FP32 vs FP64, ≤0.1 mV error"
H_MAX = 0.125
COMP = 1600
PRECISION_STEPS = 1600   # the precision workload: 200 ms of macro steps
# Algorithmic work of the persistent per-cell kernel, per unit, counted from
# the method as this design runs it (s = 4 bytes per FP32 value).  The stage
# loop, the error estimate and the A rebuilds run out of registers and LDS,
# so HBM is touched only by:
#   compartment-ion-update (a dendritic compartment: 9 of the 18 gates belong
#     to channels present there): reads the type word, area, leak, rev, V_old,
#     shell, the Poisson conductance and 9 gates (16 x 4 B) and Ca (FP64,
#     8 B); writes 9 gates, rev, V_old, b (12 x 4 B) and Ca (8 B): 128 B;
#     ~200 flop (~10 exp, ~15 reciprocals)
#   compartment-macro-step (Poisson drive): reads the type word and the
#     conductance, writes the conductance: 12 B
#   compartment-launch (launch edges): reads y_n hi/lo, F_0, Y_{m-2}, the type
#     word, b, g_parent, 1/C, topology (9 x 4 B); writes y_n hi/lo, F_0,
#     Y_{m-2} (16 B) and V in FP64 (8 B): 60 B
# On chip, per compartment-RHS (one RHS evaluation fused with its stage
# update): LDS reads of topology, g_parent, 1/C, the own and ~2 neighbour
# D values and their conductances plus one LDS write: ~32 B; ~16 flop.
# Per compartment per accepted step (error estimate): ~5 flop, registers.
BYTES_ION = 128.0
BYTES_MACRO = 12.0
BYTES_LAUNCH = 60.0
LDS_RHS = 32.0
FLOP_RHS = 16.0
FLOP_STEP = 5.0
FLOP_ION = 200.0
PEAK_HBM_GBS = 8000.0
# FP32 vector peak without packed FP32 instructions (which this build does not
# use): 256 CUs x 64 lanes x 2 flop (FMA) x 2.4 GHz
PEAK_FP32_TFLOPS = 78.6
# LDS: 128 B per clock per CU x 256 CUs x 2.4 GHz
PEAK_LDS_GBS = 78643.2


def philox_params(seed, n):
    """Per-cell initial V (-68 + N(0,1)) and soma current U(0.2, 1.0) nA from
    Philox4x32-10 (ctr = (cell, 0, 0, 0), key = (seed, 0x43454c4c))."""
    import cbrkc
    v0 = np.zeros(n)
    amp = np.zeros(n)
    for c in range(n):
        r = cbrkc.philox4x32_10([c, 0, 0, 0], [seed & 0xffffffff, 0x43454C4C])
        u1 = ((r[0] >> 5) * 67108864.0 + (r[1] >> 6)) / 9007199254740992.0
        u2 = ((r[2] >> 5) * 67108864.0 + (r[3] >> 6)) / 9007199254740992.0
        g = math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)
        v0[c] = -68.0 + g
        r2 = cbrkc.philox4x32_10([c, 1, 0, 0], [seed & 0xffffffff, 0x43454C4C])
        amp[c] = (0.2 + 0.8 * (((r2[0] >> 5) * 67108864.0 + (r2[1] >> 6)) / 9007199254740992.0)) * 1e-3
    return v0, amp


def philox_dscale(seed, n):
    """Per-cell diameter factors U(0.5, 2.0) from Philox4x32-10
    (ctr = (cell, 2, 0, 0), key = (seed, 0x43454c4c))."""
    import cbrkc
    d = np.zeros(n)
    for c in range(n):
        r = cbrkc.philox4x32_10([c, 2, 0, 0], [seed & 0xffffffff, 0x43454C4C])
        d[c] = 0.5 + 1.5 * (((r[0] >> 5) * 67108864.0 + (r[1] >> 6)) / 9007199254740992.0)
    return d


def make_pop(cbrkc, engine, morph, n, v0, amp, args):
    return cbrkc.Population(
        cbrkc.PKJ, n, 1, morph=morph, v0=v0, atol=args.tol, rtol=args.tol, stim=(1599, 0.0, 1e9, amp),
        engine=engine, poisson=dict(rate_hz=5.0, g_mS=1e-6, tau_ms=2.0, e_mV=0.0, seed=args.seed),
        **cbrkc.STABLE)


def timed_run(p, steps, chunk, t0, record=True):
    """steps macro steps in launches of `chunk` (<= 0: one launch for all of
    them); returns (trace, wall seconds)."""
    tr = np.zeros((steps, p.n)) if record else None
    done = 0
    t = time.perf_counter()
    while done < steps:
        k = steps - done if chunk <= 0 else min(chunk, steps - done)
        out = p.run(k, t0 + done * H_MAX, 1599, record=record)
        if record:
            tr[done:done + k] = out
        done += k
    return tr, time.perf_counter() - t


def work(s0, s1):
    return {k: s1[k] - s0[k] for k in s0}


def error_report(compare, ref, tr, name, early_ms=25.0):
    """Spike counts, spike-time shifts and subthreshold |dV| of trace `tr`
    against `ref` over the whole run, plus the pointwise |dV| over the first
    `early_ms` of the timed run (before chaotic divergence can grow)."""
    counts, dt, sub = compare(ref, tr)
    k = min(len(ref), int(round(early_ms / H_MAX)))
    return {
        "spike_counts_identical": bool(np.all(counts[:, 0] == counts[:, 1])),
        "spikes_fp64": int(counts[:, 0].sum()), f"spikes_{name}": int(counts[:, 1].sum()),
        "cells_with_count_mismatch": int(np.sum(counts[:, 0] != counts[:, 1])),
        "max_spike_time_diff_ms": float(dt.max()),
        "median_spike_time_diff_ms": float(np.median(dt)),
        "max_subthreshold_abs_dV_mV": float(sub.max()),
        "median_per_cell_subthreshold_abs_dV_mV": float(np.median(sub)),
        "cells_subthreshold_within_0.1mV": int(np.sum(sub <= 0.1)),
        f"max_abs_dV_mV_first_{early_ms:g}ms": float(np.max(np.abs(ref[:k] - tr[:k]))),
        "max_abs_dV_mV_all_samples": float(np.max(np.abs(ref - tr)))}


def measured_valu_busy(argv_key):
    """Fraction of SIMD cycles issuing VALU instructions in the timed launch,
    from the committed rocprofv3 SQ counter summary of this same command;
    None when made with other arguments or absent."""
    f = os.path.join(ROOT, "profiles", "k_cell_run_1024pkj_valu_lds.md")
    if not os.path.exists(f):
        return None
    key, val = None, None
    with open(f) as fh:
        for line in fh:
            if line.startswith("| bench arguments |"):
                key = line.split("|")[2].strip().strip("`")
            if line.startswith("| VALU busy"):
                val = float(line.split("|")[2].strip().strip("*"))
    return val if key == argv_key else None


def measured_traffic(argv_key):
    """HBM bytes of the timed launch of the run kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this same command (scaled by
    the plain-copy calibration recorded there); None when the summary was
    made with other arguments or is absent."""
    f = os.path.join(ROOT, "profiles", "k_cell_run_1024pkj_traffic.md")
    if not os.path.exists(f):
        return None
    key, val = None, None
    with open(f) as fh:
        for line in fh:
            if line.startswith("| bench arguments |"):
                key = line.split("|")[2].strip().strip("`")
            if line.startswith("| HBM bytes, timed launch |"):
                val = float(line.split("|")[2])
    return round(val) if key == argv_key and val else None


def cpu_baseline(args):
    """The oracle (FP64 C restatement of the reference RKC with its
    population-global controller, OpenMP over compartments) on the plain
    workload: 9 seeded synthetic Purkinje cells, rtol 1e-4 / atol 1e-6,
    0.5 nA at the soma over [20, 80) ms, 100 ms; on all host threads for up
    to --cpu-seconds, then on one thread for a third of that."""
    import cbrkc
    import orc
    n, seed = 9, 101
    morph = f"/tmp/cbrkc_bench_config1_s{seed}.txt"
    if not os.path.exists(morph):
        cbrkc.synth_pkj_morph(morph, seed=seed, ncell=n)
    v0, _ = philox_params(seed, n)
    steps_total = 800

    def timed(threads, seconds):
        o = orc.Pop.pkj(morph, n, 1, v0=v0, atol=1e-6, rtol=1e-4, libc_rand=False)
        o.stable_split()
        o.set_stim(1599, 20.0, 80.0, np.full(n, 0.5e-3))
        o.threads(threads)
        t = time.perf_counter()
        steps = 0
        while steps < steps_total and time.perf_counter() - t < seconds:
            o.set_current(steps * H_MAX)
            o.advance()
            steps += 1
        return steps, time.perf_counter() - t

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    steps, dt = timed(threads, args.cpu_seconds)
    steps1, dt1 = timed(1, args.cpu_seconds / 3)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": n * COMP * steps / dt, "unit": "compartment-steps/s", "cores": threads, "kind": "port",
            "sample": f"plain workload: 9 synthetic Purkinje cells x 1600 compartments, FP64 RKC rtol 1e-4 "
                      f"atol 1e-6, 0.5 nA over [20, 80) ms; {steps} of its 800 macro steps "
                      f"({steps * H_MAX:g} ms simulated) in {dt:.1f} s",
            "value_1_thread": n * COMP * steps1 / dt1,
            "sample_1_thread": f"{steps1} macro steps ({steps1 * H_MAX:g} ms) in {dt1:.1f} s",
            "cpu_model": model, "nproc": os.cpu_count(),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed bench steps (20 x 10 ms = 200 ms)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed bench steps before them")
    ap.add_argument("--step-macro", type=int, default=80,
                    help="macro steps (0.125 ms each) per bench step (80 = 10 ms simulated)")
    ap.add_argument("--scale-steps", type=int, default=400, help="scale workload: macro steps (400 = 50 ms)")
    ap.add_argument("--chunk", type=int, default=0,
                    help="macro steps per persistent launch (0: the whole timed run in one launch)")
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--tol", type=float, default=1e-6, help="atol = rtol (the reference's TOL_PKJ)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--plain-steps", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--threads", type=int, default=0,
                    help="threads per cell of the mixed engine (0: autotune among 256/320/512)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-noise-floor", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--morph", default=None, help="reuse an existing morphology file")
    ap.add_argument("--workload", choices=("precision", "scale", "rows", "batching", "network"), default="precision",
                    help="precision: the headline 1024-cell run; scale: the 8192-cell scale-stress comparison; "
                         "rows: every cell type x solver on the reference's morphologies, with the oracle beside it")
    ap.add_argument("--rows", default="PKJ,GR,GO,IO", help="cell types of the rows workload")
    ap.add_argument("--mix-scale", type=int, default=1, help="batching: divide every population by this")
    ap.add_argument("--mix-intervals", type=int, default=1000, help="batching: timed 0.1 ms intervals (100 ms)")
    ap.add_argument("--mix-trial", type=int, default=8, help="batching: autotuner trial intervals")
    ap.add_argument("--net-intervals", type=int, default=800, help="network: timed 0.125 ms intervals (100 ms)")
    ap.add_argument("--net-warmup", type=int, default=16, help="network: untimed intervals before them")
    ap.add_argument("--net-scale", type=int, default=1, help="network: divide every population by this")
    ap.add_argument("--net-cpu-intervals", type=int, default=16, help="network: intervals of the oracle sample")
    ap.add_argument("--net-gr", choices=("rkc", "cn-tree"), default="rkc",
                    help="network: the granule population's solver (the reference's RKC, or Crank-Nicolson with "
                         "the exact tree solve)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration << 20, "repeats": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the cells would be sharded across GPUs; that path is not built)", file=sys.stderr)
        sys.exit(2)

    import cbrkc
    from trace_metrics import compare

    if args.workload == "scale":
        return scale_stress(cbrkc, compare, args)
    if args.workload == "rows":
        return rows(cbrkc, args)
    if args.workload == "batching":
        return batching(cbrkc, args)
    if args.workload == "network":
        return network(cbrkc, args)
    n = args.cells
    morph = args.morph or f"/tmp/cbrkc_bench_pkj{n}_s{args.seed}.txt"
    if not os.path.exists(morph):
        cbrkc.synth_pkj_morph(morph, seed=args.seed, ncell=n)
    v0, amp = philox_params(args.seed, n)
    warm = args.warmup * args.step_macro        # macro steps
    timed = args.steps * args.step_macro
    t0 = warm * H_MAX

    # ---- mixed engine (the measured value) ----
    print(f"[bench] mixed engine: {n} cells, warmup {warm}, timed {timed} macro steps", file=sys.stderr, flush=True)
    pm = make_pop(cbrkc, cbrkc.ENGINE_CELL_MIXED, morph, n, v0, amp, args)
    nt = pm.autotune(4, threads=args.threads)
    tw_m, _ = timed_run(pm, warm, args.chunk, 0.0)
    pm.kernel_time(reset=True)
    s0 = pm.stats()
    tr_m, wall_m = timed_run(pm, timed, args.chunk, t0)
    kms, nlaunch = pm.kernel_time(reset=True)
    wm = work(s0, pm.stats())
    value = n * COMP * timed / wall_m

    # roofline of the persistent kernel over the timed region: the HBM bytes,
    # LDS bytes and FP32 flops its work needs (per-unit figures above), over
    # its HIP-event-timed duration on the engine's stream
    rhs = wm["stages"] - wm["steps"] - wm["rejects"] + wm["rho_iters"]
    bytes_ = COMP * (BYTES_ION * wm["ion_updates"] + BYTES_MACRO * n * timed + BYTES_LAUNCH * n * nlaunch)
    lds = COMP * LDS_RHS * rhs
    flops = COMP * (FLOP_RHS * rhs + FLOP_STEP * wm["steps"] + FLOP_ION * wm["ion_updates"])
    ksec = kms * 1e-3
    gbs = bytes_ / ksec / 1e9
    argv_key = f"--steps {args.steps} --warmup {args.warmup} --step-macro {args.step_macro} --cells {n}"
    traffic = measured_traffic(argv_key)
    roof = {"bound": "hbm", "achieved": round(gbs, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(gbs / PEAK_HBM_GBS, 5), "traffic": traffic,
            "kernel": f"k_cell_run<float,{nt}> (persistent, one workgroup per cell)",
            "kernel_ms_per_launch": round(kms / nlaunch, 4), "launches": nlaunch,
            "macro_steps_per_launch": timed / nlaunch,
            "algorithmic_hbm_bytes_per_launch": round(bytes_ / nlaunch),
            "compartment_ion_updates_per_launch": round(COMP * wm["ion_updates"] / nlaunch),
            "compartment_rhs_per_launch": round(COMP * rhs / nlaunch)}
    if traffic:
        # HBM bytes that actually moved (rocprofv3 counters of this command):
        # above the algorithmic bytes means re-reads and spill traffic
        roof["traffic_rate_GBs"] = round(traffic / (kms / nlaunch * 1e-3) / 1e9, 1)
        roof["traffic_frac"] = round(roof["traffic_rate_GBs"] / PEAK_HBM_GBS, 4)
        roof["traffic_over_algorithmic"] = round(traffic / (bytes_ / nlaunch), 3)
    on_chip = {"lds": {"achieved": round(lds / ksec / 1e9, 1), "peak": PEAK_LDS_GBS, "unit": "GB/s",
                       "frac": round(lds / ksec / 1e9 / PEAK_LDS_GBS, 5)},
               "fp32_valu": {"achieved": round(flops / ksec / 1e12, 4), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                             "frac": round(flops / ksec / 1e12 / PEAK_FP32_TFLOPS, 5)}}
    busy = measured_valu_busy(argv_key)
    if busy is not None:
        # issue-slot occupancy of the vector ALU (all instructions, not only
        # the algorithmic flops), rocprofv3 SQ counters of this command
        on_chip["valu_busy_measured"] = busy
    fr = {"hbm": roof["frac"], "lds": on_chip["lds"]["frac"], "fp32_valu": on_chip["fp32_valu"]["frac"]}
    if busy is not None:
        fr["valu_issue"] = busy
    roof["binding"] = max(fr, key=fr.get)

    out = {
        "metric": METRIC, "value": value, "unit": "compartment-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall_m / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "precision": "f32 stages/gates/currents, compensated f32 voltage, f64 step control and Ca",
        "data": "synthetic",
        "config": {"workload": f"{n} synthetic Purkinje cells x 1600 compartments (Pkj.txt format, per-cell seeded "
                               f"trees), RKC atol=rtol={args.tol:g}, soma current U(0.2,1.0) nA, dendritic Poisson "
                               "5 Hz x 1 nS tau 2 ms; safeguards: ion interval 2^-8 ms, fn refresh, spectral-radius "
                               f"floor; one step = {args.step_macro} macro steps ({args.step_macro * H_MAX:g} ms)",
                   "cells": n, "compartments": n * COMP, "sim_ms": timed * H_MAX, "macro_steps": timed,
                   "threads_per_cell": nt},
        "sim_ms_per_s": timed * H_MAX / wall_m,
        "mixed": {"wall_s": wall_m, "work": wm},
        "roofline": roof, "roofline_on_chip": on_chip,
    }

    # ---- per-cell FP64 engine: error of the mixed run ----
    if not args.no_fp64:
        print("[bench] per-cell FP64 engine on the same run", file=sys.stderr, flush=True)
        pd = make_pop(cbrkc, cbrkc.ENGINE_CELL_FP64, morph, n, v0, amp, args)
        pd.autotune(0, threads=256)
        tw_d, _ = timed_run(pd, warm, args.chunk, 0.0)
        s0 = pd.stats()
        tr_d, wall_d = timed_run(pd, timed, args.chunk, t0)
        out["fp64_cell"] = {"value": n * COMP * timed / wall_d, "wall_s": wall_d, "work": work(s0, pd.stats())}
        # the errors are stated over the precision workload itself: the
        # first 200 ms from t = 0 (warm-up included), whatever the warm-up
        win = min(PRECISION_STEPS, warm + timed)
        full_d = np.vstack([tw_d, tr_d])[:win]
        out["error_window_ms"] = [0.0, win * H_MAX]
        out["error_vs_fp64"] = error_report(compare, full_d, np.vstack([tw_m, tr_m])[:win], "mixed")
        del pd
        if not args.no_noise_floor:
            # the same FP64 run with another thread mapping: only the order of
            # the norm reductions differs, so this is the FP64 model's own
            # sensitivity to rounding over the same horizon
            print("[bench] FP64 noise floor (FP64 engine, 320 threads per cell)", file=sys.stderr, flush=True)
            pf = make_pop(cbrkc, cbrkc.ENGINE_CELL_FP64, morph, n, v0, amp, args)
            pf.autotune(0, threads=320)
            tw_f, _ = timed_run(pf, warm, args.chunk, 0.0)
            tr_f, _ = timed_run(pf, timed, args.chunk, t0)
            out["fp64_noise_floor"] = error_report(compare, full_d, np.vstack([tw_f, tr_f])[:win], "fp64_reordered")
            del pf

    # ---- plain FP64 population-global RKC ----
    if not args.no_plain:
        print(f"[bench] plain FP64 global engine, {args.plain_steps} macro steps", file=sys.stderr, flush=True)
        pg = make_pop(cbrkc, cbrkc.ENGINE_GLOBAL_FP64, morph, n, v0, amp, args)
        pg.run(2, 0.0, 1599, record=False)
        s0 = pg.stats()
        _, wall_g = timed_run(pg, args.plain_steps, args.plain_steps, 2 * H_MAX, record=False)
        vg = n * COMP * args.plain_steps / wall_g
        out["plain_fp64_global"] = {"value": vg, "wall_s": wall_g, "macro_steps": args.plain_steps,
                                    "work": work(s0, pg.stats())}
        out["speedup_vs_plain_fp64"] = value / vg
        del pg

    # ---- CPU oracle ----
    if not args.no_cpu:
        print("[bench] CPU oracle sample", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


def scale_stress(cbrkc, compare, args):
    """8192 synthetic Purkinje cells (diameters x U(0.5, 2.0) per cell, so the
    cells' spectral radii and stage counts spread), soma currents U(0.2, 1.0)
    nA, no dendritic drive, 50 ms: the mixed engine (measured), the per-cell
    FP64 RKC engine (the accuracy reference), the plain FP64 global RKC on a
    sample of macro steps, and the FP64 CN engine with the exact per-cell tree
    solve at the reference's 0.025 ms step and at 2^-5 .. 2^-8 ms; the CN
    comparison point is the fastest step whose median per-cell subthreshold
    error against the FP64 RKC run is within 0.1 mV."""
    n = 8192 if args.cells == 1024 else args.cells
    steps = args.scale_steps
    morph = args.morph or f"/tmp/cbrkc_scale_pkj{n}_s{args.seed}.txt"
    if not os.path.exists(morph):
        cbrkc.synth_pkj_morph(morph, seed=args.seed, ncell=n, dscale=philox_dscale(args.seed, n))
    v0, amp = philox_params(args.seed, n)
    stim = (1599, 0.0, 1e9, amp)

    def rkc(engine):
        return cbrkc.Population(cbrkc.PKJ, n, 1, morph=morph, v0=v0, atol=args.tol, rtol=args.tol, stim=stim,
                                engine=engine, **cbrkc.STABLE)

    out = {"metric": f"compartment-steps/s, {n} PCs scale stress (diameters x U(0.5,2.0)), {steps * H_MAX:g} ms",
           "unit": "compartment-steps/s", "n_gpus": 1, "steps": steps, "higher_is_better": True, "data": "synthetic",
           "config": {"workload": f"{n} synthetic Purkinje cells x 1600 compartments, diameters scaled per cell by "
                                  f"U(0.5,2.0), soma current U(0.2,1.0) nA, RKC atol=rtol={args.tol:g} with the "
                                  "safeguards; CN: DT 0.025 ms, exact per-cell tree solve",
                      "cells": n, "compartments": n * COMP, "sim_ms": steps * H_MAX}}
    print(f"[scale] mixed engine, {n} cells, {steps} macro steps", file=sys.stderr, flush=True)
    pm = rkc(cbrkc.ENGINE_CELL_MIXED)
    nt = pm.autotune(4, threads=args.threads)
    tr_m, wall_m = timed_run(pm, steps, args.chunk, 0.0)
    out["value"] = n * COMP * steps / wall_m
    out["mixed"] = {"value": out["value"], "wall_s": wall_m, "threads_per_cell": nt, "work": pm.stats()}
    del pm
    print("[scale] per-cell FP64 RKC engine", file=sys.stderr, flush=True)
    pd = rkc(cbrkc.ENGINE_CELL_FP64)
    pd.autotune(0, threads=256)
    tr_d, wall_d = timed_run(pd, steps, args.chunk, 0.0)
    out["fp64_cell"] = {"value": n * COMP * steps / wall_d, "wall_s": wall_d, "work": pd.stats()}
    del pd
    out["error_mixed_vs_fp64"] = error_report(compare, tr_d, tr_m, "mixed")
    print(f"[scale] plain FP64 global RKC, {args.plain_steps} macro steps", file=sys.stderr, flush=True)
    pg = rkc(cbrkc.ENGINE_GLOBAL_FP64)
    _, wall_g = timed_run(pg, args.plain_steps, args.plain_steps, 0.0, record=False)
    out["plain_fp64_global"] = {"value": n * COMP * args.plain_steps / wall_g, "wall_s": wall_g,
                                "macro_steps": args.plain_steps}
    del pg
    # FP64 CN with the exact per-cell tree solve at the reference's step and at
    # finer ones: its error against the FP64 RKC run and its throughput; the
    # reference's 0.025 ms step is not stable on these stiff synthetic somata
    # (non-finite cells are counted instead of compared)
    cn = []
    runs = [(cbrkc.CN, dt) for dt in (0.025, 2.0 ** -5, 2.0 ** -6, 2.0 ** -7, 2.0 ** -8)]
    runs += [(cbrkc.BE, 0.025), (cbrkc.BE, 2.0 ** -8)]   # backward Euler: L-stable, first order
    for solver, dt in runs:
        name = "CN" if solver == cbrkc.CN else "BE"
        print(f"[scale] FP64 {name} engine, exact per-cell tree solve, step {dt:g} ms", file=sys.stderr, flush=True)
        pc = cbrkc.Population(cbrkc.PKJ, n, 1, solver=solver, engine=cbrkc.ENGINE_TREE_FP64, morph=morph, v0=v0,
                              stim=stim, step_ms=None if dt == 0.025 else dt)
        tr_c, wall_c = timed_run(pc, steps, args.chunk, 0.0)
        entry = {"solver": name, "step_ms": dt, "value": n * COMP * steps / wall_c, "wall_s": wall_c,
                 "steps": pc.stats()["steps"]}
        del pc
        bad = ~np.all(np.isfinite(tr_c), axis=0) | np.any(np.abs(tr_c) > 1e3, axis=0)
        entry["cells_nonfinite_or_beyond_1V"] = int(bad.sum())
        if not bad.any():
            entry["error_vs_fp64_rkc"] = error_report(compare, tr_d, tr_c, "cn_tree")
        cn.append(entry)
    out["fp64_cn_tree"] = cn
    ok = [e for e in cn if e["solver"] == "CN" and "error_vs_fp64_rkc" in e
          and e["error_vs_fp64_rkc"]["median_per_cell_subthreshold_abs_dV_mV"] <= 0.1]
    out["speedup_mixed_vs_plain_fp64"] = out["value"] / out["plain_fp64_global"]["value"]
    if ok:
        best = max(ok, key=lambda e: e["value"])
        out["fp64_cn_tree_at_0.1mV"] = {"step_ms": best["step_ms"], "value": best["value"]}
        out["speedup_mixed_vs_fp64_cn_tree_at_0.1mV"] = out["value"] / best["value"]
    print(json.dumps(out), flush=True)


REF_DATA = os.path.join(ROOT, "tests", "golden", "ref_data")


def uniform(seed, stream, n, lo, hi):
    """n draws of U(lo, hi) from Philox4x32-10 (ctr = (i, stream, 0, 0), key = (seed, 0x4d495820))"""
    import cbrkc
    out = np.zeros(n)
    for i in range(n):
        r = cbrkc.philox4x32_10([i, stream, 0, 0], [seed & 0xffffffff, 0x4D495820])
        out[i] = lo + (hi - lo) * (((r[0] >> 5) * 67108864.0 + (r[1] >> 6)) / 9007199254740992.0)
    return out


def batching(cbrkc, args):
    """The uneven mix in one persistent launch per 0.1 ms exchange interval:
    16384 granule cells (4-compartment synthetic chains), 2048 Golgi cells
    (synthetic 131-compartment trees), 1024 Purkinje cells (1600-compartment
    synthetic trees, mixed precision) and 2048 olive cells (3 compartments,
    each joined to 6 random partners, g ~ U(0.5, 1.5) nS), constant per-cell
    currents from per-type uniform ranges, 100 ms.  The lanes per small cell
    are chosen by the autotuner; chunks of 64 intervals replay a captured
    graph."""
    scale = args.mix_scale
    n = {cbrkc.GR: 16384 // scale, cbrkc.GO: 2048 // scale, cbrkc.PKJ: 1024 // scale, cbrkc.IO: 2048 // scale}
    comp = {cbrkc.GR: 4, cbrkc.GO: 131, cbrkc.PKJ: 1600, cbrkc.IO: 3}
    interval = 0.1
    nint = args.mix_intervals
    seed = args.seed
    d = "/tmp/cbrkc_mix"
    os.makedirs(d, exist_ok=True)
    grf, gof, pkf, gjf = (os.path.join(d, f) for f in (f"gr{n[0]}.txt", f"go{n[1]}.txt", f"pkj{n[2]}.txt",
                                                        f"io{n[3]}_gj.csv"))
    cbrkc.synth_gr_chains(grf, os.path.join(REF_DATA, "Gr.txt"), seed, n[cbrkc.GR])
    cbrkc.synth_go_trees(gof, os.path.join(REF_DATA, "Go_131.txt"), seed, n[cbrkc.GO])
    if not os.path.exists(pkf):
        cbrkc.synth_pkj_morph(pkf, seed=seed, ncell=n[cbrkc.PKJ])
    cbrkc.synth_io_gj(gjf, seed, n[cbrkc.IO], k=6, g_min_nS=0.5, g_max_nS=1.5)
    pv0, pamp = philox_params(seed, n[cbrkc.PKJ])
    pops = {
        cbrkc.GR: dict(n=n[cbrkc.GR], morph=grf, comp=4, v0=uniform(seed, 10, n[cbrkc.GR], -72.0, -68.0),
                       amp=uniform(seed, 11, n[cbrkc.GR], 0.008e-3, 0.016e-3)),
        cbrkc.GO: dict(n=n[cbrkc.GO], morph=gof, v0=uniform(seed, 12, n[cbrkc.GO], -62.0, -58.0),
                       amp=uniform(seed, 13, n[cbrkc.GO], 0.1e-3, 0.3e-3)),
        cbrkc.PKJ: dict(n=n[cbrkc.PKJ], morph=pkf, v0=pv0, amp=pamp),
        cbrkc.IO: dict(n=n[cbrkc.IO], morph=os.path.join(REF_DATA, "IO.txt"),
                       v0=uniform(seed, 14, n[cbrkc.IO], -62.0, -58.0), amp=uniform(seed, 15, n[cbrkc.IO], 3.0, 6.0)),
    }
    print(f"[batching] building the mix {n}", file=sys.stderr, flush=True)
    m = cbrkc.Mix(pops, io_gj=gjf, interval_ms=interval, pkj_tol=args.tol)
    lanes = m.autotune(args.mix_trial)
    print(f"[batching] lanes per cell (GR, GO, PKJ, IO) {lanes}", file=sys.stderr, flush=True)
    m.run(args.warmup, record=False)
    s0 = {t: m.stats(t) for t in range(4)}
    m.kernel_time(reset=True)
    t = time.perf_counter()
    rec, spk = m.run(nint)
    wall = time.perf_counter() - t
    kms, nl = m.kernel_time(reset=True)
    s1 = {t: m.stats(t) for t in range(4)}
    total_comp = sum(n[t] * comp[t] for t in range(4))
    sim_ms = nint * interval
    value = total_comp * sim_ms / 0.125 / wall
    names = {cbrkc.GR: "granule", cbrkc.GO: "Golgi", cbrkc.PKJ: "Purkinje", cbrkc.IO: "olive"}
    per = {}
    for tt in range(4):
        w = work(s0[tt], s1[tt])
        sp = int(m.spikes(spk, tt).sum())
        per[names[tt]] = {"cells": n[tt], "compartments": n[tt] * comp[tt], "lanes_per_cell": lanes[tt],
                          "spikes": sp, "work": w}
    out = {"metric": "compartment-steps/s, batched uneven mix (GR/GO/PKJ/IO), one launch per 0.1 ms interval",
           "value": value, "unit": "compartment-steps/s (one compartment over 0.125 ms simulated)", "n_gpus": 1,
           "higher_is_better": True, "data": "synthetic", "dtype": "f32 (Purkinje) / f64 (granule, Golgi, olive)",
           "config": {"workload": f"{n[0]} granule (4-compartment chains), {n[1]} Golgi (131), {n[2]} Purkinje "
                                  f"(1600), {n[3]} olive (3, 6 random junctions per cell, U(0.5,1.5) nS); constant "
                                  f"currents; 0.1 ms exchange intervals; {sim_ms:g} ms",
                      "compartments": total_comp, "intervals": nint, "interval_ms": interval},
           "wall_s": wall, "sim_ms_per_s": sim_ms / wall,
           "kernel_ms": kms, "launch_chunks": nl, "populations": per,
           "all_finite": bool(np.all(np.isfinite(rec)))}
    if not args.no_cpu:
        out["cpu_baseline"] = batching_cpu(cbrkc, pops, gjf, interval, args)
    print(json.dumps(out), flush=True)


def network(cbrkc, args):
    """The fixed in-degree network: 2048 granule cells (4-compartment
    synthetic chains), 512 Golgi cells (synthetic 131-compartment trees), 512
    Purkinje cells (synthetic 1600-compartment trees), 780404 synapses
    (4 mossy-fibre and 4 Golgi inputs per granule cell, 100 granule inputs
    per Golgi cell, 1392-1393 per Purkinje cell), weights U(0.5, 1.5) nS,
    AMPA 1.5 ms / GABA_A 6 ms, 1 ms delay, mossy fibres at 5 Hz.  Granule and
    Golgi cells on the global FP64 RKC engine (the reference's tolerances),
    Purkinje cells on the per-cell mixed engine; the same network with the
    Purkinje cells on the per-cell FP64 engine is run beside it for the
    raster comparison."""
    import net_oracle as NO
    sc = args.net_scale
    n = (2048 // sc, 512 // sc, 512 // sc)
    comp = (4, 131, 1600)
    d = "/tmp/cbrkc_net"
    os.makedirs(d, exist_ok=True)
    files = NO.synth_files(d, *n, seed=args.seed)
    rng_v0 = (uniform(args.seed, 20, n[0], -72.0, -68.0), uniform(args.seed, 21, n[1], -62.0, -58.0),
              philox_params(args.seed, n[2])[0])
    cfg = cbrkc.net_config()
    if sc != 1:
        cfg.total = 0
    nint, warm = args.net_intervals, args.net_warmup

    def build(pkj_engine):
        if args.net_gr == "cn-tree":
            gr = cbrkc.Population(cbrkc.GR, n[0], 1, solver=cbrkc.CN, engine=cbrkc.ENGINE_TREE_FP64, morph=files[0],
                                  comp=4, v0=rng_v0[0], stim="none")
            gr.autotune(4)
        else:
            gr = cbrkc.Population(cbrkc.GR, n[0], 1, morph=files[0], comp=4, v0=rng_v0[0], stim="none")
        go = cbrkc.Population(cbrkc.GO, n[1], 1, morph=files[1], v0=rng_v0[1], stim="none")
        pk = cbrkc.Population(cbrkc.PKJ, n[2], 1, morph=files[2], gj="", v0=rng_v0[2], atol=args.tol, rtol=args.tol,
                              stim="none", engine=pkj_engine, **cbrkc.STABLE)
        threads = pk.autotune(4, threads=args.threads)
        return (gr, go, pk), threads

    res = {}
    for name, eng in (("mixed", cbrkc.ENGINE_CELL_MIXED), ("fp64", cbrkc.ENGINE_CELL_FP64)):
        if name == "fp64" and args.no_fp64:
            continue
        print(f"[network] {n} cells, Purkinje engine {name}: warm-up {warm}, timed {nint} intervals",
              file=sys.stderr, flush=True)
        pops, threads = build(eng)
        net = cbrkc.Network(*pops, cfg=cfg)
        st = net.stats()
        net.run(warm)
        t0 = net.times()
        s0 = [p.stats() for p in pops]
        t = time.perf_counter()
        done = 0
        while done < nint:   # in pieces of 100 ms, with a progress line after each
            k = min(800, nint - done)
            net.run(k)
            done += k
            print(f"[network] {name}: {done} of {nint} intervals, {time.perf_counter() - t:.1f} s",
                  file=sys.stderr, flush=True)
        wall = time.perf_counter() - t
        t1 = net.times()
        s1 = [p.stats() for p in pops]
        spikes = net.spikes()
        net.close()
        for p in pops:
            p.close()
        res[name] = dict(wall=wall, stats=st, threads=threads, spikes=spikes,
                         pop_ms=[a - b for a, b in zip(t1["populations"], t0["populations"])],
                         exchange_ms=t1["exchange"] - t0["exchange"],
                         work=[work(a, b) for a, b in zip(s0, s1)])
    m = res["mixed"]
    total_comp = sum(a * b for a, b in zip(n, comp))
    sim_ms = nint * H_MAX
    names = ("granule", "Golgi", "Purkinje")
    t_m, p_m, c_m = m["spikes"]
    per = {}
    for q in range(3):
        per[names[q]] = {"cells": n[q], "compartments": n[q] * comp[q],
                         "host_ms_in_own_loop": round(m["pop_ms"][q], 1), "spikes": int(np.sum(p_m == q)),
                         "work": m["work"][q]}
    out = {"metric": "compartment-steps/s, GR/GO/PKJ network (780404 synapses), 0.125 ms exchange, 1 ms delay",
           "value": total_comp * nint / m["wall"], "unit": "compartment-steps/s (one compartment over 0.125 ms "
           "simulated)", "n_gpus": 1, "higher_is_better": True, "data": "synthetic",
           "dtype": "f32 (Purkinje, mixed) / f64 (granule, Golgi)",
           "config": {"workload": f"{n[0]} granule (4-compartment chains, "
                                  f"{'RKC' if args.net_gr == 'rkc' else 'CN with the exact tree solve'}), {n[1]} Golgi "
                                  f"(131), {n[2]} Purkinje (1600); fixed in-degree network, "
                                  f"{sum(m['stats']['synapses'])} synapses; {sim_ms:g} ms timed after {warm * H_MAX:g} ms",
                      "compartments": total_comp, "intervals": nint, "synapses": m["stats"]["synapses"],
                      "purkinje_threads_per_cell": m["threads"]},
           "wall_s": m["wall"], "sim_ms_per_s": sim_ms / m["wall"], "populations": per,
           "exchange_ms": round(m["exchange_ms"], 1)}
    if "fp64" in res:
        f = res["fp64"]
        a = set(zip(np.rint(f["spikes"][0] / H_MAX).astype(int).tolist(), f["spikes"][1].tolist(),
                    f["spikes"][2].tolist()))
        b = set(zip(np.rint(t_m / H_MAX).astype(int).tolist(), p_m.tolist(), c_m.tolist()))
        small_equal = {s for s in a if s[1] < 2} == {s for s in b if s[1] < 2}
        pk_a = sorted(s for s in a if s[1] == 2)
        pk_b = sorted(s for s in b if s[1] == 2)
        cnt_a = np.bincount([s[2] for s in pk_a], minlength=n[2])
        cnt_b = np.bincount([s[2] for s in pk_b], minlength=n[2])
        shifts = []
        for c in np.nonzero(cnt_a == cnt_b)[0]:
            ta = [s[0] for s in pk_a if s[2] == c]
            tb = [s[0] for s in pk_b if s[2] == c]
            shifts += [abs(x - y) for x, y in zip(ta, tb)]
        out["vs_fp64_purkinje_run"] = {
            "fp64_wall_s": f["wall"],
            "granule_golgi_rasters_identical": bool(small_equal),
            "purkinje_spikes_fp64": len(pk_a), "purkinje_spikes_mixed": len(pk_b),
            "purkinje_cells_with_count_mismatch": int(np.sum(cnt_a != cnt_b)),
            "purkinje_spike_shift_intervals_max": int(max(shifts)) if shifts else 0,
            "purkinje_spikes_in_the_same_interval": int(sum(1 for x in shifts if x == 0))}
    if not args.no_cpu:
        out["cpu_baseline"] = network_cpu(cfg, files, n, rng_v0, args)
    print(json.dumps(out), flush=True)


def network_cpu(cfg, files, n, v0, args):
    """The oracle's network (FP64 RKC, the reference's population-global
    controllers; the granule and Golgi oracle readers take cell 0's
    geometry for all cells) over the same populations for the first
    --net-cpu-intervals intervals, every population on the host's threads."""
    import ctypes
    import net_oracle as NO
    lib = NO.L()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    opops = NO.oracle_pops(files, v0, pkj_tol=args.tol)
    for p in opops:
        p.threads(threads)
    h = lib.or_net_create(opops[0].h, opops[1].h, opops[2].h, ctypes.byref(NO.spec_of(cfg)))
    k = args.net_cpu_intervals
    t = time.perf_counter()
    rc = lib.or_net_run(h, k, None)
    wall = time.perf_counter() - t
    lib.or_net_free(h)
    total = sum(a * b for a, b in zip(n, (4, 131, 1600)))
    return {"value": total * k / wall, "unit": "compartment-steps/s", "cores": threads, "kind": "port",
            "sample": f"the oracle's network over the same {n} cells, first {k} intervals ({k * H_MAX:g} ms) "
                      f"in {wall:.1f} s, status {rc}"}


def batching_cpu(cbrkc, pops, gjf, interval, args):
    """The oracle on a bounded sample of the mix, each cell the reference's
    RKC on a population of one (as the engine advances it), for up to
    --cpu-seconds per type: compartment-steps/s of the sample on one host
    thread per cell, all cells of the sample in parallel host threads."""
    import ctypes
    from concurrent.futures import ThreadPoolExecutor
    import orc
    from test_gpu_mix import _L, _blocks
    lib = _L()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    d = "/tmp/cbrkc_mix_cpu"
    os.makedirs(d, exist_ok=True)
    sample = {cbrkc.GR: 16, cbrkc.GO: 4, cbrkc.PKJ: 2, cbrkc.IO: 16}
    comp = {cbrkc.GR: 4, cbrkc.GO: 131, cbrkc.PKJ: 1600, cbrkc.IO: 3}
    tot_cs, tot_s = 0.0, 0.0
    per = {}
    for t, k in sample.items():
        spec = pops[t]
        if t == cbrkc.GR:
            files = _blocks(spec["morph"], 4, spec["n"], d, "gr")[:k]
            mk = lambda c: orc.Pop(lib.or_gr_create_comp(files[c].encode(), 4, 1, 1, orc._dp(spec["v0"][c:c + 1]),
                                                         1e-9, 1e-9, 0), 1, 4)
        elif t == cbrkc.GO:
            files = _blocks(spec["morph"], 131, spec["n"], d, "go")[:k]
            mk = lambda c: orc.Pop(lib.or_go_create(files[c].encode(), 1, 1, orc._dp(spec["v0"][c:c + 1]), 1e-6,
                                                    1e-6, 0), 1, 131)
        elif t == cbrkc.PKJ:
            files = orc.split_morph(spec["morph"], spec["n"], 1600, d)[:k]
            def mk(c, files=files, spec=spec):
                o = orc.Pop.pkj(files[c], 1, 1, v0=spec["v0"][c:c + 1], atol=args.tol, rtol=args.tol,
                                libc_rand=False)
                o.stable_split()
                return o
        else:
            mk = lambda c: orc.Pop(lib.or_io_create(spec["morph"].encode(), 1, 1, orc._dp(spec["v0"][c:c + 1]),
                                                    1e-6, 1e-6, 0), 1, 3)

        stim_comp = {cbrkc.GR: 1, cbrkc.GO: 30, cbrkc.PKJ: 1599, cbrkc.IO: 0}[t]

        def one(c, mk=mk, t=t, stim_comp=stim_comp, spec=spec):
            o = mk(c)
            o.set_stim(stim_comp, 0.0, 1e300, np.array([spec["amp"][c]]))
            o.set_current(0.0)
            o.threads(1)
            steps = 0
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < args.cpu_seconds / 4:
                o.advance()
                steps += 1
            return steps, time.perf_counter() - t0

        with ThreadPoolExecutor(max_workers=min(k, threads)) as ex:
            res = list(ex.map(one, range(k)))
        cs = sum(st * comp[t] for st, _ in res)
        sec = max(w for _, w in res)
        per[t] = {"cells": k, "macro_steps": [st for st, _ in res], "wall_s": sec}
        tot_cs += cs
        tot_s += sec
    return {"value": tot_cs / tot_s, "unit": "compartment-steps/s (0.125 ms output intervals)", "cores": threads,
            "kind": "port", "sample": "oracle, each cell its own population (the reference's RKC per cell): "
                                      "16 granule chains, 4 Golgi trees, 2 Purkinje cells, 16 olive cells "
                                      "(uncoupled), one host thread per cell, the four types one after another, "
                                      f"about {args.cpu_seconds / 4:g} s each",
            "per_type": {str(k): v for k, v in per.items()}}


def rows(cbrkc, args):
    """Every cell type with every solver on the reference's own morphology
    files (reference semantics: population-global RKC, the CN / BE Krylov
    solves, and the exact per-cell tree solve), each timed on the GPU and, on
    a bounded sample, in the oracle on the host cores."""
    types = {  # name: (cell type, file, nx, ny, initial V, macro steps)
        "PKJ": (cbrkc.PKJ, "Pkj.txt", 1, 64, -65.0, 8),
        "GR": (cbrkc.GR, "Gr.txt", 16, 16, -70.0, 4),
        "GO": (cbrkc.GO, "Go_131.txt", 16, 16, -60.0, 8),
        "IO": (cbrkc.IO, "IO.txt", 32, 32, -60.0, 8),
    }
    solvers = (("RKC", cbrkc.RKC, cbrkc.ENGINE_GLOBAL_FP64), ("CN", cbrkc.CN, cbrkc.ENGINE_GLOBAL_FP64),
               ("CN tree", cbrkc.CN, cbrkc.ENGINE_TREE_FP64), ("BE", cbrkc.BE, cbrkc.ENGINE_GLOBAL_FP64),
               ("BE tree", cbrkc.BE, cbrkc.ENGINE_TREE_FP64))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = []
    for name in args.rows.split(","):
        ctype, fname, nx, ny, v, k = types[name]
        n = nx * ny
        morph = os.path.join(REF_DATA, fname)
        for sname, solver, engine in solvers:
            row = {"cells": name, "solver": sname, "n": n}
            print(f"[rows] {name} x {n}, {sname}", file=sys.stderr, flush=True)
            try:
                p = cbrkc.Population(ctype, nx, ny, solver=solver, engine=engine, morph=morph, v0=np.full(n, v))
            except cbrkc.CbError as e:
                row["refused"] = str(e)
                out.append(row)
                continue
            rec = {cbrkc.PKJ: 1599, cbrkc.GR: 16, cbrkc.GO: 30, cbrkc.IO: 0}[ctype]
            p.run(1, 0.0, rec)                   # recording copies back, so each call ends synchronised
            t = time.perf_counter()
            p.run(k, H_MAX, rec)
            wall = time.perf_counter() - t
            row.update(compartments=p.nc, macro_steps=k, wall_s=wall, value=p.nc * k / wall,
                       work=p.stats())
            del p
            if not args.no_cpu:
                row["cpu_baseline"] = rows_cpu(ctype, morph, nx, ny, v, solver, engine, args, threads)
            out.append(row)
    print(json.dumps({"workload": "rows", "unit": "compartment-steps/s",
                      "note": "reference morphologies, reference semantics; cpu_baseline = the oracle (FP64 C, "
                              "OpenMP where it has loops over cells) on the same population",
                      "rows": out}), flush=True)


def rows_cpu(ctype, morph, nx, ny, v, solver, engine, args, threads):
    """The oracle on the same population for up to 4 macro steps or
    --cpu-seconds, whichever ends first."""
    import cbrkc
    import orc
    from test_oracle_direct import L
    from test_oracle_gr import gr_pop
    from test_oracle_go import go_pop
    from test_oracle_io import io_pop
    n = nx * ny
    v0 = np.full(n, v)
    ref = os.path.dirname(morph)
    if ctype == cbrkc.PKJ:
        o = orc.Pop.pkj(morph, nx, ny, v0=v0, libc_rand=False)
    elif ctype == cbrkc.GR:
        o = gr_pop(ref, nx, ny, v0=v0)
    elif ctype == cbrkc.GO:
        o = go_pop(ref, nx, ny, v0=v0)
    else:
        o = io_pop(ref, nx, ny, v0=v0)
    o.threads(threads)
    if solver == cbrkc.CN:
        assert L().or_cn_init(o.h) == 0
        adv = L().or_cn_advance
    elif solver == cbrkc.BE:
        assert L().or_be_init(o.h) == 0
        adv = L().or_be_advance
    if solver != cbrkc.RKC and engine == cbrkc.ENGINE_TREE_FP64:
        assert L().or_cn_set_direct(o.h, 1) == 0
    t = time.perf_counter()
    k = 0
    while k < 4 and (k == 0 or time.perf_counter() - t < args.cpu_seconds):
        if solver == cbrkc.RKC:
            o.set_current(k * H_MAX)
            o.advance()
        else:
            adv(o.h, k * H_MAX)
        k += 1
    dt = time.perf_counter() - t
    return {"value": o.nc * k / dt, "cores": threads, "kind": "port", "macro_steps": k, "wall_s": dt,
            "threaded": "all vector loops" if solver == cbrkc.RKC else "ion updates (the linear solve is serial)"}


if __name__ == "__main__":
    main()
