"""DBS evaluated flips/s on one MI355X (gfx950).

Workload (default): 21x21 tiles, disjoint layout (N = 10590), 9 frequencies
26..34 GHz, synthetic complex128 G generated on the device, Bernoulli(0.5)
initial pattern from std::mt19937(0).  One step = one DBS run from that
initial state to convergence (no flip of the raster cycle improves mean |S11|^2),
the window the reference times (pngf-opt/main.cpp:333-464).  Between steps the
optimizer state is restored from a device checkpoint outside the timed region.

value = committed evaluations (the reference's flip_count) / seconds of the timed steps,
each measured on the host around the DBS call (the device is idle when it starts and the
call returns after synchronising its stream).  Printed as one JSON line with, beside the
required fields: roofline (the column-strip kernel, timed by HIP events recorded on its
stream around every launch of the timed steps; HBM traffic, MFMA busy and clock from rocprofv3
counter passes over the same launches), cpu_baseline (the oracle on a bounded sample,
projected to the run's counts), config_7x7_1f (the reference's CPU-runnable 7x7 case: the GPU
beside the oracle), reference_config (the reference program's own default configuration:
LOWRANK and PLAIN on the GPU),
plain_fp64_gpu (the reference formulation on the GPU, same start), speedup_vs_plain,
vs_oracle_run (the run against the oracle's recorded whole run of the default workload),
best_improvement_variant, time_to_solution_s, per-kernel-category statistics, autotuner
choice and the symmetric-accept decision.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import pngf_amd as P  # noqa: E402
from pngf_amd.problem import MULTI16, WORKLOADS, initial_pattern  # noqa: E402

METRIC = "DBS evaluated flips/s, 21×21 tiles × 9 freqs, vs plain FP64 GPU at matched S11"
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
DX, DY, DZ = 0.01 / 60, 0.01 / 60, 0.01 / 180
HEADLINE_ACCEPTS = 183  # accepts of the default workload's run (deterministic)
# rocprofv3 --pmc counts of the column strip over the default run's 183 launches (python3 bench.py
# --counter-probe under FETCH_SIZE / WRITE_SIZE / MfmaUtil passes, summarised by --counter-summary;
# profiles/colblk_counters_antenna21x21_9f.txt)
STRIP_COUNTED = dict(N=10590, freqs=9, variant=4, launches=HEADLINE_ACCEPTS, alg_bytes_per_launch=3411258624.0,
                     bytes_per_launch=3507488159.6, mfma_util=0.7615, clock_ghz=2.223,
                     source="rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE (copy-calibrated), MfmaUtil over the same "
                            "183 strip launches (profiles/colblk_counters_antenna21x21_9f.txt)")


# the oracle's whole DBS runs on the default workload, and on its G perturbed non-reciprocally at
# 1e-7 (tests/golden/make_oracle_sequence.py: the C++ restatement of pngfopt.hpp to convergence on
# the host, start C^-1 by LAPACK)
ORACLE_RUN = os.path.join(ROOT, "tests", "golden", "oracle_antenna21x21_9f.npz")
ORACLE_RUN_NONRECIPROCAL = os.path.join(ROOT, "tests", "golden", "oracle_antenna21x21_9f_nonreciprocal.npz")


def oracle_run_check(r, path=ORACLE_RUN):
    """A run of the default workload against the oracle's recorded whole run: same accepted flips
    at the same evaluations (else the index of the first accept where they part), and the
    relative difference of the final objectives."""
    if not os.path.exists(path):
        return None
    z = np.load(path)
    same = (int(z["flip_count"]) == r["flip_count"] and np.array_equal(z["acc_tiles"], r["acc_tiles"]) and
            np.array_equal(z["acc_eval"], r["acc_eval"]))
    # the first accept (index into the oracle's list) where the two runs part, if they do
    first = None
    n = min(len(z["acc_tiles"]), len(r["acc_tiles"]))
    for i in range(n):
        if z["acc_tiles"][i] != r["acc_tiles"][i] or z["acc_eval"][i] != r["acc_eval"][i]:
            first = i
            break
    if first is None and not same:
        first = n
    return dict(same_accepts=bool(same), first_difference=first, accepts=int(len(z["acc_tiles"])),
                flip_count=int(z["flip_count"]), F_final=float(z["F_final"]),
                F_rel=abs(r["F_final"] - float(z["F_final"])) / float(z["F_final"]))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make(wl, method, max_deferred=0, symmetry=P.SYM_AUTO, tile_variant=0, perturb=0.0, precision=P.PREC_C128):
    T, freqs = wl["T"], wl["freqs"]
    opt = P.Optimizer(wl["layout"], T, num_freq=len(freqs), method=method, max_deferred=max_deferred,
                      symmetry=symmetry, tile_variant=tile_variant, precision=precision)
    for f, hz in enumerate(freqs):
        opt.set_frequency(f, hz)
        opt.synth_g(f)
        if perturb > 0.0:
            opt.perturb_g(f, perturb, seed=f + 1)
    opt.set_tile_map(initial_pattern(T, 0))
    mid = (T - 1) // 2
    opt.set_source_port(mid, mid, 1, 2)
    t0 = time.perf_counter()
    opt.init_state()
    return opt, time.perf_counter() - t0


def roofline(opt, variant=None):
    """Dominant kernel category by event-timed duration -> roofline record."""
    n_comp = opt.N
    best = None
    for k in range(6):
        s = opt.kernel_stats(k)
        if s["launches"] and (best is None or s["total_ms"] > best[1]["total_ms"]):
            best = (k, s)
    if best is None:
        return None, {}
    k, s = best
    avg_ms = s["total_ms"] / s["launches"]
    bytes_l = s["bytes"] / s["launches"]
    flops_l = s["flops"] / s["launches"]
    ai = flops_l / max(bytes_l, 1.0)
    if ai > PEAK_FP64_TFLOPS * 1e12 / (PEAK_HBM_GBS * 1e9):
        ach = flops_l / (avg_ms * 1e-3) / 1e12
        rec = dict(bound="mfma", achieved=ach, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                   frac=ach / PEAK_FP64_TFLOPS, traffic=None)
    else:
        ach = bytes_l / (avg_ms * 1e-3) / 1e9
        rec = dict(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s", frac=ach / PEAK_HBM_GBS,
                   traffic=None)
    rec.update(kernel=P.K_NAMES[k], avg_ms=avg_ms, launches=s["launches"], algo_bytes=bytes_l,
               algo_flops=flops_l)
    c = STRIP_COUNTED
    if (P.K_NAMES[k] == "colblk" and n_comp == c["N"] and opt.nf == c["freqs"] and variant == c["variant"]
            and s["launches"] % c["launches"] == 0 and abs(bytes_l / c["alg_bytes_per_launch"] - 1.0) < 1e-9):
        # the same launches (N, frequencies, k, K = 0, 24, ..., 24 (L - 1), tile variant) as the
        # counted strip probe: HBM bytes per launch from its copy-calibrated FETCH_SIZE + WRITE_SIZE,
        # and its MFMA-busy fraction (MfmaUtil) and clock beside the event-timed fraction of peak
        rec["traffic"] = c["bytes_per_launch"]
        rec["traffic_source"] = c["source"]
        rec["mfma_busy"] = c["mfma_util"]
        rec["clock_ghz"] = c["clock_ghz"]
    allk = {P.K_NAMES[i]: opt.kernel_stats(i) for i in range(6)}
    return rec, allk


def strip_stats(opt, peak=PEAK_FP64_TFLOPS):
    """Event-timed column and row strips of a profiled run: launches, average us and the
    fraction of the MFMA peak (FP64, or FP32 for the complex64 mode) at their algorithmic flops."""
    out = {}
    for k in (P.K_COLBLK, P.K_ROWBLK):
        s = opt.kernel_stats(k)
        if s["launches"]:
            avg = s["total_ms"] / s["launches"]
            tf = s["flops"] / (s["total_ms"] * 1e-3) / 1e12
            out[P.K_NAMES[k]] = dict(launches=s["launches"], avg_us=avg * 1e3, tflops=tf, frac=tf / peak,
                                     peak_tflops=peak)
    return out


def variant(wl, name, symmetry, perturb, headline, precision=P.PREC_C128, oracle_path=None):
    """The LOWRANK run from the same start on another accept path (or in complex64): one untimed
    warmup, one run for the rate, one profiled run for the strips' event times."""
    opt, init_s = make(wl, P.METHOD_LOWRANK, symmetry=symmetry, perturb=perturb, precision=precision)
    try:
        used, asym = opt.symmetry()
        opt.checkpoint()
        opt.dbs_run(max_evals=wl["max_evals"])
        opt.restore()
        r = opt.dbs_run(max_evals=wl["max_evals"])
        opt.restore()
        opt.profile(True)
        opt.dbs_run(max_evals=wl["max_evals"])
        strips = strip_stats(opt, PEAK_FP32_TFLOPS if precision == P.PREC_C64 else PEAK_FP64_TFLOPS)
        opt.profile(False)
        final_map = opt.tile_map()
    finally:
        opt.close()
    out = dict(value=r["flip_count"] / r["seconds"], seconds=r["seconds"], flip_count=r["flip_count"],
               accepts=r["accepts"], F_final=r["F_final"], symmetric_accept=used, rel_asym=asym,
               init_s=init_s, strips=strips, F_final_vs_headline=r["F_final"] / headline["F_final"],
               same_accept_sequence=bool(np.array_equal(r["acc_tiles"], headline["acc_tiles"])),
               vs_oracle_run=oracle_run_check(r, oracle_path) if oracle_path else None)
    if precision == P.PREC_C64:
        # the complex64 state's objective against a complex128 init_state of the design it reached
        fresh = P.Optimizer(wl["layout"], wl["T"], num_freq=len(wl["freqs"]), method=P.METHOD_LOWRANK)
        try:
            for f, hz in enumerate(wl["freqs"]):
                fresh.set_frequency(f, hz)
                fresh.synth_g(f)
            fresh.set_tile_map(final_map)
            mid = (wl["T"] - 1) // 2
            fresh.set_source_port(mid, mid, 1, 2)
            fresh.init_state()
            F64 = fresh.state()["F"]
        finally:
            fresh.close()
        out["F_final_fp64_of_design"] = F64
        out["F_error_vs_fp64"] = abs(r["F_final"] - F64) / F64
    log("[bench] %s: %d evals %d accepts %.3f s (%.0f flips/s), strips %s" % (
        name, r["flip_count"], r["accepts"], r["seconds"], out["value"],
        {k: "%.1f us %.3f" % (v["avg_us"], v["frac"]) for k, v in strips.items()}))
    return out


def plain_ceiling_s(wl, run):
    """Seconds the PLAIN path needs at 8 TB/s for a run's committed work: 16 N (2k + 1 + ne) B per
    evaluation and frequency (k rows of G, k columns of C^-1, S, Gobj) and 48 N^2 B per accept and
    frequency (V C^-1 reads C^-1, the rank-k update reads and writes it)."""
    N = P.layout_size(wl["layout"], wl["T"])
    nf, k = len(wl["freqs"]), 24
    b = nf * (run["flip_count"] * 16.0 * N * (2 * k + 2) + run["accepts"] * 48.0 * N * N)
    return b / (PEAK_HBM_GBS * 1e9)


def cpu_baseline(wl, g0, cinv0, budget_s, gpu_run):
    """Reference-structure FP64 CPU path (the oracle) on a bounded sample: one frequency of the
    same problem, same G and initial state, DBS loop timed for about budget_s seconds.  The
    sample's early phase accepts far more often than a whole run does, and a CPU accept
    (O(k N^2)) costs much more than an evaluation (O(k^2 N)), so the sample's own flips/s is not
    the run's: the sample gives the time per evaluation and per accept at one frequency, and the
    value is the GPU run's evaluation and accept counts priced at those times, at every
    frequency (the reference scores frequencies one after another)."""
    import oracle as O
    T = wl["T"]
    mid = (T - 1) // 2
    o = O.Optimizer(wl["layout"], T, g0, ne=1)
    o.set_tile_map(initial_pattern(T, 0))
    o.set_source_port(mid, mid, 1, 2)
    o.set_cinv(cinv0)
    o.compute_PEinc()
    o.compute_R_matrix(1)
    r = O.dbs_run([o], O.OBJ_S11, [wl["freqs"][0]], DX, DY, DZ, max_seconds=budget_s)
    nf = len(wl["freqs"])
    t_eval = r["eval_s"] / max(1, r["flip_count"])
    t_acc = (r["loop_s"] - r["eval_s"]) / max(1, r["accepts"])
    run_s = nf * (gpu_run["flip_count"] * t_eval + gpu_run["accepts"] * t_acc)
    return dict(value=gpu_run["flip_count"] / run_s, unit="flips/s", cores=O.num_threads(), kind="port",
                projected_run_s=run_s, eval_ms=1e3 * t_eval, accept_ms=1e3 * t_acc,
                sample="oracle (C++ restatement of pngfopt.hpp, FP64, OpenMP over %d threads, its own loops: "
                       "no BLAS library) on 1 of %d frequencies, same G and initial state: %d evaluations and %d "
                       "accepts in %.1f s give %.3f ms per evaluation and %.1f ms per accept; value = the GPU "
                       "run's %d evaluations and %d accepts at those times x %d frequencies (%.0f s)"
                       % (O.num_threads(), nf, r["flip_count"], r["accepts"], r["loop_s"], 1e3 * t_eval,
                          1e3 * t_acc, gpu_run["flip_count"], gpu_run["accepts"], nf, run_s))


def reference_config(max_evals=3000):
    """The reference program's own default (main.cpp:116-117, 154-160, 181-241): 21x21 tiles on the
    shared Yee grid (N = 8064), 5 frequencies 25..35 GHz, ne = 3, the directivity objective,
    MagicHat order with mirror-symmetric 2/4-tile flips, seed -514213645, port at tile (10, 10),
    synthetic G at the default softening, start initial_pattern(21, 0).  LOWRANK and PLAIN on the
    GPU for the first `max_evals` evaluations (one warmup run each, restored from a checkpoint):
    evaluated flips/s, accepts, and whether the two take the same flips."""
    T, ne, mid = 21, 3, 10
    freqs = [25e9, 27.5e9, 30e9, 32.5e9, 35e9]
    out = dict(workload="reference default: 21x21 shared grid, N = 8064, 5 frequencies, ne = 3, directivity "
                        "objective, MagicHat 2/4-tile flips", max_evals=max_evals)
    runs = {}
    for name, method in (("lowrank", P.METHOD_LOWRANK), ("plain", P.METHOD_PLAIN)):
        opt = P.Optimizer(P.LAYOUT_SHARED, T, num_freq=len(freqs), num_eobj=ne, max_flips=4, method=method,
                          objective=P.OBJ_REF, driver=P.DRIVER_REF)
        try:
            for f, hz in enumerate(freqs):
                opt.set_frequency(f, hz)
                opt.synth_g(f)
            opt.set_tile_map(initial_pattern(T, 0))
            opt.set_source_port(mid, mid, 1, 2)
            opt.init_state()
            opt.checkpoint()
            opt.dbs_run(max_evals=max_evals, seed=-514213645)
            opt.restore()
            r = opt.dbs_run(max_evals=max_evals, seed=-514213645)
        finally:
            opt.close()
        runs[name] = r
        out[name] = dict(value=r["flip_count"] / r["seconds"], seconds=r["seconds"], flip_count=r["flip_count"],
                         accepts=r["accepts"], F_final=r["F_final"])
        log("[bench] reference config, %s: %d evals %d accepts %.3f s (%.0f flips/s) F %.6g" % (
            name, r["flip_count"], r["accepts"], r["seconds"], out[name]["value"], r["F_final"]))
    out["speedup_vs_plain"] = out["lowrank"]["value"] / out["plain"]["value"]
    out["same_accept_sequence"] = bool(np.array_equal(runs["lowrank"]["acc_tiles"], runs["plain"]["acc_tiles"]))
    return out


def config_small(steps=5):
    """The reference's CPU-runnable case (7x7 tiles, N = 1182, one frequency at 30 GHz, the
    benchmark's start convention): DBS to convergence on the GPU (LOWRANK, device-driven loop;
    one warmup, then `steps` timed runs from a device checkpoint) and on the CPU (the oracle on
    the device's G, read back through a PLAIN handle, one run).  Same accepted sequence expected; both rates in
    evaluated flips/s."""
    import oracle as O
    wl = WORKLOADS["antenna7x7_1f"]
    T, freqs = wl["T"], wl["freqs"]
    mid = (T - 1) // 2
    tm = initial_pattern(T, 0)
    # the device's G for the oracle, read back from a PLAIN handle (the method that keeps G)
    pl = P.Optimizer(wl["layout"], T, num_freq=1, method=P.METHOD_PLAIN)
    try:
        pl.set_frequency(0, freqs[0])
        pl.synth_g(0)
        pl.set_tile_map(tm)
        pl.set_source_port(mid, mid, 1, 2)
        pl.init_state()
        G = pl.g(0)
    finally:
        pl.close()
    opt = P.Optimizer(wl["layout"], T, num_freq=1, method=P.METHOD_LOWRANK)
    try:
        opt.set_frequency(0, freqs[0])
        opt.synth_g(0)
        opt.set_tile_map(tm)
        opt.set_source_port(mid, mid, 1, 2)
        opt.init_state()
        opt.checkpoint()
        opt.dbs_run()
        secs, runs = 0.0, []
        for _ in range(steps):
            opt.restore()
            t0 = time.perf_counter()
            runs.append(opt.dbs_run())
            secs += time.perf_counter() - t0
    finally:
        opt.close()
    g = runs[-1]
    o = O.Optimizer(wl["layout"], T, G, ne=1)
    assert o.init_state(tm, (mid, mid, 1, 2)) == 0
    c = O.dbs_run([o], O.OBJ_S11, freqs, DX, DY, DZ)
    out = dict(workload="antenna7x7_1f", N=P.layout_size(wl["layout"], T),
               gpu=dict(value=sum(r["flip_count"] for r in runs) / secs, seconds=secs / steps,
                        flip_count=g["flip_count"], accepts=g["accepts"], F_final=g["F_final"]),
               cpu_oracle=dict(value=c["flip_count"] / c["loop_s"], seconds=c["loop_s"], flip_count=c["flip_count"],
                               accepts=c["accepts"], F_final=c["F_final"], cores=O.num_threads(),
                               kind="port (C++ restatement of pngfopt.hpp, FP64, OpenMP, no BLAS library)"),
               same_accept_sequence=bool(np.array_equal(g["acc_tiles"], c["acc_tiles"]) and
                                         np.array_equal(g["acc_eval"], c["acc_eval"])),
               F_final_rel_diff=abs(g["F_final"] - c["F_final"]) / c["F_final"])
    log("[bench] 7x7 x 1: GPU %d evals %d accepts %.2f ms (%.0f flips/s); CPU oracle %.1f ms (%.0f flips/s, %d "
        "threads); same sequence %s" % (g["flip_count"], g["accepts"], 1e3 * out["gpu"]["seconds"], out["gpu"]["value"],
                                        1e3 * c["loop_s"], out["cpu_oracle"]["value"], O.num_threads(),
                                        out["same_accept_sequence"]))
    return out


def copy_probe():
    """Three device-to-device copies of 1 GiB (the runtime's copy kernel): the reference reading
    for the HBM byte counters (FETCH_SIZE / WRITE_SIZE) that scale the kernels' measured traffic."""
    import ctypes as ct
    try:
        hip = ct.CDLL("libamdhip64.so")
    except OSError:
        hip = ct.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    hip.hipMemset.argtypes = [ct.c_void_p, ct.c_int, ct.c_size_t]
    hip.hipFree.argtypes = [ct.c_void_p]
    n = 1 << 30
    a, b = ct.c_void_p(), ct.c_void_p()
    assert hip.hipMalloc(ct.byref(a), n) == 0 and hip.hipMalloc(ct.byref(b), n) == 0
    assert hip.hipMemset(a, 1, n) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": n, "copies": 3}))


def counter_probe(steps=1500):
    """The workload's column-strip launches in a process without rocSOLVER work, for HBM and MFMA
    counter passes (rocprofv3 --pmc ... --kernel-include-regex lr_skinny): the 21x21-tile,
    9-frequency problem from the empty design (no metal component: C = I, nothing to factor), whose
    init runs the autotuner's strips at factor depth K = 1536 and whose DBS (the first `steps`
    evaluations) accepts flips at K = 24 j, the factor depths of the benchmark's run.
    A strip's work depends on N, K, k and the frequency count only.  Prints one JSON line: the
    dispatches' (K, algorithmic bytes, flops) in order."""
    wl = WORKLOADS["antenna21x21_9f"]
    T, freqs = wl["T"], wl["freqs"]
    opt = P.Optimizer(wl["layout"], T, num_freq=len(freqs), method=P.METHOD_LOWRANK)
    for f, hz in enumerate(freqs):
        opt.set_frequency(f, hz)
        opt.synth_g(f)
    mid = (T - 1) // 2
    opt.set_tile_map(np.zeros(T * T, dtype=np.int32))
    opt.set_source_port(mid, mid, 1, 2)
    opt.init_state()
    tune = opt.autotune_result()
    # host-driven loop: one strip dispatch per accept, no graph replay under the counter passes
    r = opt.dbs_run(max_evals=steps, loop=P.LOOP_HOST)
    N, nf, k = opt.N, len(freqs), 24
    # the run converges before the benchmark's accept count: further explicit accepts (single
    # tiles in raster order, the port tile skipped) continue the factor depths K = 24 j
    extra = max(0, HEADLINE_ACCEPTS + 8 - r["accepts"])
    tiles = [t for t in range(T * T) if t != mid * T + mid]
    for i in range(extra):
        opt.accept([tiles[i % len(tiles)]])
    opt.state()  # a read back: the stream has drained
    tune_K = min(opt.factor_capacity()[0] - k, 64 * k)

    def algo(K):
        return dict(K=K, bytes=16.0 * nf * (N * K + 2 * N * k + K * k), flops=8.0 * nf * N * K * k)
    # the autotuner times every variant with 3 dispatches each, in variant order; the run then
    # dispatches the chosen one once per accept
    out = dict(N=N, freqs=nf, k=k, autotune_variants=len(tune["ms"]), autotune_chosen=tune["chosen"],
               autotune_K=tune_K, autotune=algo(tune_K), accepts=[algo(k * j) for j in range(r["accepts"] + extra)],
               flip_count=r["flip_count"], dbs_accepts=r["accepts"])
    print(json.dumps(out))


def _pmc_csv(path):
    """rocprofv3 --pmc ... -f csv output under `path`: per dispatch (in order) its kernel name and
    counter values (summed over the rows of one counter)."""
    import csv
    import glob
    files = glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True)
    if len(files) != 1:
        raise SystemExit("expected one counter_collection.csv under %s, found %d" % (path, len(files)))
    per = {}
    with open(files[0]) as fh:
        for row in csv.DictReader(fh):
            d = per.setdefault(int(row["Dispatch_Id"]), {
                "name": row["Kernel_Name"], "ns": int(row["End_Timestamp"]) - int(row["Start_Timestamp"])})
            d[row["Counter_Name"]] = d.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    return [per[i] for i in sorted(per)]


def counter_summary(path):
    """Summary of the counter passes over --counter-probe and --copy-probe, read from `path`:
    probe.json (the probe's stdout) and the rocprofv3 CSV folders pmc_copy_fetch, pmc_copy_write,
    pmc_strip_fetch, pmc_strip_write and pmc_strip_mfma.  HBM bytes = FETCH_SIZE / WRITE_SIZE
    (KiB) x 1024 x the factor that makes the 1 GiB copies read and write 1 GiB.  Prints a text
    table and, last, one JSON line with the per-launch figures over the first L dispatches of the
    run (the benchmark's L = 183 launches and all of the probe's)."""
    with open(os.path.join(path, "probe.json")) as fh:
        probe = json.load(fh)
    cf, cw = _pmc_csv(os.path.join(path, "pmc_copy_fetch")), _pmc_csv(os.path.join(path, "pmc_copy_write"))
    gib_kib = float(1 << 20)
    f_fetch = gib_kib * len(cf) / sum(d["FETCH_SIZE"] for d in cf)
    f_write = gib_kib * len(cw) / sum(d["WRITE_SIZE"] for d in cw)
    print("copy calibration (%d x 1 GiB %s): FETCH_SIZE %.1f KiB per copy -> factor %.3f; WRITE_SIZE %.1f KiB "
          "-> factor %.3f" % (len(cf), cf[0]["name"], gib_kib / f_fetch, f_fetch, gib_kib / f_write, f_write))
    sf, sw, sm = (_pmc_csv(os.path.join(path, n)) for n in ("pmc_strip_fetch", "pmc_strip_write", "pmc_strip_mfma"))
    nv, N, nf, k = probe["autotune_variants"], probe["N"], probe["freqs"], probe["k"]
    acc = probe["accepts"]
    if not (len(sf) == len(sw) == len(sm) == 3 * nv + len(acc)):
        raise SystemExit("dispatch counts %d/%d/%d, expected %d" % (len(sf), len(sw), len(sm), 3 * nv + len(acc)))
    rows = []
    for i, (a, b, c) in enumerate(zip(sf, sw, sm)):
        al = probe["autotune"] if i < 3 * nv else acc[i - 3 * nv]
        rows.append(dict(K=al["K"], rd=a["FETCH_SIZE"] * 1024 * f_fetch, wr=b["WRITE_SIZE"] * 1024 * f_write,
                         alg=al["bytes"], algf=al["flops"], mfma_flops=c["SQ_INSTS_VALU_MFMA_MOPS_F64"] * 512,
                         busy=c["SQ_VALU_MFMA_BUSY_CYCLES"], active=c["GRBM_GUI_ACTIVE"], util=c["MfmaUtil"] / 100,
                         ns=c["ns"], name=a["name"]))
    wr_alg = 16.0 * nf * N * k
    print("\ncolumn strip, N = %d, %d frequencies, k = %d\nautotuner: 3 dispatches per variant at K = %d "
          "(algorithmic read %.0f KiB, write %.0f KiB)" % (N, nf, k, probe["autotune_K"],
                                                         (probe["autotune"]["bytes"] - wr_alg) / 1024, wr_alg / 1024))
    print("  variant  read/alg  write/alg  MFMA flops/alg  MfmaUtil  kernel")
    for i in range(3 * nv):
        r = rows[i]
        nm = r["name"]
        nm = nm[nm.find("<"):nm.find(">") + 1] if "<" in nm else nm
        print("  %d.%d      %.4f    %.4f     %.4f          %.3f     %s" % (
            i // 3, i % 3, r["rd"] / (r["alg"] - wr_alg), r["wr"] / wr_alg, r["mfma_flops"] / r["algf"], r["util"], nm))
    dbs = rows[3 * nv:]
    print("\nDBS from the empty design, chosen variant %d: %d strips at K = 0, 24, ..., %d" % (
        probe["autotune_chosen"], len(dbs), dbs[-1]["K"]))
    print("  K      (read+write)/alg  MFMA flops/alg  MfmaUtil  us     GHz")
    for r in dbs[::max(1, len(dbs) // 16)]:
        print("  %5d  %.4f            %.4f          %.3f     %6.1f  %.3f" % (
            r["K"], (r["rd"] + r["wr"]) / r["alg"], r["mfma_flops"] / max(r["algf"], 1.0), r["util"], r["ns"] * 1e-3,
            r["active"] / 8.0 / r["ns"]))
    fit = dict(N=N, freqs=nf, k=k, variant=probe["autotune_chosen"])
    # the benchmark's run launches the strip once per accept at K = 0, 24, ..., 24 (L - 1): the
    # probe's first L dispatches are the same launches
    fit["prefix"] = {}
    for L in sorted({L for L in (HEADLINE_ACCEPTS, len(dbs)) if L <= len(dbs)}):
        sel = dbs[:L]
        t = sum(r["rd"] + r["wr"] for r in sel)
        u = sum(r["util"] * r["active"] for r in sel) / sum(r["active"] for r in sel)
        alg = sum(r["alg"] for r in sel)
        fl = sum(r["mfma_flops"] for r in sel) / sum(r["algf"] for r in sel)
        # GRBM_GUI_ACTIVE counts the cycles of each of the 8 XCDs: / 8 / dispatch time = clock
        ghz = sum(r["active"] for r in sel) / 8.0 / sum(r["ns"] for r in sel)
        fit["prefix"][str(L)] = dict(bytes_per_launch=t / L, alg_bytes_per_launch=alg / L, mfma_util=u,
                                     mfma_flops_per_alg=fl, clock_ghz=ghz)
        print("first %d strips (K = 0 .. %d): HBM bytes %.4f x algorithmic (%.6g B per launch), MfmaUtil %.3f "
              "(active-cycle weighted), MFMA flops %.4f x algorithmic, clock %.3f GHz (GRBM_GUI_ACTIVE / 8 / "
              "dispatch time)" % (L, sel[-1]["K"], t / alg, t / L, u, fl, ghz))
    print(json.dumps(fit))


def trace_summary(path):
    """Summary of a rocprofv3 --kernel-trace CSV of `bench.py --steps 1 --warmup 0 --no-cpu`:
    the timed step (from its first loop_batch_kernel through the device loop's kernels, up to the
    first copy of 50 us or more: the checkpoint restore that follows the step) and the PLAIN
    phase (every dispatch after the last rocSOLVER dispatch, the PLAIN handle's init).  Per
    kernel: dispatches, counted dispatches (column strips under 10 us are iterations without an
    accept, which return at once), average / min / max us and total ms."""
    import collections
    import csv
    rows = []
    with open(path) as fh:
        for r in csv.DictReader(fh):
            n = r["Kernel_Name"].replace("void ", "").replace("pngf::", "").replace("(anonymous namespace)::", "")
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), n.split("(")[0][:60]))
    rows.sort()
    loop = ("lr_eval_kernel", "loop_control_kernel", "lr_prep_kernel", "lr_gather", "lr_skinny_mfma_kernel",
            "lr_finish_kernel", "loop_batch_kernel", "__amd_rocclr_copyBuffer", "mask_update_kernel")

    def table(sel, title):
        st = collections.OrderedDict()
        for s0, e0, n in sel:
            d = (e0 - s0) / 1e3
            v = st.setdefault(n, [0, 0, 0.0, 1e30, 0.0])
            v[0] += 1
            if n.startswith("lr_skinny") and d < 10.0:
                continue
            v[1] += 1
            v[2] += d
            v[3] = min(v[3], d)
            v[4] = max(v[4], d)
        print(title)
        print("    %-60s %10s %8s %10s %9s %9s %10s" % ("kernel", "dispatches", "counted", "avg_us", "min_us",
                                                       "max_us", "total_ms"))
        for n, v in sorted(st.items(), key=lambda kv: -kv[1][2]):
            if v[1]:
                print("    %-60s %10d %8d %10.1f %9.1f %9.1f %10.2f" % (n, v[0], v[1], v[2] / v[1], v[3], v[4],
                                                                       v[2] / 1e3))
        print()

    i0 = next(i for i, r in enumerate(rows) if r[2].startswith("loop_batch_kernel"))
    i1 = i0 + 1
    while (i1 < len(rows) and any(rows[i1][2].startswith(x) for x in loop) and
           not (rows[i1][2].startswith("__amd_rocclr_copyBuffer") and rows[i1][1] - rows[i1][0] >= 50000)):
        i1 += 1
    win = rows[i0:i1]
    print("timed step window: %d dispatches, %.1f ms (first scoring launch to last accept kernel)" % (
        len(win), (win[-1][1] - win[0][0]) / 1e6))
    table(win, "## Timed step")
    solver = [i for i, r in enumerate(rows) if "getr" in r[2] or "rocsolver" in r[2].lower() or "getf2" in r[2]]
    if solver:
        plain = rows[solver[-1] + 1:]
        print("PLAIN phase: %d dispatches after the PLAIN handle's init, %.1f ms of kernel time" % (
            len(plain), sum(e0 - s0 for s0, e0, n in plain) / 1e6))
        table(plain, "## PLAIN runs (two DBS runs: one for the rate, one event-timed)")


def load_probe(nfreq=3):
    """acmefdfd-layout G files (raw column-major complex128, M = N, 21x21 tiles: 1.79 GB each)
    through pngf_load_g, the reader that replaces pngfopt.hpp:121-184: the files are written
    from the host synthesiser to a temporary directory, then loaded into one handle, one
    frequency after another, and the bytes per second of the loads are printed as one JSON line,
    beside the rate of a plain host read of the same files into memory.  The files were just
    written, so they come from the page cache: this is the loader's host-to-device rate, not the
    disk's."""
    import shutil
    import tempfile
    wl = WORKLOADS["antenna21x21_9f"]
    T, layout, freqs = wl["T"], wl["layout"], wl["freqs"][:nfreq]
    N = P.layout_size(layout, T)
    tmp = tempfile.mkdtemp(prefix="pngf_g_")
    try:
        paths = []
        for f, hz in enumerate(freqs):
            G = P.synth_g_host(layout, T, 1, hz)
            path = os.path.join(tmp, "Gmat_sub_%02d.bin" % (f + 1))
            G.ravel(order="K").tofile(path)
            paths.append(path)
            del G
        nbytes = sum(os.path.getsize(p) for p in paths)
        t0 = time.perf_counter()
        for p in paths:
            with open(p, "rb") as fh:
                np.fromfile(fh, dtype=np.complex128)
        read_s = time.perf_counter() - t0
        opt = P.Optimizer(layout, T, num_freq=len(freqs), method=P.METHOD_LOWRANK)
        try:
            for f, hz in enumerate(freqs):
                opt.set_frequency(f, hz)
            t0 = time.perf_counter()
            for f, p in enumerate(paths):
                opt.load_g(f, p)
            load_s = time.perf_counter() - t0
        finally:
            opt.close()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    out = {"metric": "G file load rate (acmefdfd layout, page-cached files, host to device)",
           "value": nbytes / load_s / 1e9, "unit": "GB/s", "files": len(paths), "N": N, "bytes": nbytes,
           "seconds": load_s, "host_read_GBps": nbytes / read_s / 1e9, "data": "synthetic"}
    log("[load] %d files, %.2f GB: pngf_load_g %.3f s (%.1f GB/s); host read alone %.1f GB/s" % (
        len(paths), nbytes / 1e9, load_s, out["value"], out["host_read_GBps"]))
    print(json.dumps(out))


def bench_multi(workers=4):
    """16 DBS instances of uneven size (11..25 tiles, 5 frequencies) on one GPU at once.
    Prints its own JSON line: total committed evaluations / wall seconds of the concurrent
    run (pngf_dbs_run_many: host worker threads, largest instances first)."""
    m = MULTI16
    opts = []
    t0 = time.perf_counter()
    # largest first: the factor capacity each handle takes at init is capped by free memory
    for T in sorted(m["sizes"], reverse=True):
        for s in m["seeds"]:
            o = P.Optimizer(m["layout"], T, num_freq=len(m["freqs"]), method=P.METHOD_LOWRANK,
                            max_deferred=m["max_deferred"][T])
            for f, hz in enumerate(m["freqs"]):
                o.set_frequency(f, hz)
                o.synth_g(f)
            o.set_tile_map(initial_pattern(T, s))
            o.set_source_port((T - 1) // 2, (T - 1) // 2, 1, 2)
            o.init_state()
            opts.append(o)
            log("[multi] init %dx%d seed %d (%.1f s)" % (T, T, s, time.perf_counter() - t0))
    init_s = time.perf_counter() - t0
    t1 = time.perf_counter()
    rs = P.dbs_run_many(opts, workers=workers)
    log("[multi] %d host workers" % workers)
    conc_s = time.perf_counter() - t1
    flips = sum(r["flip_count"] for r in rs)
    log("[multi] %d evaluations, %d accepts, %d folds in %.3f s" % (
        flips, sum(r["accepts"] for r in rs), sum(r["flushes"] for r in rs), conc_s))
    out = {"metric": "multi-instance DBS evaluated flips/s (16 instances, 11..25 tiles, 5 freqs)",
           "value": flips / conc_s, "unit": "flips/s", "n_gpus": 1, "seconds": conc_s, "init_s": init_s,
           "flip_count": flips, "accepts": sum(r["accepts"] for r in rs),
           "folds": sum(r["flushes"] for r in rs),
           "instances": [dict(T=opts[i].T, flips=r["flip_count"], accepts=r["accepts"], seconds=r["seconds"],
                              flushes=r["flushes"], F_final=r["F_final"]) for i, r in enumerate(rs)],
           "sum_instance_seconds": sum(r["seconds"] for r in rs), "dtype": "f64", "data": "synthetic"}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="antenna21x21_9f", choices=sorted(WORKLOADS) + ["multi16"])
    ap.add_argument("--method", default="lowrank", choices=["lowrank", "plain"])
    ap.add_argument("--max-deferred", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 GPU comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--tiles", type=int, default=0, help="override the workload's tile count")
    ap.add_argument("--nfreq", type=int, default=0, help="override the workload's frequency count")
    ap.add_argument("--tile-variant", type=int, default=0,
                    help="skinny-product tile shape: 0 = autotune at init, v+1 = use variant v")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run three 1 GiB device-to-device copies (HBM counter calibration)")
    ap.add_argument("--counter-probe", action="store_true",
                    help="only run the column strips of the 21x21 x 9 workload from the empty design (counter passes)")
    ap.add_argument("--counter-summary", metavar="DIR",
                    help="only summarise the counter passes of --counter-probe / --copy-probe found under DIR")
    ap.add_argument("--trace-summary", metavar="CSV",
                    help="only summarise a rocprofv3 kernel_trace.csv of bench.py --steps 1 --warmup 0 --no-cpu")
    ap.add_argument("--workers", type=int, default=4, help="host worker threads of the multi16 workload")
    ap.add_argument("--symmetry", default="auto", choices=["auto", "off"],
                    help="LOWRANK accept: row strip from the column strip when G is symmetric (auto)")
    ap.add_argument("--load-probe", action="store_true",
                    help="only time pngf_load_g on acmefdfd-layout G files (21x21 tiles, 3 frequencies)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the two-strip and non-reciprocal-G runs")
    ap.add_argument("--reference-config", action="store_true",
                    help="run only the reference program's default configuration (LOWRANK and PLAIN) and "
                         "print its result as one JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.counter_probe:
        counter_probe()
        return
    if args.counter_summary:
        counter_summary(args.counter_summary)
        return
    if args.trace_summary:
        trace_summary(args.trace_summary)
        return
    if args.load_probe:
        load_probe()
        return
    if args.reference_config:
        print(json.dumps(reference_config()), flush=True)
        return
    if args.workload == "multi16":
        bench_multi(args.workers)
        return
    wl = dict(WORKLOADS[args.workload])
    if args.tiles:
        wl["T"] = args.tiles
    if args.nfreq:
        lo, hi = min(wl["freqs"]), max(wl["freqs"])
        wl["freqs"] = list(np.linspace(lo, hi, args.nfreq)) if args.nfreq > 1 else [0.5 * (lo + hi)]
    wl_name = args.workload
    if args.tiles or args.nfreq:
        wl_name = "antenna%dx%d_%df" % (wl["T"], wl["T"], len(wl["freqs"]))
    method = P.METHOD_LOWRANK if args.method == "lowrank" else P.METHOD_PLAIN
    # the workload whose whole run the oracle recorded
    default_run = wl_name == "antenna21x21_9f" and wl["max_evals"] == 0

    log("[bench] init %s method=%s" % (wl_name, args.method))
    opt, init_s = make(wl, method, args.max_deferred, P.SYM_AUTO if args.symmetry == "auto" else P.SYM_OFF,
                       args.tile_variant)
    opt.checkpoint()
    tune = opt.autotune_result() if method == P.METHOD_LOWRANK else None
    sym_used, sym_asym = opt.symmetry()
    log("[bench] init %.2f s, N=%d, tile variant %s, symmetric accept %s (asym %.3g)" % (
        init_s, opt.N, tune, sym_used, sym_asym))
    # HIP events around every column-strip launch of the timed steps (on the stream it runs on,
    # inside the device-driven loop); switched on before the warmup so that the warmup already
    # captures the loop graph with its events, and reset before the timed steps
    profile = not args.no_profile
    opt.profile(profile)
    for w in range(args.warmup):
        r = opt.dbs_run(max_evals=wl["max_evals"])
        opt.restore()
        log("[bench] warmup %d: %d evals %d accepts %.3f s" % (w, r["flip_count"], r["accepts"], r["seconds"]))
    opt.profile(profile)
    flips, secs, runs = 0, 0.0, []
    loop_secs = 0.0
    for s in range(args.steps):
        opt.restore()  # returns with the device idle (stream synchronised)
        t0 = time.perf_counter()
        r = opt.dbs_run(max_evals=wl["max_evals"])  # returns after its final stream synchronise
        dt = time.perf_counter() - t0
        flips += r["flip_count"]
        secs += dt
        loop_secs += r["seconds"]
        runs.append(r)
        log("[bench] step %d: %d evals %d accepts %.3f s (loop %.3f s) F %.6g -> %.6g" % (
            s, r["flip_count"], r["accepts"], dt, r["seconds"], r["F_start"], r["F_final"]))
    value = flips / secs
    last = runs[-1]
    roof, allk, strip_share = None, {}, None
    if profile:
        roof, allk = roofline(opt, tune["chosen"] if tune else None)
        opt.profile(False)
        if roof:
            strip_share = roof["avg_ms"] * roof["launches"] * 1e-3 / loop_secs
            log("[bench] timed steps: %d %s launches, %.1f us average (HIP events)" % (
                roof["launches"], roof["kernel"], roof["avg_ms"] * 1e3))
    # the same start with the best-improvement accept rule (a different accepted sequence)
    opt.restore()
    rb = opt.dbs_run(max_evals=wl["max_evals"], order=P.ORDER_BEST)
    best = dict(value=rb["flip_count"] / rb["seconds"], seconds=rb["seconds"], flip_count=rb["flip_count"],
                accepts=rb["accepts"], F_final=rb["F_final"])
    log("[bench] best-improvement: %d evals %d accepts %.3f s F %.6g" % (rb["flip_count"], rb["accepts"],
                                                                          rb["seconds"], rb["F_final"]))
    # sweep time: one scoring launch over every candidate at every frequency, HIP-event timed
    # (each step of the best-improvement run is one such launch)
    sweep = None
    if profile:
        opt.restore()
        opt.profile(True)
        opt.dbs_run(max_evals=wl["max_evals"], order=P.ORDER_BEST)
        ev = opt.kernel_stats(0)
        opt.profile(False)
        if ev["launches"]:
            ncand = rb["flip_count"] // max(1, rb["accepts"] + 1)
            sweep_ms = ev["total_ms"] / ev["launches"]
            # algorithmic traffic of a score: the k x k block and k-vectors, 16 (k^2 + k (ne + 2)) B
            gbps = ev["bytes"] / (ev["total_ms"] * 1e-3) / 1e9
            sweep = dict(candidates=ncand, freqs=len(wl["freqs"]), ms=sweep_ms, launches=ev["launches"],
                         flips_per_s=ncand / (sweep_ms * 1e-3), algo_bytes=ev["bytes"] / ev["launches"],
                         hbm_GBps=gbps, hbm_frac=gbps / PEAK_HBM_GBS)
            log("[bench] sweep of %d candidates x %d frequencies: %.1f us (HIP events)" % (
                ncand, len(wl["freqs"]), sweep_ms * 1e3))
    opt.close()
    del opt

    # the accept path a real (non-reciprocal) G takes: both strips, forced on the symmetric G and
    # taken by itself on a G perturbed at the single-precision level
    variants = None
    if not args.no_variants and method == P.METHOD_LOWRANK:
        orc = default_run
        variants = {"symmetry_off": variant(wl, "symmetry off", P.SYM_OFF, 0.0, last,
                                            oracle_path=ORACLE_RUN if orc else None),
                    "nonreciprocal_1e-7": variant(wl, "G perturbed 1e-7", P.SYM_AUTO, 1e-7, last,
                                                  oracle_path=ORACLE_RUN_NONRECIPROCAL if orc else None),
                    "nonreciprocal_1e-7_projected": variant(wl, "G perturbed 1e-7, projected", P.SYM_PROJECT,
                                                            1e-7, last),
                    "precision_c64": variant(wl, "complex64", P.SYM_AUTO, 0.0, last, P.PREC_C64)}
        # the projected run against the exact (two-strip) run on the same perturbed G
        pr, ex = variants["nonreciprocal_1e-7_projected"], variants["nonreciprocal_1e-7"]
        pr["F_final_vs_unprojected"] = pr["F_final"] / ex["F_final"]
        pr["speedup_vs_unprojected"] = pr["value"] / ex["value"]

    refcfg = None
    if not args.no_variants and method == P.METHOD_LOWRANK:
        refcfg = reference_config()

    plain = None
    g0 = cinv0 = None
    if not args.no_plain or not args.no_cpu:
        popt, pinit = make(wl, P.METHOD_PLAIN)
        if not args.no_cpu:
            g0 = popt.g(0)
            cinv0 = popt.cinv(0)
        if not args.no_plain:
            popt.checkpoint()
            pr = popt.dbs_run(max_evals=wl["max_evals"])
            # a second run with event timing for the kernel breakdown (its accepts synchronise)
            popt.restore()
            popt.profile(True)
            popt.dbs_run(max_evals=wl["max_evals"])
            pk = {P.K_NAMES[i]: popt.kernel_stats(i) for i in range(6)}
            popt.profile(False)
            ceil_s = plain_ceiling_s(wl, pr)
            plain = dict(value=pr["flip_count"] / pr["seconds"], seconds=pr["seconds"],
                         flip_count=pr["flip_count"], accepts=pr["accepts"], F_final=pr["F_final"],
                         speculative_evals=pr["speculative_evals"], init_s=pinit,
                         same_accept_sequence=bool(np.array_equal(pr["acc_tiles"], last["acc_tiles"])),
                         vs_oracle_run=oracle_run_check(pr) if default_run else None,
                         hbm_ceiling_s=ceil_s, hbm_ceiling_value=pr["flip_count"] / ceil_s,
                         kernels={k: dict(v, avg_ms=v["total_ms"] / v["launches"] if v["launches"] else 0.0,
                                          GBps=v["bytes"] / (v["total_ms"] * 1e-3) / 1e9 if v["total_ms"] else 0.0)
                                  for k, v in pk.items() if v["launches"]})
            log("[bench] plain: %d evals %d accepts %.3f s F %.6g (HBM ceiling %.3f s)" % (
                pr["flip_count"], pr["accepts"], pr["seconds"], pr["F_final"], ceil_s))
        popt.close()
        del popt
    cpu = None
    small = None
    if not args.no_cpu:
        log("[bench] cpu baseline (%.0f s budget)" % args.cpu_seconds)
        cpu = cpu_baseline(wl, g0, cinv0, args.cpu_seconds, last)
        del g0, cinv0
        small = config_small()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "flips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps,
        "dbs_loop_ms_per_step": 1e3 * loop_secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl_name, "tiles": wl["T"], "N": P.layout_size(wl["layout"], wl["T"]),
                   "freqs": len(wl["freqs"]), "layout": "disjoint" if wl["layout"] == 1 else "shared",
                   "method": args.method, "objective": "mean |S11|^2", "driver": "raster",
                   "max_evals": wl["max_evals"], "symmetry": args.symmetry},
        "flip_count": last["flip_count"],
        "accepts": last["accepts"],
        "F_start": last["F_start"],
        "F_final": last["F_final"],
        "colblk_seconds_share": strip_share,
        "symmetric_accept": {"used": sym_used, "rel_asym": sym_asym},
        "init_s": init_s,
        "autotune": tune,
        "roofline": roof,
        "cpu_baseline": cpu,
        "config_7x7_1f": small,
        "reference_config": refcfg,
        "plain_fp64_gpu": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "speedup_vs_plain_ceiling": (value / plain["hbm_ceiling_value"]) if plain else None,
        "accept_path_variants": variants,
        "F_final_vs_plain": (last["F_final"] / plain["F_final"]) if plain else None,
        "vs_oracle_run": oracle_run_check(last) if default_run else None,
        "best_improvement_variant": best,
        "sweep": sweep,
        "time_to_solution_s": {"first_improvement": last["seconds"], "best_improvement": rb["seconds"],
                               "plain_fp64_gpu": plain["seconds"] if plain else None},
        "kernels": allk,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
