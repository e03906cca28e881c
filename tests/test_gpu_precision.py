"""The complex64 mode (PNGF_PREC_C64, the reference's USE_SINGLE_PRECISION build: pngf-opt/main.cpp:58-72,
pngfopt.hpp:128-156, 285-347, 488-510, linalg.hpp:25-78) against the complex64 build of the oracle, and its
error against the complex128 path.

Bar: the two complex64 implementations (the oracle restates the reference formulation, the GPU keeps
W = (G - I) C^-1 and scores from its blocks) round differently, so they agree to single-precision
accuracy times the conditioning, not bitwise: on the well-conditioned G (softening 0.4 cells,
cond(C) ~ 1e3) every score within RTOL32 of the oracle's, and the same accepted flips.  Against
complex128 the error of each result is measured and stated (printed) with a bound."""
import numpy as np
import pytest

import oracle as O
import pngf_amd as P
from pngf_amd.problem import WORKLOADS, freq_list, initial_pattern

pytestmark = pytest.mark.gpu

RTOL32 = 2e-4
SOFT_WELL = 0.4
DX, DY, DZ = 0.01 / 60, 0.01 / 60, 0.01 / 180
REF_SEED = -514213645


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def pair(layout, T, freqs, ne=1, driver=P.DRIVER_RASTER, objective=P.OBJ_S11, max_flips=1, seed=0,
         soft=SOFT_WELL, symmetry=P.SYM_AUTO, factors=P.FACTORS_AUTO, max_deferred=0):
    Gs = [P.synth_g_host(layout, T, ne, hz, soft=soft) for hz in freqs]
    tm = initial_pattern(T, seed)
    mid = (T - 1) // 2
    gpu = P.Optimizer(layout, T, num_freq=len(freqs), num_eobj=ne, max_flips=max_flips, objective=objective,
                      driver=driver, symmetry=symmetry, factors=factors, max_deferred=max_deferred,
                      precision=P.PREC_C64)
    ors = []
    for f, (hz, G) in enumerate(zip(freqs, Gs)):
        gpu.set_frequency(f, hz)
        gpu.set_g(f, G)
        o = O.Optimizer(layout, T, G, ne=ne, max_flips=max_flips, c64=True)
        assert o.init_state(tm, (mid, mid, 1, 2)) == 0
        ors.append(o)
    gpu.set_tile_map(tm)
    gpu.set_source_port(mid, mid, 1, 2)
    gpu.init_state()
    return gpu, ors


def test_c64_needs_lowrank():
    with pytest.raises(P.PNGFError) as e:
        P.Optimizer(P.LAYOUT_DISJOINT, 5, method=P.METHOD_PLAIN, precision=P.PREC_C64)
    assert e.value.code == -1


@pytest.mark.parametrize("layout", [P.LAYOUT_DISJOINT, P.LAYOUT_SHARED])
def test_c64_scores_match_c64_oracle(layout):
    """Init state and the score of every single-tile flip, 7x7 tiles at 2 frequencies."""
    T = 7
    freqs = freq_list(2, 29e9, 31e9)
    gpu, ors = pair(layout, T, freqs)
    E = np.array([o.scalars()["Eobj"] for o in ors])
    assert rel(gpu.state()["Eobj"], E) < RTOL32
    cands = [t for t in range(T * T) if t != (T - 1) // 2 * (T + 1)]
    F, e = gpu.eval_batch([[t] for t in cands])
    worst = 0.0
    for i, t in enumerate(cands):
        ref = np.array([o.compute_updated_Eobj([t]) for o in ors])
        worst = max(worst, rel(e[i], ref))
    print("[c64] %s: worst score error against the complex64 oracle %.2e" % (layout, worst))
    assert worst < RTOL32
    gpu.close()


@pytest.mark.parametrize("case", ["raster_device_loop", "raster_folds", "reference_driver"])
def test_c64_dbs_matches_c64_oracle(case):
    """Same accepted flips at the same evaluations as the complex64 oracle, objective within RTOL32:
    the raster driver on the device-driven loop (FP32-MFMA strips, symmetric W), with the factors
    folded every 3 accepts (complex64 gemm), and the reference driver (MagicHat order, 2/4-tile
    flips up to 96 components, REF objective, ne = 3)."""
    T = 7
    if case == "reference_driver":
        freqs = freq_list(2, 29e9, 31e9)
        gpu, ors = pair(P.LAYOUT_SHARED, T, freqs, ne=3, driver=P.DRIVER_REF, objective=P.OBJ_REF, max_flips=4,
                        seed=1)
        g = gpu.dbs_run(seed=REF_SEED)
        o = O.dbs_run(ors, O.OBJ_REF, freqs, DX, DY, DZ, seed=REF_SEED)
    else:
        freqs = freq_list(3, 28e9, 32e9)
        gpu, ors = pair(P.LAYOUT_DISJOINT, T, freqs, max_deferred=3 if case == "raster_folds" else 0)
        g = gpu.dbs_run()
        o = O.dbs_run(ors, O.OBJ_S11, freqs, DX, DY, DZ)
        if case == "raster_folds":
            assert g["flushes"] > 0
    print("[c64] %s: %d accepts, F %.7g (oracle %.7g)" % (case, g["accepts"], g["F_final"], o["F_final"]))
    assert g["accepts"] == o["accepts"] and g["accepts"] > 0
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert np.array_equal(g["acc_eval"], o["acc_eval"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL32
    gpu.close()


def test_c64_with_symmetric_projection():
    """The complex64 mode with PNGF_SYM_PROJECT on a G perturbed non-reciprocally at 1e-7: the
    projection is applied to the complex128 G before the cast, the accept uses W = W^T, and the
    scores and the DBS run follow the complex64 oracle on (G + G^T) / 2 (formed on the host with
    the same sums)."""
    T = 7
    freqs = freq_list(2, 29e9, 31e9)
    N = P.layout_size(P.LAYOUT_DISJOINT, T)
    Gs = [P.perturb_block(P.synth_g_host(P.LAYOUT_DISJOINT, T, 1, hz, soft=SOFT_WELL), N, 1e-7, seed=f + 1)
          for f, hz in enumerate(freqs)]
    tm = initial_pattern(T, 0)
    gpu = P.Optimizer(P.LAYOUT_DISJOINT, T, num_freq=len(freqs), symmetry=P.SYM_PROJECT, precision=P.PREC_C64)
    ors = []
    for f, (hz, G) in enumerate(zip(freqs, Gs)):
        gpu.set_frequency(f, hz)
        gpu.set_g(f, G)
        o = O.Optimizer(P.LAYOUT_DISJOINT, T, np.asfortranarray(0.5 * (G + G.T)), c64=True)
        assert o.init_state(tm, (3, 3, 1, 2)) == 0
        ors.append(o)
    gpu.set_tile_map(tm)
    gpu.set_source_port(3, 3, 1, 2)
    gpu.init_state()
    try:
        used, asym = gpu.symmetry()
        assert used and 1e-8 < asym < 5e-7
        cands = [t for t in range(T * T) if t != 24]
        F, e = gpu.eval_batch([[t] for t in cands])
        worst = max(rel(e[i], np.array([o.compute_updated_Eobj([t]) for o in ors])) for i, t in enumerate(cands))
        assert worst < RTOL32
        g = gpu.dbs_run()
        o = O.dbs_run(ors, O.OBJ_S11, freqs, DX, DY, DZ)
        assert g["accepts"] == o["accepts"] and g["accepts"] > 0
        assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
        assert rel(g["acc_F"], o["acc_F"]) < RTOL32
    finally:
        gpu.close()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("wl,soft,evals", [("antenna21x21_9f", SOFT_WELL, 1500), ("antenna31x31_7f", 0.0, 0)])
def test_c64_error_against_complex128(wl, soft, evals):
    """The complex64 state after DBS against a complex128 init_state of the same final tile map (the
    FP64 truth of that design), and the complex64 run's final objective against the complex128
    run's from the same start: 21x21 x 9 at softening 0.4 for 1500 evaluations, and the 31x31 x 7
    workload (the precision form, default softening, cond(C) ~ 1e5-1e6) to convergence.  Errors
    printed.  Measured on one MI355X: at softening 0.4 the maintained objective is 4.7e-6 from the
    FP64 objective of its design (bound 1e-3); at the default softening single precision fails: the
    c64 run converges to F = 0.0561 where the FP64 objective of that design is 0.080 (30 % off,
    Eobj wrong by more than its size), so the complex64 mode is not fit for the near-PEC regime and
    FP64 stays the default.  The bound there (error above 1e-2) records that finding."""
    w = WORKLOADS[wl]
    T, freqs = w["T"], w["freqs"]

    def build(tm, precision):
        o = P.Optimizer(P.LAYOUT_DISJOINT, T, num_freq=len(freqs), synth_soft=soft, precision=precision)
        for f, hz in enumerate(freqs):
            o.set_frequency(f, hz)
            o.synth_g(f)
        o.set_tile_map(tm)
        o.set_source_port((T - 1) // 2, (T - 1) // 2, 1, 2)
        o.init_state()
        return o

    tm0 = initial_pattern(T, 0)
    a = build(tm0, P.PREC_C64)
    try:
        r32 = a.dbs_run(max_evals=evals)
        s32, tm = a.state(), a.tile_map()
    finally:
        a.close()
    b = build(tm, P.PREC_C128)
    try:
        s64 = b.state()
    finally:
        b.close()
    err_state = abs(s32["F"] - s64["F"]) / s64["F"]
    c = build(tm0, P.PREC_C128)
    try:
        r64 = c.dbs_run(max_evals=evals)
    finally:
        c.close()
    err_run = abs(r32["F_final"] - r64["F_final"]) / r64["F_final"]
    same = np.array_equal(r32["acc_tiles"], r64["acc_tiles"])
    print("[c64] %s: %d accepts (FP64 %d, same sequence %s), F %.7g vs FP64 run %.7g (rel %.2e); objective of "
          "the c64 state vs FP64 init of its design: rel %.2e, Eobj rel %.2e" % (
              wl, r32["accepts"], r64["accepts"], same, r32["F_final"], r64["F_final"], err_run, err_state,
              rel(s32["Eobj"], s64["Eobj"])))
    assert r32["accepts"] > 20
    if soft > 0:
        assert err_state < 1e-3
    else:
        assert r32["converged"] and r64["converged"] and err_state > 1e-2
