"""GPU parity: the HIP library against the CPU oracle on the same G and seed.

Bar (FP64): newEobj, Eobj and the objective within 1e-10 relative of the oracle,
and the identical accepted-flip sequence (tile and evaluation index) in DBS.

The 1e-10 bar is checked on a synthetic G with softening radius 0.4 cells, where
cond(C) ~ 1e3.  The default synthetic G (softening 1 cell, the near-PEC regime)
has cond(C) ~ 1e5..1e6; there two FP64 implementations that round differently
agree only to ~cond * eps * (accepts), so that case is checked for the identical
accepted sequence and the objective within RTOL_ILL.
"""
import os

import numpy as np
import pytest

import oracle as O
import pngf_amd as P
from pngf_amd.problem import freq_list, initial_pattern

pytestmark = pytest.mark.gpu

RTOL = 1e-10
RTOL_ILL = 1e-5
SOFT_WELL = 0.4
DX, DY, DZ = 0.01 / 60, 0.01 / 60, 0.01 / 180


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def setup_pair(layout, T, freqs, method, ne=1, driver=P.DRIVER_RASTER, objective=P.OBJ_S11, max_flips=1,
               max_deferred=0, seed=0, tm=None, soft=SOFT_WELL, tile_variant=0, symmetry=P.SYM_AUTO, Gs=None,
               factors=P.FACTORS_AUTO):
    if Gs is None:
        Gs = [P.synth_g_host(layout, T, ne, hz, soft=soft) for hz in freqs]
    if tm is None:
        tm = initial_pattern(T, seed)
    mid = (T - 1) // 2
    gpu = P.Optimizer(layout, T, num_freq=len(freqs), num_eobj=ne, max_flips=max_flips, method=method,
                      objective=objective, driver=driver, max_deferred=max_deferred, tile_variant=tile_variant,
                      symmetry=symmetry, factors=factors)
    ors = []
    for f, (hz, G) in enumerate(zip(freqs, Gs)):
        gpu.set_frequency(f, hz)
        gpu.set_g(f, G)
        o = O.Optimizer(layout, T, G, ne=ne, max_flips=max_flips)
        assert o.init_state(tm, (mid, mid, 1, 2)) == 0
        ors.append(o)
    gpu.set_tile_map(tm)
    gpu.set_source_port(mid, mid, 1, 2)
    gpu.init_state()
    return gpu, ors, tm


def oracle_state(ors):
    E = np.array([o.scalars()["Eobj"] for o in ors])
    V = np.array([o.scalars()["Vinc"] for o in ors])
    return E, V


def candidates(layout, T):
    port = set(O.layout_port(layout, T, (T - 1) // 2, (T - 1) // 2).tolist())
    out = []
    for t in range(T * T):
        tm = np.zeros(T * T, dtype=np.int32)
        tm[t] = 1
        comps = set(np.nonzero(O.layout_mask(layout, T, tm))[0].tolist())
        if not comps & port:
            out.append(t)
    return out


@pytest.mark.parametrize("layout", [P.LAYOUT_DISJOINT, P.LAYOUT_SHARED])
@pytest.mark.parametrize("method", [P.METHOD_LOWRANK, P.METHOD_PLAIN])
def test_init_and_single_flips(layout, method):
    T = 7
    freqs = [30e9]
    gpu, ors, tm = setup_pair(layout, T, freqs, method)
    st = gpu.state()
    E, V = oracle_state(ors)
    assert rel(st["Eobj"], E) < RTOL
    assert rel(st["Vinc"], V) < RTOL
    cands = candidates(layout, T)
    F, e = gpu.eval_batch([[t] for t in cands])
    for i, t in enumerate(cands):
        ref = np.array([o.compute_updated_Eobj([t]) for o in ors])
        assert rel(e[i], ref) < RTOL, (t, e[i], ref)
        Fo, _, _ = O.objective(O.OBJ_S11, ref, V, None, freqs, DX, DY, DZ)
        assert abs(F[i] - Fo) <= RTOL * abs(Fo)


@pytest.mark.parametrize("method", [P.METHOD_LOWRANK, P.METHOD_PLAIN])
def test_accept_sequence(method):
    T = 7
    freqs = freq_list(3, 28e9, 32e9)
    layout = P.LAYOUT_DISJOINT
    gpu, ors, tm = setup_pair(layout, T, freqs, method)
    cands = candidates(layout, T)
    for t in [3, 17, 30, 17, 44, 8]:
        for o in ors:
            o.compute_updated_Eobj([t])
            o.accept()
        gpu.accept([t])
        E, V = oracle_state(ors)
        assert rel(gpu.state()["Eobj"], E) < RTOL
    assert np.array_equal(gpu.tile_map(), ors[0].tile_map())
    F, e = gpu.eval_batch([[t] for t in cands])
    for i, t in enumerate(cands):
        ref = np.array([o.compute_updated_Eobj([t]) for o in ors])
        assert rel(e[i], ref) < RTOL


@pytest.mark.parametrize("layout,flips", [
    # disjoint: 24 components a tile, so k = 24, 48, 96 (each tile-shape bucket of the MFMA
    # kernels); N = 1182 is not a multiple of their 16-, 32- or 128-wide tiles
    (P.LAYOUT_DISJOINT, [[3], [17, 30], [17, 30], [8, 9, 15, 16], [40], [1, 2, 45, 47], [1, 2, 45, 47]]),
    # shared edges: k = 12 .. 24 for single tiles
    (P.LAYOUT_SHARED, [[3], [17], [30], [30], [44], [8]]),
])
def test_plain_cinv_after_accepts(layout, flips):
    """The PLAIN accept (V C^-1 and C^-1 -= CU M on the FP64-MFMA kernels) leaves the whole of C^-1
    equal to the oracle's and to a dense inverse of the current C (rows of G on metal components,
    identity rows elsewhere) after each accept of 1-, 2- and 4-tile flips at 3 frequencies; a
    flip accepted twice in a row (the metal set back as it was) returns C^-1 to its value before
    the pair.  Bars: 1e-10 until a flip undoes the one before it.  That update cancels the
    previous correction on every entry it touched, and the oracle's own C^-1 then moves 2.4e-10
    (first such flip) and 5.2e-10 (second) from the dense inverse on the disjoint sequence (1.2e-12
    before), so from there both are held to 2e-9 of the dense inverse and of each other."""
    T = 7
    freqs = freq_list(3, 28e9, 32e9)
    Gs = [P.synth_g_host(layout, T, 1, hz, soft=SOFT_WELL) for hz in freqs]
    gpu, ors, tm = setup_pair(layout, T, freqs, P.METHOD_PLAIN, max_flips=4, Gs=Gs)
    N = Gs[0].shape[0]
    port = set(O.layout_port(layout, T, 3, 3).tolist())
    tm = tm.copy()
    hist = [[gpu.cinv(f) for f in range(len(freqs))]]
    tol = RTOL
    for n, fl in enumerate(flips):
        tmf = np.zeros(T * T, dtype=np.int32)
        tmf[fl] = 1
        assert not set(np.nonzero(O.layout_mask(layout, T, tmf))[0].tolist()) & port
        if n > 0 and flips[n - 1] == fl:
            tol = 2e-9
        tm[fl] ^= 1
        mask = O.layout_mask(layout, T, tm).astype(bool)
        for o in ors:
            o.compute_updated_Eobj(fl)
            o.accept()
        gpu.accept(fl)
        hist.append([gpu.cinv(f) for f in range(len(freqs))])
        for f, o in enumerate(ors):
            C = np.eye(N, dtype=np.complex128)
            C[mask, :] = Gs[f][mask, :N]
            dense = np.linalg.inv(C)
            assert rel(hist[-1][f], o.cinv()) < tol, (n, fl, f)
            assert rel(hist[-1][f], dense) < tol, (n, fl, f)
            assert rel(o.cinv(), dense) < tol, (n, fl, f)
        E, _ = oracle_state(ors)
        assert rel(gpu.state()["Eobj"], E) < RTOL
        if n > 0 and flips[n - 1] == fl:
            for f in range(len(freqs)):
                assert rel(hist[-1][f], hist[-3][f]) < tol, (n, fl, f)


def run_both(gpu, ors, freqs, objective, seed=0, max_evals=0):
    g = gpu.dbs_run(max_evals=max_evals, seed=seed)
    o = O.dbs_run(ors, objective, freqs, DX, DY, DZ, seed=seed, max_evals=max_evals)
    return g, o


@pytest.mark.parametrize("layout", [P.LAYOUT_DISJOINT, P.LAYOUT_SHARED])
@pytest.mark.parametrize("method", [P.METHOD_LOWRANK, P.METHOD_PLAIN])
def test_dbs_raster_identical_sequence(layout, method):
    T = 7
    freqs = [30e9]
    gpu, ors, tm = setup_pair(layout, T, freqs, method)
    g, o = run_both(gpu, ors, freqs, O.OBJ_S11)
    assert g["accepts"] == o["accepts"] and g["accepts"] > 0
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert np.array_equal(g["acc_eval"], o["acc_eval"])
    assert g["flip_count"] == o["flip_count"]
    assert rel(g["acc_F"], o["acc_F"]) < RTOL
    assert np.array_equal(gpu.tile_map(), o["final_map"])
    assert g["converged"]


@pytest.mark.parametrize("symmetry,factors", [(P.SYM_AUTO, P.FACTORS_AUTO), (P.SYM_AUTO, P.FACTORS_REBUILT),
                                              (P.SYM_OFF, P.FACTORS_AUTO)])
def test_dbs_multifreq_and_flush(symmetry, factors):
    """3 frequencies, deferred factors folded into W every 3 accepts; the accept with the
    row strip taken from the column strip (symmetric G, M stored or rebuilt from A and Lam')
    and with both strips formed."""
    T = 7
    freqs = freq_list(3, 28e9, 32e9)
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, P.METHOD_LOWRANK, max_deferred=3, symmetry=symmetry,
                              factors=factors)
    assert gpu.symmetry()[0] == (symmetry == P.SYM_AUTO)
    assert gpu.factor_capacity() == (3 * 24, factors == P.FACTORS_REBUILT)
    g, o = run_both(gpu, ors, freqs, O.OBJ_S11)
    assert g["flushes"] > 0
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert np.array_equal(g["acc_eval"], o["acc_eval"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL


@pytest.mark.parametrize("method,symmetry", [(P.METHOD_LOWRANK, P.SYM_AUTO), (P.METHOD_LOWRANK, P.SYM_OFF),
                                             (P.METHOD_PLAIN, P.SYM_AUTO)])
def test_dbs_reference_driver(method, symmetry):
    """MagicHat order, mirror-symmetric 2/4-tile flips (k up to 96), REF objective, ne = 3."""
    T = 7
    freqs = freq_list(2, 29e9, 31e9)
    seed = -514213645
    gpu, ors, tm = setup_pair(P.LAYOUT_SHARED, T, freqs, method, ne=3, driver=P.DRIVER_REF,
                              objective=P.OBJ_REF, max_flips=4, seed=1, symmetry=symmetry)
    g, o = run_both(gpu, ors, freqs, O.OBJ_REF, seed=seed)
    assert g["accepts"] == o["accepts"] and g["accepts"] > 0
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert np.array_equal(g["acc_eval"], o["acc_eval"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL


@pytest.mark.timeout(900)
def test_reference_default_configuration():
    """The reference program's own configuration at its own size (main.cpp:116-117, 154-160,
    181-241): 21x21 tiles on the shared Yee grid (N = 8064), 5 frequencies 25..35 GHz, ne = 3
    (port row and two NF2FF rows), the directivity objective, MagicHat order with mirror-
    symmetric 2/4-tile flips (k up to 96), seed -514213645, port at tile (10, 10); softening 0.4.
    The LOWRANK run's first 60 evaluations against the oracle's: same accepted flips at the same
    evaluations, objective to RTOL.  The oracle starts from the C^-1 that a PLAIN handle formed
    on the device (the oracle's own inversion at this N takes minutes on the host) and then
    runs the reference's Woodbury updates from it."""
    T, ne, mid, nev = 21, 3, 10, 60
    freqs = [25e9, 27.5e9, 30e9, 32.5e9, 35e9]
    tm = initial_pattern(T, 1)
    Gs = [P.synth_g_host(P.LAYOUT_SHARED, T, ne, hz, soft=SOFT_WELL) for hz in freqs]

    def handle(method):
        h = P.Optimizer(P.LAYOUT_SHARED, T, num_freq=len(freqs), num_eobj=ne, max_flips=4, method=method,
                        objective=P.OBJ_REF, driver=P.DRIVER_REF)
        for f, (hz, G) in enumerate(zip(freqs, Gs)):
            h.set_frequency(f, hz)
            h.set_g(f, G)
        h.set_tile_map(tm)
        h.set_source_port(mid, mid, 1, 2)
        h.init_state()
        return h

    pl = handle(P.METHOD_PLAIN)
    try:
        assert pl.N == 8064
        cinvs = [pl.cinv(f) for f in range(len(freqs))]
    finally:
        pl.close()
    ors = []
    for G, ci in zip(Gs, cinvs):
        o = O.Optimizer(P.LAYOUT_SHARED, T, G, ne=ne, max_flips=4)
        o.set_tile_map(tm)
        o.set_source_port(mid, mid, 1, 2)
        o.set_cinv(ci)
        o.compute_PEinc()
        o.compute_R_matrix(1)
        ors.append(o)
    del cinvs
    gpu = handle(P.METHOD_LOWRANK)
    try:
        E, V = oracle_state(ors)
        assert rel(gpu.state()["Eobj"], E) < RTOL and rel(gpu.state()["Vinc"], V) < RTOL
        g, o = run_both(gpu, ors, freqs, O.OBJ_REF, seed=-514213645, max_evals=nev)
    finally:
        gpu.close()
    print("[reference config] %d evaluations, %d accepts, F %.10g (oracle %.10g)" % (
        g["flip_count"], g["accepts"], g["F_final"], o["F_final"]))
    assert g["flip_count"] == o["flip_count"] == nev
    assert g["accepts"] == o["accepts"] and g["accepts"] > 0
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert np.array_equal(g["acc_eval"], o["acc_eval"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL


@pytest.mark.parametrize("case", ["raster_s11_lowrank", "raster_s11_plain", "ref_lowrank"])
def test_history_matches_oracle(case):
    """The reference Logger's records (start state, then one per evaluation: tile, decision,
    F, F_best, per-frequency Re S11, Im S11, D) from the GPU loop equal the oracle's: same
    tiles and decisions, values within RTOL."""
    freqs = freq_list(2, 29e9, 31e9)
    if case.startswith("raster"):
        method = P.METHOD_LOWRANK if case.endswith("lowrank") else P.METHOD_PLAIN
        gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, 7, freqs, method)
        mode, seed, cap = O.OBJ_S11, 0, 400
    else:
        gpu, ors, tm = setup_pair(P.LAYOUT_SHARED, 7, freqs, P.METHOD_LOWRANK, ne=3, driver=P.DRIVER_REF,
                                  objective=P.OBJ_REF, max_flips=4, seed=1)
        mode, seed, cap = O.OBJ_REF, -514213645, 400
    g = gpu.dbs_run(seed=seed, max_evals=cap - 1, hist_cap=cap)
    o = O.dbs_run(ors, mode, freqs, DX, DY, DZ, seed=seed, max_evals=cap - 1, hist_cap=cap)
    n = len(o["hist_tile"])
    assert n == o["flip_count"] + 1 and len(g["hist_tile"]) == n and g["flip_count"] == o["flip_count"]
    assert np.array_equal(g["hist_tile"], o["hist_tile"])
    assert np.array_equal(g["hist_dec"], o["hist_dec"])
    assert g["hist_perf"].shape == o["hist_perf"].shape == (n, 2 + 3 * len(freqs))
    assert rel(g["hist_perf"], o["hist_perf"]) < RTOL
    assert np.array_equal(np.nonzero(g["hist_dec"][1:] == 1)[0], g["acc_eval"])


def test_history_file_from_cli(tmp_path):
    """pngf_opt --log 3 writes log_3.txt in the Logger's layout: 'seed records', the start map,
    one line per record (index, tile, decision, 2 + 3 nf values), the final map."""
    import subprocess
    exe = os.path.join(os.path.dirname(P.__file__), "cli", "pngf_opt")
    cmd = [exe, "--tiles", "7", "--layout", "disjoint", "--driver", "raster", "--objective", "s11",
           "--freqs", "29,31", "--log", "3"]
    out = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = (tmp_path / "log_3.txt").read_text().splitlines()
    seed, n = map(int, lines[0].split())
    assert seed == -514213645 and n == len(lines) - 3 and n > 1
    assert len(lines[1]) == 49 and set(lines[1]) <= set("x.")
    assert len(lines[-1]) == 49 and set(lines[-1]) <= set("x.")
    flips = int(out.stdout.split("Converged after ")[1].split()[0])
    assert n == flips + 1
    rec = [l.split() for l in lines[2:-1]]
    assert all(len(r) == 3 + 2 + 3 * 2 for r in rec)
    assert rec[0][:3] == ["0", "-1", "-1"]
    assert [int(r[0]) for r in rec] == list(range(n))


def test_cli_symmetry_project_is_a_no_op_on_a_symmetric_g(tmp_path):
    """The synthetic G is exactly symmetric, so projecting it ((a + a) / 2 = a) changes nothing:
    pngf_opt --symmetry project writes the same history as --symmetry auto, byte for byte, on
    the default (ill-conditioned) synthetic G."""
    import subprocess
    exe = os.path.join(os.path.dirname(P.__file__), "cli", "pngf_opt")
    logs = []
    for i, mode in enumerate(("auto", "project")):
        cmd = [exe, "--tiles", "7", "--layout", "disjoint", "--driver", "raster", "--objective", "s11",
               "--freqs", "29,31", "--symmetry", mode, "--log", str(i)]
        out = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        logs.append((tmp_path / ("log_%d.txt" % i)).read_bytes())
    assert len(logs[0]) > 100 and logs[0] == logs[1]


@pytest.mark.parametrize("method", [P.METHOD_LOWRANK, P.METHOD_PLAIN])
def test_dbs_ill_conditioned_identical_sequence(method):
    """Default synthetic G (cond(C) ~ 1e6): same accepted flips, objective to RTOL_ILL."""
    T = 7
    freqs = freq_list(3, 28e9, 32e9)
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, method, soft=0.0)
    g, o = run_both(gpu, ors, freqs, O.OBJ_S11)
    assert g["accepts"] == o["accepts"] and g["accepts"] > 10
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert np.array_equal(g["acc_eval"], o["acc_eval"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL_ILL


@pytest.mark.parametrize("method", [P.METHOD_LOWRANK, P.METHOD_PLAIN])
def test_empty_design_init(method):
    """No metal component: C = I is its own inverse and init_state skips the factorisation.  The
    state is the closed form (Eobj = 0: PEinc = 0) and scores and DBS follow the oracle, which
    factors the identity."""
    T = 7
    freqs = freq_list(2, 29e9, 31e9)
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, method, tm=np.zeros(T * T, dtype=np.int32))
    assert np.all(gpu.state()["Eobj"] == 0)
    cands = candidates(P.LAYOUT_DISJOINT, T)
    F, e = gpu.eval_batch([[t] for t in cands])
    for i, t in enumerate(cands):
        ref = np.array([o.compute_updated_Eobj([t]) for o in ors])
        assert rel(e[i], ref) < RTOL
    g, o = run_both(gpu, ors, freqs, O.OBJ_S11, max_evals=150)
    assert g["accepts"] == o["accepts"] and g["accepts"] > 0
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL
    gpu.close()


@pytest.mark.parametrize("method", [P.METHOD_LOWRANK, P.METHOD_PLAIN])
def test_checkpoint_restore_repeats_run(method):
    """A device checkpoint taken mid-run (after 20 evaluations, accepts already in the state)
    and restored after a second run went on to convergence: the run from the restored state
    repeats the second one bitwise, its state is the oracle's at the checkpoint (Eobj to RTOL),
    and it is the oracle's continuation from there (same accepted flips at the same
    evaluations, objective to RTOL)."""
    T = 7
    freqs = [30e9]
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, method)
    try:
        g1 = gpu.dbs_run(max_evals=20)
        o1 = O.dbs_run(ors, O.OBJ_S11, freqs, DX, DY, DZ, max_evals=20)
        assert g1["accepts"] == o1["accepts"] > 0
        assert np.array_equal(g1["acc_tiles"], o1["acc_tiles"])
        gpu.checkpoint()
        a = gpu.dbs_run()
        gpu.restore()
        E, _ = oracle_state(ors)
        assert rel(gpu.state()["Eobj"], E) < RTOL
        b = gpu.dbs_run()
        o2 = O.dbs_run(ors, O.OBJ_S11, freqs, DX, DY, DZ)
    finally:
        gpu.close()
    assert np.array_equal(a["acc_tiles"], b["acc_tiles"])
    assert a["acc_F"].tolist() == b["acc_F"].tolist()
    assert b["accepts"] == o2["accepts"] > 0
    assert np.array_equal(b["acc_tiles"], o2["acc_tiles"])
    assert np.array_equal(b["acc_eval"], o2["acc_eval"])
    assert rel(b["acc_F"], o2["acc_F"]) < RTOL


def test_dbs_run_many_matches_sequential():
    """Instances of uneven size run concurrently give the same results as one at a time."""
    freqs = freq_list(2, 29e9, 31e9)
    sizes = [5, 7, 9]
    opts = []
    for i, T in enumerate(sizes):
        o = P.Optimizer(P.LAYOUT_DISJOINT, T, num_freq=len(freqs), method=P.METHOD_LOWRANK, synth_soft=SOFT_WELL)
        for f, hz in enumerate(freqs):
            o.set_frequency(f, hz)
            o.synth_g(f)
        o.set_tile_map(initial_pattern(T, i))
        o.set_source_port((T - 1) // 2, (T - 1) // 2, 1, 2)
        o.init_state()
        o.checkpoint()
        opts.append(o)
    seq = []
    for o in opts:
        seq.append(o.dbs_run())
        o.restore()
    conc = P.dbs_run_many(opts)
    for a, b in zip(seq, conc):
        assert a["accepts"] > 0
        assert np.array_equal(a["acc_tiles"], b["acc_tiles"])
        assert a["acc_F"].tolist() == b["acc_F"].tolist()


@pytest.mark.parametrize("variant", range(6))
def test_every_tile_variant_gives_the_same_run(variant):
    """Each tile shape the autotuner may pick reproduces the oracle's DBS (k = 24)."""
    T = 7
    freqs = freq_list(2, 29e9, 31e9)
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, P.METHOD_LOWRANK, tile_variant=variant + 1)
    g, o = run_both(gpu, ors, freqs, O.OBJ_S11)
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL


def test_tile_variants_are_bitwise_identical():
    """Ill-conditioned G (cond(C) ~ 1e6), where a different rounding changes the accepted
    sequence: every tile shape the autotuner may pick gives bitwise the same run, so timing
    noise in the autotuner cannot change results."""
    T = 7
    freqs = freq_list(3, 28e9, 32e9)
    runs = []
    for variant in range(6):
        gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, P.METHOD_LOWRANK, soft=0.0, tile_variant=variant + 1)
        runs.append(gpu.dbs_run())
        gpu.close()
    for r in runs[1:]:
        assert np.array_equal(r["acc_tiles"], runs[0]["acc_tiles"])
        assert r["acc_F"].tolist() == runs[0]["acc_F"].tolist()


@pytest.mark.parametrize("driver", [P.DRIVER_RASTER, P.DRIVER_REF])
def test_sweep_is_first_improvement_of_sequential_evals(driver):
    """pngf_sweep over a shuffled visit sequence returns the first flip whose objective, scored
    alone, is strictly below F_best (main.cpp:429), with that objective bitwise and
    the number of flips scored; with F_best below every score it returns -1 after the whole
    sequence.  Raster driver: single tiles, scored in batches; reference driver: single tiles
    are ad-hoc flips (scored one at a time) between two of its mirrored pairs."""
    T = 7
    mid = (T - 1) // 2
    freqs = freq_list(2, 29e9, 31e9)
    mf = 1 if driver == P.DRIVER_RASTER else 4
    gpu, ors, tm = setup_pair(P.LAYOUT_SHARED, T, freqs, P.METHOD_LOWRANK, driver=driver, max_flips=mf,
                              objective=P.OBJ_S11)
    rng = np.random.default_rng(5)
    cands = [[int(t)] for t in rng.permutation(T * T) if t != mid * (T + 1)]
    if driver == P.DRIVER_REF:
        cands = cands[:12] + [[mid * T + 0, mid * T + T - 1], [mid * T + 1, mid * T + T - 2]] + cands[12:]
    F = [gpu.eval_batch([c])[0][0] for c in cands]
    F0 = gpu.state()["F"]
    first = next(i for i, x in enumerate(F) if x < F0)
    idx, Fs, ev = gpu.sweep(cands, F0)
    assert (idx, ev) == (first, first + 1) and Fs == F[first]
    idx, Fs, ev = gpu.sweep(cands, min(F) * (1 - 1e-12))
    assert (idx, ev) == (-1, len(cands)) and Fs == min(F) * (1 - 1e-12)
    assert gpu.sweep([], F0) == (-1, F0, 0)
    gpu.close()


@pytest.mark.parametrize("layout", [P.LAYOUT_DISJOINT, P.LAYOUT_SHARED])
@pytest.mark.parametrize("T,nf", [(15, 10), (17, 9)])
def test_wave_scoring_matches_workgroup_scoring(layout, T, nf):
    """All candidates of T x T tiles at 10 or 9 frequencies in one batch (2240 or 2592 pairs:
    scored one pair per wave, the capacitance matrix in registers) and in two batches under 2048
    pairs each (the two-wave workgroup kernel, matrix in LDS): bitwise the same scores, on the
    default synthetic G (cond(C) ~ 1e5, where any change in rounding shows)."""
    gpu = P.Optimizer(layout, T, num_freq=nf, method=P.METHOD_LOWRANK)
    for f, hz in enumerate(freq_list(nf, 26e9, 34e9)):
        gpu.set_frequency(f, hz)
        gpu.synth_g(f)
    gpu.set_tile_map(initial_pattern(T, 3))
    gpu.set_source_port((T - 1) // 2, (T - 1) // 2, 1, 2)
    gpu.init_state()
    cands = [[t] for t in range(T * T) if t != (T - 1) // 2 * (T + 1)]
    h = len(cands) // 2
    F1, e1 = gpu.eval_batch(cands)
    F2a, e2a = gpu.eval_batch(cands[:h])
    F2b, e2b = gpu.eval_batch(cands[h:])
    assert len(cands) * nf >= 2048 and (len(cands) - h) * nf < 2048
    assert np.array_equal(F1, np.concatenate([F2a, F2b]))
    assert np.array_equal(e1, np.concatenate([e2a, e2b]))
    assert np.all(np.isfinite(F1))
    gpu.close()


def test_autotune_reports_variants():
    T = 7
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, [30e9], P.METHOD_LOWRANK)
    r = gpu.autotune_result()
    assert len(r["ms"]) == 6 and 0 <= r["chosen"] < 6
    assert r["ms"][r["chosen"]] == min(r["ms"])


def test_symmetry_detected_for_reciprocal_g():
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, 7, freq_list(2, 29e9, 31e9), P.METHOD_LOWRANK)
    used, asym = gpu.symmetry()
    assert used and asym == 0.0
    plain, _, _ = setup_pair(P.LAYOUT_DISJOINT, 7, [30e9], P.METHOD_PLAIN)
    assert not plain.symmetry()[0]


def test_asymmetric_g_forms_both_strips():
    """A G that is not symmetric (a random non-reciprocal perturbation of 1e-3 of its largest
    entry) is detected, and the accept then forms both strips: same run as the oracle."""
    T = 7
    freqs = freq_list(2, 29e9, 31e9)
    rng = np.random.default_rng(7)
    Gs = []
    for hz in freqs:
        G = P.synth_g_host(P.LAYOUT_DISJOINT, T, 1, hz, soft=SOFT_WELL)
        N = G.shape[0]
        G = G + 1e-3 * np.abs(G).max() * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / N
        Gs.append(np.asfortranarray(G))
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, P.METHOD_LOWRANK, max_deferred=4, Gs=Gs)
    used, asym = gpu.symmetry()
    assert not used and 1e-7 < asym < 1e-2
    g, o = run_both(gpu, ors, freqs, O.OBJ_S11)
    assert g["accepts"] == o["accepts"] and g["accepts"] > 0 and g["flushes"] > 0
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL


def _perturbed_gs(T, freqs, rel_noise):
    N = P.layout_size(P.LAYOUT_DISJOINT, T)
    return [P.perturb_block(P.synth_g_host(P.LAYOUT_DISJOINT, T, 1, hz, soft=SOFT_WELL), N, rel_noise, seed=f + 1)
            for f, hz in enumerate(freqs)]


def _oracles(Gs, T, tm):
    ors = []
    for G in Gs:
        o = O.Optimizer(P.LAYOUT_DISJOINT, T, G)
        assert o.init_state(tm, ((T - 1) // 2, (T - 1) // 2, 1, 2)) == 0
        ors.append(o)
    return ors


def test_projection_runs_the_projected_g():
    """PNGF_SYM_PROJECT on a G perturbed non-reciprocally at 1e-7 of its largest entry (the level
    of acmefdfd's single-precision Schur complement): the reported asymmetry is that of G as
    given, the accept uses W = W^T, and init state, scores and the DBS run are the oracle's on
    (G + G^T) / 2 (formed on the host with the same sums, so the same matrix bitwise).  Against
    the oracle on the unprojected G the objective of the run moves by the projection only:
    printed, and bounded by 1e-5."""
    T = 7
    freqs = freq_list(2, 29e9, 31e9)
    Gs = _perturbed_gs(T, freqs, 1e-7)
    gpu, ors_given, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, P.METHOD_LOWRANK, Gs=Gs,
                                    symmetry=P.SYM_PROJECT)
    used, asym = gpu.symmetry()
    assert used and 1e-8 < asym < 5e-7
    ors = _oracles([np.asfortranarray(0.5 * (G + G.T)) for G in Gs], T, tm)
    E, _ = oracle_state(ors)
    assert rel(gpu.state()["Eobj"], E) < RTOL
    cands = candidates(P.LAYOUT_DISJOINT, T)
    F, e = gpu.eval_batch([[t] for t in cands])
    for i, t in enumerate(cands):
        assert rel(e[i], np.array([o.compute_updated_Eobj([t]) for o in ors])) < RTOL
    g, o = run_both(gpu, ors, freqs, O.OBJ_S11)
    assert g["accepts"] == o["accepts"] and g["accepts"] > 0
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL
    ref = O.dbs_run(ors_given, O.OBJ_S11, freqs, DX, DY, DZ)
    dF = abs(g["F_final"] - ref["F_final"]) / ref["F_final"]
    print("[projection] asym %.2e: %d accepts (unprojected %d, same sequence %s), F %.10g vs %.10g (rel %.2e)" % (
        asym, g["accepts"], ref["accepts"], np.array_equal(g["acc_tiles"], ref["acc_tiles"]), g["F_final"],
        ref["F_final"], dF))
    assert dF < 1e-5
    gpu.close()


def test_projection_declined_beyond_its_tolerance():
    """A G non-reciprocal at 1e-4 of its largest entry is beyond PNGF_SYM_PROJECT's 1e-5: it is
    not projected, the accept forms both strips, and the run is the oracle's on G as given."""
    T = 7
    freqs = freq_list(2, 29e9, 31e9)
    Gs = _perturbed_gs(T, freqs, 1e-4)
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, P.METHOD_LOWRANK, Gs=Gs, symmetry=P.SYM_PROJECT)
    used, asym = gpu.symmetry()
    assert not used and 1e-5 < asym < 1e-3
    g, o = run_both(gpu, ors, freqs, O.OBJ_S11)
    assert g["accepts"] == o["accepts"] and g["accepts"] > 0
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL
    gpu.close()


def test_device_perturbation_matches_host_and_oracle():
    """pngf_perturb_g at 1e-7 of the largest entry (the rounding level of acmefdfd's single-
    precision Schur complement): the device's G equals perturb_block's on the host bitwise, the
    perturbed G is detected as non-reciprocal, the accept forms both strips, and the DBS run is
    the oracle's on the host-perturbed G."""
    T = 7
    layout = P.LAYOUT_DISJOINT
    freqs = freq_list(2, 29e9, 31e9)
    N = P.layout_size(layout, T)
    Gs = [P.synth_g_host(layout, T, 1, hz, soft=SOFT_WELL) for hz in freqs]
    Gp = [P.perturb_block(G, N, 1e-7, seed=f + 1) for f, G in enumerate(Gs)]
    tm = initial_pattern(T, 0)
    plain = P.Optimizer(layout, T, num_freq=len(freqs), method=P.METHOD_PLAIN)
    gpu = P.Optimizer(layout, T, num_freq=len(freqs), method=P.METHOD_LOWRANK)
    for o in (plain, gpu):
        for f, (hz, G) in enumerate(zip(freqs, Gs)):
            o.set_frequency(f, hz)
            o.set_g(f, G)
            o.perturb_g(f, 1e-7, seed=f + 1)
        o.set_tile_map(tm)
        o.set_source_port(3, 3, 1, 2)
        o.init_state()
    for f in range(len(freqs)):
        assert np.array_equal(plain.g(f), Gp[f][:N, :N])
    plain.close()
    used, asym = gpu.symmetry()
    assert not used and 1e-8 < asym < 5e-7
    ors = []
    for G in Gp:
        o = O.Optimizer(layout, T, G)
        assert o.init_state(tm, (3, 3, 1, 2)) == 0
        ors.append(o)
    g, o = run_both(gpu, ors, freqs, O.OBJ_S11)
    assert g["accepts"] == o["accepts"] and g["accepts"] > 0
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert np.array_equal(g["acc_eval"], o["acc_eval"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL
    gpu.close()


@pytest.mark.parametrize("asym,factors", [(False, P.FACTORS_AUTO), (False, P.FACTORS_REBUILT),
                                          (True, P.FACTORS_REBUILT)])
def test_get_w_after_folds_and_pending_factors(asym, factors):
    """W = (G - I) C^-1 read back after accepts that folded the deferred factors into W0 (the
    in-place fold) and left some pending (the fold into a copy): equal to the oracle's
    (G - I) C^-1, row-major, for a symmetric G (M stored or rebuilt) and a non-symmetric one (M
    stored, whatever the request).  (Capacity 2 accepts of 24 components: the 5 accepts fold
    twice and leave one pending.)"""
    T = 7
    freqs = freq_list(2, 29e9, 31e9)
    rng = np.random.default_rng(11)
    Gs = []
    for hz in freqs:
        G = P.synth_g_host(P.LAYOUT_DISJOINT, T, 1, hz, soft=SOFT_WELL)
        if asym:
            N = G.shape[0]
            G = G + 1e-3 * np.abs(G).max() * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / N
        Gs.append(np.asfortranarray(G))
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, P.METHOD_LOWRANK, max_deferred=2, Gs=Gs,
                              factors=factors)
    assert gpu.symmetry()[0] == (not asym)
    assert gpu.factor_capacity()[1] == (not asym and factors == P.FACTORS_REBUILT)
    for t in [3, 17, 30, 44, 8]:
        for o in ors:
            o.compute_updated_Eobj([t])
            o.accept()
        gpu.accept([t])
    for f, G in enumerate(Gs):
        N = ors[f].N
        W_ref = (G[:N, :N] - np.eye(N)) @ ors[f].cinv()
        assert rel(gpu.w(f), W_ref) < RTOL
    E, V = oracle_state(ors)
    assert rel(gpu.state()["Eobj"], E) < RTOL


def test_full_size_state_matches_fresh_init():
    """Bench workload size (21x21 tiles, N = 10590, 9 frequencies; synthetic G with softening
    0.4 cells, cond(C) ~ 1e3): after 1500 evaluations of DBS the maintained state (W0 plus
    deferred factors, symmetric accept) gives the same Eobj and objective as a fresh init_state
    from the final tile map (C inverted again), and a flip accepted twice returns Eobj to where
    it was.  (With the default softening, cond(C) ~ 1e5-1e6, the two agree to ~1e-4 after a
    few dozen accepts: the drift any FP64 update scheme has there.)"""
    from pngf_amd.problem import WORKLOADS
    wl = WORKLOADS["antenna21x21_9f"]
    T, freqs = wl["T"], wl["freqs"]

    def make(tm):
        o = P.Optimizer(P.LAYOUT_DISJOINT, T, num_freq=len(freqs), method=P.METHOD_LOWRANK,
                        synth_soft=SOFT_WELL)
        for f, hz in enumerate(freqs):
            o.set_frequency(f, hz)
            o.synth_g(f)
        o.set_tile_map(tm)
        o.set_source_port((T - 1) // 2, (T - 1) // 2, 1, 2)
        o.init_state()
        return o

    a = make(initial_pattern(T, 0))
    assert a.symmetry()[0]
    r = a.dbs_run(max_evals=1500)
    assert r["accepts"] > 20
    sa = a.state()
    b = make(a.tile_map())
    sb = b.state()
    b.close()
    assert rel(sa["Eobj"], sb["Eobj"]) < 1e-9
    assert abs(sa["F"] - sb["F"]) < 1e-9 * sb["F"]
    t = int(r["acc_tiles"][0])
    a.accept([t])
    a.accept([t])
    assert rel(a.state()["Eobj"], sa["Eobj"]) < 1e-9
    a.close()


@pytest.mark.parametrize("layout,max_evals,factors", [(P.LAYOUT_DISJOINT, 0, P.FACTORS_AUTO),
                                                      (P.LAYOUT_SHARED, 0, P.FACTORS_AUTO),
                                                      (P.LAYOUT_DISJOINT, 60, P.FACTORS_AUTO),
                                                      (P.LAYOUT_DISJOINT, 0, P.FACTORS_REBUILT)])
def test_device_loop_matches_host_loop(layout, max_evals, factors):
    """The device-driven raster loop (batches, selection and accept bookkeeping in kernels) and
    the host-driven one give bitwise the same run, on the ill-conditioned G where any
    difference in rounding would show up as a different accepted sequence."""
    T = 7
    freqs = freq_list(3, 28e9, 32e9)
    gpu, ors, tm = setup_pair(layout, T, freqs, P.METHOD_LOWRANK, soft=0.0, factors=factors)
    assert gpu.factor_capacity()[1] == (factors == P.FACTORS_REBUILT)
    gpu.checkpoint()
    a = gpu.dbs_run(max_evals=max_evals, loop=P.LOOP_HOST)
    map_a, st_a = gpu.tile_map(), gpu.state()
    gpu.restore()
    b = gpu.dbs_run(max_evals=max_evals, loop=P.LOOP_AUTO)
    assert a["accepts"] > 0
    for key in ("flip_count", "accepts", "speculative_evals", "batches", "converged", "F_final"):
        assert a[key] == b[key], key
    assert np.array_equal(a["acc_tiles"], b["acc_tiles"])
    assert np.array_equal(a["acc_eval"], b["acc_eval"])
    assert a["acc_F"].tolist() == b["acc_F"].tolist()
    assert np.array_equal(map_a, gpu.tile_map())
    assert np.array_equal(st_a["Eobj"], gpu.state()["Eobj"])
    gpu.close()


def test_device_loop_profiled_run():
    """With profiling on, the device-driven loop times every column strip that ran (one per
    accept) with HIP events and counts its algorithmic flops (8 N K k per frequency, K = 24 j for
    the j-th accept of 24-component flips); the run itself is bitwise the one without events."""
    T = 7
    freqs = freq_list(3, 28e9, 32e9)
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, P.METHOD_LOWRANK)
    gpu.checkpoint()
    gpu.profile(True)
    a = gpu.dbs_run()
    st = gpu.kernel_stats(1)
    gpu.profile(False)
    gpu.restore()
    b = gpu.dbs_run()
    assert a["accepts"] > 0 and st["launches"] == a["accepts"] and st["total_ms"] > 0
    N, nf, J = gpu.N, len(freqs), a["accepts"]
    assert st["flops"] == sum(8.0 * N * (24 * j) * 24 * nf for j in range(J))
    assert np.array_equal(a["acc_tiles"], b["acc_tiles"]) and a["acc_F"].tolist() == b["acc_F"].tolist()
    gpu.close()


@pytest.mark.parametrize("factors", [P.FACTORS_AUTO, P.FACTORS_REBUILT])
def test_device_loop_folds_factors(factors):
    """Factor capacity of 3 accepts: the device-driven loop folds between its chunks and still
    follows the oracle's run (M stored, or rebuilt from A and Lam')."""
    T = 7
    freqs = freq_list(3, 28e9, 32e9)
    gpu, ors, tm = setup_pair(P.LAYOUT_DISJOINT, T, freqs, P.METHOD_LOWRANK, max_deferred=3, factors=factors)
    g = gpu.dbs_run(loop=P.LOOP_AUTO)
    o = O.dbs_run(ors, O.OBJ_S11, freqs, DX, DY, DZ)
    assert g["flushes"] > 0
    assert np.array_equal(g["acc_tiles"], o["acc_tiles"])
    assert np.array_equal(g["acc_eval"], o["acc_eval"])
    assert rel(g["acc_F"], o["acc_F"]) < RTOL
    gpu.close()
