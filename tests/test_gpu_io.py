"""The drop-in's file path and failure handling, and resuming a DBS run from disk.

* acmefdfd's raw G files (pngf-opt/pngfopt.hpp:121-184 reads what acmefdfd/main.cpp:533-538
  writes: M x M complex128, column-major, no header, M = N + 2 with the two NF2FF rows) through
  pngf_load_g give bitwise the state and scores pngf_set_g gives from the same buffer, and the
  reference driver's run on them is the oracle's.
* Every status the ABI promises for bad input: missing file, wrong size, NaN or Inf in G or in a
  Gobj row, a singular C, a device-memory cap, calls out of order and bad tiles.
* Resume: a run stopped mid-way, saved as tile map + accepted flips + driver position, then
  continued on a handle rebuilt from the tile map, reproduces the uninterrupted run and the
  oracle's (reference driver: the MagicHat generator position; raster driver: the place in
  the visit cycle).
"""
import numpy as np
import pytest

import oracle as O
import pngf_amd as P
from pngf_amd.problem import freq_list, initial_pattern

pytestmark = pytest.mark.gpu

RTOL = 1e-10
SOFT_WELL = 0.4
DX, DY, DZ = 0.01 / 60, 0.01 / 60, 0.01 / 180
REF_SEED = -514213645


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def write_g(path, G):
    """acmefdfd's layout: the column-major M x M complex128 array, raw (acmefdfd/main.cpp:536-538)."""
    np.asfortranarray(G, dtype=np.complex128).ravel(order="K").tofile(path)


def ref_optimizer(T, nf, method=P.METHOD_LOWRANK, **kw):
    return P.Optimizer(P.LAYOUT_SHARED, T, num_freq=nf, num_eobj=3, max_flips=4, method=method,
                       objective=P.OBJ_REF, driver=P.DRIVER_REF, **kw)


def finish_init(opt, tm, T):
    mid = (T - 1) // 2
    opt.set_tile_map(tm)
    opt.set_source_port(mid, mid, 1, 2)
    opt.init_state()


@pytest.mark.parametrize("layout,T,ne", [(P.LAYOUT_DISJOINT, 21, 1), (P.LAYOUT_SHARED, 11, 3)])
def test_device_synth_g_matches_host(layout, T, ne):
    """The synthetic G the device forms (pngf_synth_g, what the benchmark and the full-size tests
    run on) against the host synthesiser's (pngf_synth_g_host, what the oracle's recorded runs
    were computed from), at the default softening: the design block agrees to 4e-16 of its
    largest entry."""
    freqs = [26e9, 34e9]
    N = P.layout_size(layout, T)
    o = P.Optimizer(layout, T, num_freq=len(freqs), num_eobj=ne, method=P.METHOD_PLAIN)
    try:
        for f, hz in enumerate(freqs):
            o.set_frequency(f, hz)
            o.synth_g(f)
        o.set_tile_map(initial_pattern(T, 0))
        o.set_source_port((T - 1) // 2, (T - 1) // 2, 1, 2)
        o.init_state()
        for f, hz in enumerate(freqs):
            Gd = o.g(f)
            Gh = P.synth_g_host(layout, T, ne, hz)[:N, :N]
            d = rel(Gd, Gh)
            print("[synth] layout %d T %d f %.0f GHz: device vs host %.3g, bitwise %s" % (
                layout, T, hz / 1e9, d, np.array_equal(Gd, Gh)))
            assert d < 4e-16
    finally:
        o.close()


def test_load_g_files_match_set_g_and_oracle(tmp_path):
    """Shared layout, ne = 3 (two NF2FF rows), reference driver and objective: load_g from
    files equals set_g bitwise (state, every flip's score, the whole DBS run) and follows the
    oracle's accepted sequence."""
    T = 7
    freqs = freq_list(2, 29e9, 31e9)
    tm = initial_pattern(T, 1)
    Gs = [P.synth_g_host(P.LAYOUT_SHARED, T, 3, hz, soft=SOFT_WELL) for hz in freqs]
    paths = []
    for f, G in enumerate(Gs):
        p = tmp_path / ("Gmat_sub_%02d.bin" % (f + 1))
        write_g(p, G)
        assert p.stat().st_size == G.shape[0] ** 2 * 16
        paths.append(str(p))
    a = ref_optimizer(T, len(freqs))
    b = ref_optimizer(T, len(freqs))
    for f, hz in enumerate(freqs):
        a.set_frequency(f, hz)
        b.set_frequency(f, hz)
        a.load_g(f, paths[f])
        b.set_g(f, Gs[f])
    finish_init(a, tm, T)
    finish_init(b, tm, T)
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa["Eobj"], sb["Eobj"]) and np.array_equal(sa["LN"], sb["LN"])
    cands = [[t] for t in range(T * T) if t != (T - 1) // 2 * (T + 1)]
    Fa, ea = a.eval_batch(cands)
    Fb, eb = b.eval_batch(cands)
    assert np.array_equal(Fa, Fb) and np.array_equal(ea, eb)
    ors = []
    mid = (T - 1) // 2
    for G in Gs:
        o = O.Optimizer(P.LAYOUT_SHARED, T, G, ne=3, max_flips=4)
        assert o.init_state(tm, (mid, mid, 1, 2)) == 0
        ors.append(o)
    E = np.array([o.scalars()["Eobj"] for o in ors])
    assert rel(sa["Eobj"], E) < RTOL
    ga = a.dbs_run(seed=REF_SEED)
    gb = b.dbs_run(seed=REF_SEED)
    o = O.dbs_run(ors, O.OBJ_REF, freqs, DX, DY, DZ, seed=REF_SEED)
    assert ga["accepts"] == o["accepts"] and ga["accepts"] > 0
    assert np.array_equal(ga["acc_tiles"], gb["acc_tiles"]) and ga["acc_F"].tolist() == gb["acc_F"].tolist()
    assert np.array_equal(ga["acc_tiles"], o["acc_tiles"])
    assert np.array_equal(ga["acc_eval"], o["acc_eval"])
    assert rel(ga["acc_F"], o["acc_F"]) < RTOL
    a.close()
    b.close()


def test_load_g_file_errors(tmp_path):
    """Missing file -> IO; a file one complex short or one complex long -> SIZE (the reference
    reads whatever is there, pngfopt.hpp:163-164); a device-memory cap -> NOMEM; and the handle
    still loads a good file after each of them."""
    T = 5
    G = P.synth_g_host(P.LAYOUT_SHARED, T, 3, 30e9)
    good = tmp_path / "good.bin"
    write_g(good, G)
    raw = good.read_bytes()
    short = tmp_path / "short.bin"
    short.write_bytes(raw[:-16])
    long_ = tmp_path / "long.bin"
    long_.write_bytes(raw + raw[:16])
    opt = ref_optimizer(T, 1)
    with pytest.raises(P.PNGFError) as e:
        opt.load_g(0, str(tmp_path / "missing.bin"))
    assert e.value.code == -2
    for bad in (short, long_):
        with pytest.raises(P.PNGFError) as e:
            opt.load_g(0, str(bad))
        assert e.value.code == -3
    with pytest.raises(P.PNGFError) as e:
        opt.init_state()
    assert e.value.code == -7  # no tile map, no G yet
    opt.load_g(0, str(good))
    finish_init(opt, initial_pattern(T, 0), T)
    assert np.isfinite(opt.state()["F"])
    opt.close()
    capped = ref_optimizer(T, 1, mem_limit=1 << 20)  # G alone needs 16 N^2 bytes > 1 MiB
    with pytest.raises(P.PNGFError) as e:
        capped.load_g(0, str(good))
    assert e.value.code == -6
    with pytest.raises(P.PNGFError) as e:
        capped.set_g(0, G)
    assert e.value.code == -6
    capped.close()


@pytest.mark.parametrize("where", ["g_nan", "g_inf", "gobj_nan"])
@pytest.mark.parametrize("source", ["file", "buffer"])
def test_nonfinite_g_is_refused(tmp_path, where, source):
    """A NaN or an infinity in G (or in a Gobj row of the file) makes init_state return
    NONFINITE instead of a run of NaN scores that would end as 'converged'."""
    T = 5
    G = P.synth_g_host(P.LAYOUT_SHARED, T, 3, 30e9).copy(order="F")
    N = P.layout_size(P.LAYOUT_SHARED, T)
    if where == "g_nan":
        G[N - 3, 7] = np.nan
    elif where == "g_inf":
        G[4, N - 1] = complex(0.0, np.inf)
    else:
        G[N + 1, 11] = np.nan
    opt = ref_optimizer(T, 1)
    if source == "file":
        p = tmp_path / "g.bin"
        write_g(p, G)
        opt.load_g(0, str(p))
    else:
        opt.set_g(0, G)
    opt.set_tile_map(initial_pattern(T, 0))
    opt.set_source_port(2, 2, 1, 2)
    with pytest.raises(P.PNGFError) as e:
        opt.init_state()
    assert e.value.code == -8
    with pytest.raises(P.PNGFError) as e:
        opt.dbs_run()
    assert e.value.code == -7
    opt.close()


@pytest.mark.parametrize("method", [P.METHOD_LOWRANK, P.METHOD_PLAIN])
def test_singular_c_is_refused(method):
    """A metal component whose row of G is zero makes C = I + P(G - I) singular (that row of C
    is zero, and stays exactly zero through the elimination): init_state returns SINGULAR,
    where the reference ignores invertMatrix's status (main.cpp:246, linalg.hpp:55-58)."""
    T = 7
    layout = P.LAYOUT_DISJOINT
    G = P.synth_g_host(layout, T, 1, 30e9, soft=SOFT_WELL).copy(order="F")
    tm = initial_pattern(T, 0)
    mask = O.layout_mask(layout, T, tm)
    port = set(O.layout_port(layout, T, 3, 3).tolist())
    p = next(i for i in np.nonzero(mask)[0] if i not in port)
    G[p, :] = 0.0
    opt = P.Optimizer(layout, T, num_freq=1, method=method)
    opt.set_g(0, G)
    opt.set_tile_map(tm)
    opt.set_source_port(3, 3, 1, 2)
    with pytest.raises(P.PNGFError) as e:
        opt.init_state()
    assert e.value.code == -4
    opt.close()


def test_argument_and_order_errors():
    """Calls out of order -> STATE; bad tiles or flips -> ARG; nothing is left half-done."""
    T = 7
    G = P.synth_g_host(P.LAYOUT_DISJOINT, T, 1, 30e9, soft=SOFT_WELL)
    opt = P.Optimizer(P.LAYOUT_DISJOINT, T, num_freq=1)
    with pytest.raises(P.PNGFError) as e:
        opt.eval_batch([[3]])
    assert e.value.code == -7
    with pytest.raises(P.PNGFError) as e:
        opt.set_source_port(3, 3, 1, 3)  # the disjoint layout has 2 port rows
    assert e.value.code == -1
    opt.set_g(0, G)
    finish = initial_pattern(T, 0)
    opt.set_tile_map(finish)
    opt.set_source_port(3, 3, 1, 2)
    opt.init_state()
    with pytest.raises(P.PNGFError) as e:
        opt.set_g(0, G)
    assert e.value.code == -7
    for bad in ([T * T], [-1], [3, 3]):
        with pytest.raises(P.PNGFError) as e:
            opt.accept(bad)
        assert e.value.code == -1
        with pytest.raises(P.PNGFError) as e:
            opt.eval_batch([[5], bad])
        assert e.value.code == -1
    assert np.array_equal(opt.tile_map(), finish)
    # empty inputs: nothing scored, nothing accepted
    F, e = opt.eval_batch([])
    assert F.shape == (0,) and e.shape == (0, 1, 1)
    assert opt.sweep([], 1.0) == (-1, 1.0, 0)
    with pytest.raises(P.PNGFError) as e:
        opt.dbs_run(order=P.ORDER_BEST, resume_from=np.zeros(P.DRIVER_STATE_WORDS, np.uint32))
    assert e.value.code == -1
    with pytest.raises(P.PNGFError) as e:
        opt.dbs_run(resume_from=np.zeros(P.DRIVER_STATE_WORDS, np.uint32))  # no saved position
    assert e.value.code == -1
    opt.close()


def _resume_case(driver):
    T = 7
    if driver == "ref":
        freqs = freq_list(2, 29e9, 31e9)
        layout, ne, mf, obj, drv, omode = P.LAYOUT_SHARED, 3, 4, P.OBJ_REF, P.DRIVER_REF, O.OBJ_REF
        seed, tm = REF_SEED, initial_pattern(T, 1)
    else:
        freqs = freq_list(3, 28e9, 32e9)
        layout, ne, mf, obj, drv, omode = P.LAYOUT_DISJOINT, 1, 1, P.OBJ_S11, P.DRIVER_RASTER, O.OBJ_S11
        seed, tm = 0, initial_pattern(T, 0)
    Gs = [P.synth_g_host(layout, T, ne, hz, soft=SOFT_WELL) for hz in freqs]

    def make(tmap):
        o = P.Optimizer(layout, T, num_freq=len(freqs), num_eobj=ne, max_flips=mf, objective=obj, driver=drv)
        for f, (hz, G) in enumerate(zip(freqs, Gs)):
            o.set_frequency(f, hz)
            o.set_g(f, G)
        finish_init(o, tmap, T)
        return o

    ors = []
    mid = (T - 1) // 2
    for G in Gs:
        o = O.Optimizer(layout, T, G, ne=ne, max_flips=mf)
        assert o.init_state(tm, (mid, mid, 1, 2)) == 0
        ors.append(o)
    oracle = O.dbs_run(ors, omode, freqs, DX, DY, DZ, seed=seed)
    return make, tm, seed, oracle


@pytest.mark.parametrize("driver,loop", [("ref", P.LOOP_HOST), ("raster", P.LOOP_AUTO), ("raster", P.LOOP_HOST)])
def test_resume_reproduces_uninterrupted_run(tmp_path, driver, loop):
    """Stop after some evaluations (mid-way between accepts), save tile map, accepted flips and
    driver position to disk, close the handle; rebuild a handle from the saved tile map
    (C inverted again) and continue: the accepted flips (tile, evaluation index), the objective
    and the final map are those of the uninterrupted run and of the oracle."""
    make, tm, seed, oracle = _resume_case(driver)
    a = make(tm)
    full = a.dbs_run(seed=seed, loop=loop)
    assert np.array_equal(full["acc_tiles"], oracle["acc_tiles"])
    assert np.array_equal(full["acc_eval"], oracle["acc_eval"])
    assert full["accepts"] >= 6
    a.close()
    cut = int(full["acc_eval"][full["accepts"] // 2]) + 3
    b = make(tm)
    part1 = b.dbs_run(seed=seed, loop=loop, max_evals=cut)
    assert part1["flip_count"] == cut and not part1["converged"]
    path = tmp_path / "resume.json"
    P.save_run(str(path), seed, b.tile_map(), [part1])
    b.close()
    rec = P.load_run(str(path))
    c = make(rec["tile_map"])
    part2 = c.dbs_run(seed=seed, loop=loop, resume_from=rec["driver_state"])
    tiles = np.concatenate([rec["acc_tiles"], part2["acc_tiles"]]).astype(int)
    evals = np.concatenate([rec["acc_eval"], part2["acc_eval"] + rec["flip_count"]]).astype(int)
    accF = np.concatenate([rec["acc_F"], part2["acc_F"]])
    assert part2["converged"]
    assert np.array_equal(tiles, full["acc_tiles"]) and np.array_equal(tiles, oracle["acc_tiles"])
    assert np.array_equal(evals, full["acc_eval"]) and np.array_equal(evals, oracle["acc_eval"])
    assert rec["flip_count"] + part2["flip_count"] == full["flip_count"] == oracle["flip_count"]
    assert rel(accF, oracle["acc_F"]) < RTOL
    assert np.array_equal(c.tile_map(), oracle["final_map"])
    c.close()
