"""Pins the CPU oracle (oracle/pngf_oracle.cpp) before the GPU tests trust it.

Each check is one a plausible mistake in the restatement would fail: dense
brute-force re-solves, numpy/scipy LAPACK, closed-form layout counts, the C++
standard's mt19937 known answer, hand-computed objective values, the golden
fixture written by tests/golden/make_golden.py, and algebraic invariants
(a flip and its reverse restore the state; DBS ends in a local minimum).
"""
import os

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
DX, DY, DZ = 0.01 / 60, 0.01 / 60, 0.01 / 180


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


# ------------------------------------------------------------------ layout --
@pytest.mark.parametrize("T", [1, 2, 7, 13, 21, 31])
def test_shared_layout_count(T):
    # pngfopt.hpp:86-94: (3T)(3T+1) Ex + (3T+1)(3T) Ey
    assert O.layout_size(O.LAYOUT_SHARED, T) == 2 * (3 * T) * (3 * T + 1)


def test_shared_single_flip_k_values():
    """A single-tile flip changes 12 interior components plus 3 per side whose
    neighbour (or the outside) is void: k = 12 + 3 * (#void sides)."""
    rng = np.random.default_rng(3)
    T = 7
    for _ in range(20):
        tm = rng.integers(0, 2, T * T).astype(np.int32)
        P0 = O.layout_mask(O.LAYOUT_SHARED, T, tm)
        for t in range(T * T):
            t2 = tm.copy()
            t2[t] ^= 1
            k = int(np.sum(O.layout_mask(O.LAYOUT_SHARED, T, t2) != P0))
            ty, tx = divmod(t, T)
            void = 0
            for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                x, y = tx + dx, ty + dy
                if not (0 <= x < T and 0 <= y < T) or tm[y * T + x] == 0:
                    void += 1
            assert k == 12 + 3 * void
            assert k in (12, 15, 18, 21, 24)


def test_port_indices():
    # pngfopt.hpp:588 with port tile (10,10), offset 1, height 2, 3 cells per tile
    got = O.layout_port(O.LAYOUT_SHARED, 21, 10, 10, 1, 2)
    want = [(30 + ii) + (30 + jj + 1) * 63 for jj in range(2) for ii in range(3)]
    assert got.tolist() == want
    assert O.layout_port(O.LAYOUT_DISJOINT, 21, 10, 10).tolist() == list(range(10584, 10590))


# ------------------------------------------------------- dense LAPACK-alike --
@pytest.mark.parametrize("n", [1, 5, 24, 60])
def test_zgetrf_matches_scipy(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    lu, ipiv, info = O.zgetrf(a)
    lu_s, piv_s, info_s = lapack.zgetrf(a)
    assert info == info_s == 0
    assert ipiv.tolist() == piv_s.tolist()   # same pivot rule (izamax: first max |Re|+|Im|)
    assert rel(lu, lu_s) < 1e-12
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = O.zgetrs(lu, ipiv, b)
    assert rel(a @ x, b) < 1e-11
    inv = O.zgetri(lu, ipiv)
    assert rel(inv, np.linalg.inv(a)) < 1e-11


def test_zgetrf_singular_info():
    a = np.ones((4, 4), dtype=complex)
    _, _, info = O.zgetrf(a)
    _, _, info_s = lapack.zgetrf(a)
    assert info == info_s and info > 0


def test_invert_dense():
    rng = np.random.default_rng(7)
    a = rng.standard_normal((150, 150)) + 1j * rng.standard_normal((150, 150)) + 10 * np.eye(150)
    inv, info = O.invert(a)
    assert info == 0
    assert rel(inv, np.linalg.inv(a)) < 1e-12


# --------------------------------------------------------- Woodbury path -----
def dense_eobj(kind, T, G, extra, tm, port):
    N = G.shape[0]
    P = O.layout_mask(kind, T, tm)
    Gobj = np.vstack([G[port, :].sum(0)] + list(extra))
    Einc = G[:, port].sum(1)
    C = np.eye(N, dtype=complex)
    C[P == 1, :] = G[P == 1, :]
    return Gobj @ np.linalg.solve(C, -P * Einc)


def random_problem(kind, T, ne, seed):
    rng = np.random.default_rng(seed)
    N = O.layout_size(kind, T)
    G = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / np.sqrt(N)
    G += np.diag(2 + rng.random(N))
    extra = rng.standard_normal((ne - 1, N)) + 1j * rng.standard_normal((ne - 1, N))
    M = N + ne - 1
    fb = np.zeros((M, M), dtype=complex, order="F")
    fb[:N, :N] = G
    fb[N:, :N] = extra
    tm = rng.integers(0, 2, T * T).astype(np.int32)
    return G, extra, fb, tm


@pytest.mark.parametrize("kind", [O.LAYOUT_SHARED, O.LAYOUT_DISJOINT])
@pytest.mark.parametrize("ne", [1, 3])
def test_eval_matches_dense_resolve(kind, ne):
    T = 5
    G, extra, fb, tm = random_problem(kind, T, ne, 11 + ne + 7 * kind)
    port = O.layout_port(kind, T, 2, 2)
    o = O.Optimizer(kind, T, fb, ne=ne, max_flips=4)
    assert o.init_state(tm, (2, 2, 1, 2)) == 0
    assert rel(o.scalars()["Eobj"], dense_eobj(kind, T, G, extra, tm, port)) < 1e-11
    rng = np.random.default_rng(5)
    for _ in range(12):
        nt = int(rng.integers(1, 5))
        tiles = sorted(rng.choice(T * T, nt, replace=False).tolist())
        e = o.compute_updated_Eobj(tiles)
        t2 = tm.copy()
        t2[tiles] ^= 1
        assert rel(e, dense_eobj(kind, T, G, extra, t2, port)) < 1e-10


@pytest.mark.parametrize("kind", [O.LAYOUT_SHARED, O.LAYOUT_DISJOINT])
def test_accept_equals_reinversion_and_flip_twice_restores(kind):
    T = 4
    G, extra, fb, tm = random_problem(kind, T, 1, 23 + kind)
    N = G.shape[0]
    o = O.Optimizer(kind, T, fb, ne=1)
    assert o.init_state(tm, (2, 2, 1, 2)) == 0
    e_start = o.scalars()["Eobj"].copy()
    c_start = o.cinv().copy()
    cur = tm.copy()
    for t in [5, 0, 15, 5, 9]:
        o.compute_updated_Eobj([t])
        o.accept()
        cur[t] ^= 1
        P = O.layout_mask(kind, T, cur)
        C = np.eye(N, dtype=complex)
        C[P == 1, :] = G[P == 1, :]
        assert rel(o.cinv(), np.linalg.inv(C)) < 1e-10
        assert np.array_equal(o.tile_map(), cur)
    # undo everything: flips applied again in reverse order restore the state
    for t in [9, 5, 15, 0, 5]:
        o.compute_updated_Eobj([t])
        o.accept()
    assert np.array_equal(o.tile_map(), tm)
    assert rel(o.scalars()["Eobj"], e_start) < 1e-10
    assert rel(o.cinv(), c_start) < 1e-10


# ------------------------------------------------- complex64 build (FP32) --
def test_c64_build_is_single_precision():
    assert O.lib(c64=True).or_is_c64() == 1 and O.lib().or_is_c64() == 0


@pytest.mark.parametrize("kind", [O.LAYOUT_SHARED, O.LAYOUT_DISJOINT])
@pytest.mark.parametrize("ne", [1, 3])
def test_c64_eval_tracks_dense_resolve(kind, ne):
    """The complex64 build (the reference's USE_SINGLE_PRECISION) scores flips to FP32 accuracy of
    the dense re-solve (a wrong term would be an O(1) error), and it does compute in FP32: its
    scores are not the complex128 build's."""
    T = 5
    G, extra, fb, tm = random_problem(kind, T, ne, 11 + ne + 7 * kind)
    port = O.layout_port(kind, T, 2, 2)
    o32 = O.Optimizer(kind, T, fb, ne=ne, max_flips=4, c64=True)
    o64 = O.Optimizer(kind, T, fb, ne=ne, max_flips=4)
    assert o32.init_state(tm, (2, 2, 1, 2)) == 0 and o64.init_state(tm, (2, 2, 1, 2)) == 0
    assert rel(o32.scalars()["Eobj"], dense_eobj(kind, T, G, extra, tm, port)) < 1e-5
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(12):
        nt = int(rng.integers(1, 5))
        tiles = sorted(rng.choice(T * T, nt, replace=False).tolist())
        e32 = o32.compute_updated_Eobj(tiles)
        e64 = o64.compute_updated_Eobj(tiles)
        t2 = tm.copy()
        t2[tiles] ^= 1
        ref = dense_eobj(kind, T, G, extra, t2, port)
        assert rel(e32, ref) < 1e-5
        worst = max(worst, rel(e32, e64))
    assert worst > 1e-9


def test_c64_accept_equals_reinversion():
    """Woodbury accepts in complex64 keep C^-1 equal to the inverse of the updated C to FP32
    accuracy (numpy complex128 inverse as the truth)."""
    kind, T = O.LAYOUT_DISJOINT, 4
    G, extra, fb, tm = random_problem(kind, T, 1, 29)
    N = G.shape[0]
    o = O.Optimizer(kind, T, fb, ne=1, c64=True)
    assert o.init_state(tm, (2, 2, 1, 2)) == 0
    cur = tm.copy()
    for t in [5, 0, 15, 5, 9]:
        o.compute_updated_Eobj([t])
        o.accept()
        cur[t] ^= 1
        P = O.layout_mask(kind, T, cur)
        C = np.eye(N, dtype=complex)
        C[P == 1, :] = G[P == 1, :]
        assert 1e-9 < rel(o.cinv(), np.linalg.inv(C)) < 1e-5


def test_golden_fixture():
    z = np.load(os.path.join(HERE, "golden", "woodbury_small.npz"))
    names = sorted({k.split("__")[0] for k in z.files})
    assert len(names) == 4
    for name in names:
        g = {k.split("__")[1]: z[k] for k in z.files if k.startswith(name + "__")}
        kind, T, ne = int(g["kind"]), int(g["T"]), int(g["ne"])
        o = O.Optimizer(kind, T, g["file"], ne=ne, max_flips=4)
        mid = T // 2
        assert o.init_state(g["tile_map"], (mid, mid, 1, 2)) == 0
        assert rel(o.scalars()["Eobj"], g["eobj0"]) < 1e-10, name
        flips = [[t for t in row if t >= 0] for row in g["flips"]]
        for f, want in zip(flips, g["eobj_flip"]):
            assert rel(o.compute_updated_Eobj(f), want) < 1e-10, (name, f)
        for f, want in zip(flips, g["eobj_seq"]):
            o.compute_updated_Eobj(f)
            o.accept()
            assert rel(o.scalars()["Eobj"], want) < 1e-10, (name, f)


@pytest.mark.parametrize("name,ntiles,raster", [("antenna21x21_9f", 21 * 21, True),
                                                 ("antenna21x21_9f_nonreciprocal", 21 * 21, True),
                                                 ("multi16_T15_s0", 15 * 15, True),
                                                 ("multi16_T21_s0", 21 * 21, True),
                                                 ("multi16_T25_s0", 25 * 25, True),
                                                 ("reference_default", 21 * 21, False)])
def test_whole_run_fixtures(name, ntiles, raster):
    """The oracle's whole runs on the full-size workloads (tests/golden/make_oracle_sequence.py)
    have the shape first-improvement DBS gives (main.cpp:338-459): accepts at strictly increasing
    evaluations, each strictly improving the objective (main.cpp:429).  The raster run ends after
    one full pass of the tiles without an improving flip; the MagicHat run after its last accept
    draws every remaining quadrant entry once."""
    z = np.load(os.path.join(HERE, "golden", "oracle_%s.npz" % name))
    t, e, F = z["acc_tiles"], z["acc_eval"], z["acc_F"]
    assert len(t) == len(e) == len(F) > 5
    assert t.min() >= 0 and t.max() < ntiles
    assert np.all(np.diff(e) > 0) and e[0] >= 0
    assert F[0] < float(z["F_start"]) and np.all(np.diff(F) < 0)
    assert F[-1] == float(z["F_final"])
    if raster:
        assert int(z["flip_count"]) == e[-1] + 1 + ntiles
    else:
        # main.cpp:181: 120 quadrant indices; the last accepted one is skipped as last_tried and
        # the port tile's group as well (main.cpp:354-361), so at most 119 draws follow
        assert 0 < int(z["flip_count"]) - e[-1] - 1 <= 119


# ---------------------------------------------------------------- objective --
def test_objective_known_answers():
    # Zin = -(E0 + Vinc) dx/2 / (2 dy dz) = -4500 (E0 + Vinc) for the reference cell
    zfac = DX * 0.5 / (2 * DY * DZ)
    assert abs(zfac - 4500.0) < 1e-9
    e0, vinc = 0.001 + 0.002j, -0.012 - 0.004j
    zin = -(e0 + vinc) * 4500.0          # 49.5 + 9 i
    s11 = (zin - 50) / (zin + 50)
    F, s, _ = O.objective(O.OBJ_S11, np.array([[e0]]), np.array([vinc]), None, [30e9], DX, DY, DZ)
    assert abs(s[0] - s11) < 1e-15 and abs(F - abs(s11) ** 2) < 1e-15
    # matched port: S11 = 0
    F0, s0, _ = O.objective(O.OBJ_S11, np.array([[0j]]), np.array([-50 / 4500 + 0j]), None, [30e9], DX, DY, DZ)
    assert abs(s0[0]) < 1e-14 and F0 < 1e-28
    # mean over frequencies
    F2, _, _ = O.objective(O.OBJ_S11, np.array([[e0], [0j]]), np.array([vinc, -50 / 4500]), None,
                           [29e9, 31e9], DX, DY, DZ)
    assert abs(F2 - 0.5 * abs(s11) ** 2) < 1e-15
    # REF objective: 200 |S11|^2 + (15 - D)^2 with D of main.cpp:281-284
    ln = np.array([[1e-8 + 2e-8j, -3e-8j]])
    e = np.array([[e0, 2e-8, 1e-8j]])
    k0 = 2 * np.pi * 30e9 / 2.997924579998211e8
    a, b = e[0, 1] + ln[0, 0], e[0, 2] + ln[0, 1]
    D = 4 * np.pi * (k0 ** 2 / (32 * np.pi ** 2 * 376.7303134118051)) * (abs(a) ** 2 + abs(b) ** 2)
    D /= np.real(-(e0 + vinc) * DX * 0.5 * (2 * DY * DZ) * 0.5)
    Fr, _, d = O.objective(O.OBJ_REF, e, np.array([vinc]), ln, [30e9], DX, DY, DZ)
    assert abs(d[0] - D) < 1e-12 * abs(D)
    assert abs(Fr - (200 * abs(s11) ** 2 + (15 - D) ** 2)) < 1e-10 * Fr


# ----------------------------------------------------------------- MagicHat --
def test_mt19937_known_answer():
    assert O.mt19937_nth(5489, 10000) == 4123659995


def test_hat_draws_without_replacement():
    h = O.Hat(-514213645, 120)
    seen = [h.draw() for _ in range(120)]
    assert sorted(seen) == list(range(120))
    assert h.draw() == -1
    h.reset()
    again = [h.draw() for _ in range(120)]
    assert sorted(again) == list(range(120)) and again != seen
    h2 = O.Hat(-514213645, 120)
    assert [h2.draw() for _ in range(120)] == seen


# --------------------------------------------------------------------- DBS --
def test_dbs_ends_in_local_minimum():
    kind, T = O.LAYOUT_DISJOINT, 3
    G, extra, fb, tm = random_problem(kind, T, 1, 99)
    # scale so the port impedance is in a sensible range
    fb *= 1e-3
    G *= 1e-3
    port = O.layout_port(kind, T, 1, 1)
    o = O.Optimizer(kind, T, fb, ne=1)
    assert o.init_state(tm, (1, 1, 1, 2)) == 0
    r = O.dbs_run([o], O.OBJ_S11, [30e9], DX, DY, DZ)
    assert r["accepts"] > 0 and np.all(np.diff(r["acc_F"]) < 0)
    vinc = G[np.ix_(port, port)].sum()

    def F_of(tmap):
        e = dense_eobj(kind, T, G, extra, tmap, port)
        return O.objective(O.OBJ_S11, e[None, :], np.array([vinc]), None, [30e9], DX, DY, DZ)[0]

    fm = r["final_map"]
    assert abs(F_of(fm) - r["F_final"]) < 1e-10 * r["F_final"]
    for t in range(T * T):
        t2 = fm.copy()
        t2[t] ^= 1
        assert F_of(t2) >= r["F_final"] * (1 - 1e-12)


def test_empty_pattern_closed_form():
    """P = 0 (every tile and the port neighbourhood void): C = I, PEinc = 0, so S = 0,
    Eobj = 0 and Zin = -Vinc dx / 2 / (2 dy dz) with Vinc = sum of G over the port block
    (pngfopt.hpp:584-617, main.cpp:277-279)."""
    kind, T = O.LAYOUT_DISJOINT, 5
    G, extra, fb, tm = random_problem(kind, T, 1, 41)
    o = O.Optimizer(kind, T, fb, ne=1)
    assert o.init_state(np.zeros(T * T, dtype=np.int32), (2, 2, 1, 2)) == 0
    port = O.layout_port(kind, T, 2, 2)
    vinc = G[np.ix_(port, port)].sum()
    sc = o.scalars()
    assert np.all(sc["Eobj"] == 0)
    assert abs(sc["Vinc"] - vinc) < 1e-12 * abs(vinc)
    assert np.all(o.vec("S") == 0) and np.all(o.vec("PEinc") == 0)
    zin = -vinc * DX * 0.5 / (2 * DY * DZ)
    F, s, _ = O.objective(O.OBJ_S11, sc["Eobj"][None, :], np.array([sc["Vinc"]]), None, [30e9], DX, DY, DZ)
    assert abs(s[0] - (zin - 50) / (zin + 50)) < 1e-12


@pytest.mark.parametrize("mode", [O.OBJ_S11, O.OBJ_REF])
def test_history_records(mode):
    """The Logger history of the oracle's DBS loop (main.cpp:304-328, 401-426): a start record,
    then one record per evaluation whose F column is the evaluation's objective, whose F_best
    column is the best objective before it, whose decision is F < F_best, whose tile is the
    accepted tile where the loop accepted, and whose per-frequency S11 / D reproduce F through
    the objective (mean |S11|^2, or sum 200 |S11|^2 + (15 - D)^2)."""
    kind, T, ne = (O.LAYOUT_DISJOINT, 3, 1) if mode == O.OBJ_S11 else (O.LAYOUT_SHARED, 5, 3)
    freqs = [29e9, 31e9]
    opts = []
    for f in range(len(freqs)):
        G, extra, fb, tm = random_problem(kind, T, ne, 40 + f)
        fb *= 1e-3
        o = O.Optimizer(kind, T, fb, ne=ne, max_flips=4)   # the reference driver flips up to 4 tiles
        mid = (T - 1) // 2
        tm0 = random_problem(kind, T, ne, 40)[3]
        assert o.init_state(tm0, (mid, mid, 1, 2)) == 0
        opts.append(o)
    r = O.dbs_run(opts, mode, freqs, DX, DY, DZ, seed=-514213645, max_evals=40, trace_cap=64, hist_cap=64)
    n = r["flip_count"]
    assert n > 0 and len(r["hist_tile"]) == n + 1
    perf, dec, tile = r["hist_perf"], r["hist_dec"], r["hist_tile"]
    assert perf.shape[1] == 2 + 3 * len(freqs)
    assert tile[0] == -1 and dec[0] == -1 and perf[0, 0] == r["F_start"] and perf[0, 1] == r["F_start"]
    assert np.array_equal(perf[1:, 0], r["F_trace"])
    best = r["F_start"]
    for i in range(1, n + 1):
        assert perf[i, 1] == best
        assert dec[i] == (1 if perf[i, 0] < best else 0)
        best = min(best, perf[i, 0])
    assert np.array_equal(tile[r["acc_eval"] + 1], r["acc_tiles"])
    assert np.array_equal(np.nonzero(dec[1:] == 1)[0], r["acc_eval"])
    s = perf[:, 2::3] + 1j * perf[:, 3::3]
    D = perf[:, 4::3]
    if mode == O.OBJ_S11:
        assert np.all(D == 0.0)
        F = np.mean(np.abs(s) ** 2, axis=1)
    else:
        F = np.sum(200 * np.abs(s) ** 2 + (15 - D) ** 2, axis=1)
    assert np.max(np.abs(F - perf[:, 0]) / np.abs(perf[:, 0])) < 1e-12
