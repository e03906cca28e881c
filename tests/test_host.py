"""Host-side helpers without a GPU: the resume record (save_run / load_run) and the host twin of
pngf_perturb_g (the deterministic non-reciprocal perturbation that models acmefdfd's
single-precision G)."""
import json

import numpy as np
import pytest

import pngf_amd as P


def fake_run(tiles, evals, F, flips, state, converged=False):
    return dict(acc_tiles=np.array(tiles, np.int32), acc_eval=np.array(evals, np.int64),
                acc_F=np.array(F, np.float64), flip_count=flips, converged=converged,
                driver_state=np.array(state, np.uint32))


def test_resume_record_round_trip(tmp_path):
    """Two runs of one DBS: the second run's evaluation indices continue the first's count, the
    driver position is the last run's, and the record reads back with numpy types."""
    a = fake_run([3, 17], [4, 40], [0.5, 0.4], 55, [1, 2, 3])
    b = fake_run([8], [2], [0.3], 20, [7, 8, 9, 10], converged=True)
    tm = np.array([0, 1, 1, 0], np.int32)
    path = tmp_path / "run.json"
    P.save_run(path, -514213645, tm, [a, b])
    rec = P.load_run(path)
    assert rec["seed"] == -514213645
    assert rec["acc_tiles"] == [3, 17, 8]
    assert rec["acc_eval"] == [4, 40, 57]
    assert rec["acc_F"] == [0.5, 0.4, 0.3]
    assert rec["flip_count"] == 75 and rec["converged"] is True
    assert rec["tile_map"].dtype == np.int32 and rec["tile_map"].tolist() == [0, 1, 1, 0]
    assert rec["driver_state"].dtype == np.uint32 and rec["driver_state"].tolist() == [7, 8, 9, 10]


def test_resume_record_format_is_checked(tmp_path):
    path = tmp_path / "other.json"
    path.write_text(json.dumps({"format": "something-else"}))
    with pytest.raises(ValueError):
        P.load_run(path)


def test_mix64_is_splitmix64():
    """The published splitmix64 generator (Steele, Lea and Flood; Vigna's reference code) from
    state 0: its first two outputs are the mixing function at 1 and 2 times the golden gamma."""
    g = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        assert int(P._mix64(np.uint64(0))) == 0xE220A8397B1DCDAF
        assert int(P._mix64(g)) == 0x6E789E6AA1B965F4


def symmetric_g(N, M, seed=0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))
    G = np.asfortranarray(A)
    G[:N, :N] = 0.5 * (A[:N, :N] + A[:N, :N].T)
    return G


def test_perturbation_bounds_locality_and_determinism():
    """Each real and imaginary part moves by at most rel * max(|Re| + |Im|) of the block, the
    entries outside the N x N block (the Gobj rows of the file) do not move, the same seed gives
    the same matrix and another seed another one, and the uniform draws fill [-1, 1)."""
    N, M, rel = 40, 42, 1e-3
    G = symmetric_g(N, M)
    amax = np.max(np.abs(G[:N, :N].real) + np.abs(G[:N, :N].imag))
    H = P.perturb_block(G, N, rel, seed=3)
    d = H[:N, :N] - G[:N, :N]
    assert np.max(np.abs(d.real)) <= rel * amax and np.max(np.abs(d.imag)) <= rel * amax
    assert np.array_equal(H[N:, :], G[N:, :]) and np.array_equal(H[:, N:], G[:, N:])
    assert np.array_equal(H, P.perturb_block(G, N, rel, seed=3))
    assert not np.array_equal(H, P.perturb_block(G, N, rel, seed=4))
    u = d.real.ravel() / (rel * amax)
    assert u.min() < -0.95 and u.max() > 0.95 and abs(u.mean()) < 0.05
    assert np.mean(np.abs(u)) == pytest.approx(0.5, abs=0.03)


def test_perturbation_breaks_reciprocity_at_its_level():
    """A complex-symmetric G stops being symmetric, by about rel relative to its largest entry
    (the level at which the symmetric accept is switched off for acmefdfd's widened FP32 G)."""
    N, rel = 60, 1e-7
    G = symmetric_g(N, N, seed=1)
    assert np.array_equal(G, G.T)
    H = P.perturb_block(G, N, rel, seed=1)
    amax = np.max(np.abs(G.real) + np.abs(G.imag))
    asym = np.max(np.abs(H - H.T)) / amax
    assert 0.3 * rel < asym <= 2 * np.sqrt(2) * rel
