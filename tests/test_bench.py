"""bench.py's host-side contract, without a GPU: the one-GPU refusal and the roofline record it
builds from the library's event-timed kernel statistics."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stdout + r.stderr
    assert r.stdout.strip() == ""  # no JSON line


class FakeHandle:
    """kernel_stats as the library reports them: launches, summed event ms, algorithmic bytes
    and flops per category (0 eval, 1 colblk, 2 rowblk, 3 acc_small, 4 vcinv, 5 rankk)."""

    def __init__(self, N, stats, nf=9):
        self.N = N
        self.nf = nf
        self.stats = stats

    def kernel_stats(self, k):
        return self.stats.get(k, dict(launches=0, total_ms=0.0, bytes=0.0, flops=0.0))


def test_roofline_of_the_column_strip():
    # 183 launches of the 21x21x9 strip: 8 N K k flop and 16 (N K + 2 N k + K k) B per
    # frequency at the average factor depth K = 24 * 91 = 2184, k = 24
    N, K, k, nf, n = 10590, 2184, 24, 9, 183
    flops = nf * 8.0 * N * K * k
    byts = nf * 16.0 * (N * K + 2 * N * k + K * k)
    h = FakeHandle(N, {0: dict(launches=220, total_ms=8.3, bytes=1e9, flops=1e9),
                       1: dict(launches=n, total_ms=n * 0.72, bytes=n * byts, flops=n * flops)})
    rec, allk = bench.roofline(h, variant=4)
    assert rec["kernel"] == "colblk" and rec["launches"] == n
    # arithmetic intensity 11.7 flop/B is above the 9.8 ridge of 78.6 TFLOP/s over 8 TB/s
    assert rec["bound"] == "mfma" and rec["unit"] == "TFLOP/s" and rec["peak"] == 78.6
    assert rec["achieved"] == pytest.approx(flops / 0.72e-3 / 1e12)
    assert rec["frac"] == pytest.approx(rec["achieved"] / 78.6)
    # the counted strip launches are these: their HBM bytes per launch (2.8 % above the
    # algorithmic bytes) and MFMA-busy fraction
    assert byts == bench.STRIP_COUNTED["alg_bytes_per_launch"]
    assert rec["traffic"] == bench.STRIP_COUNTED["bytes_per_launch"]
    assert 1.0 < rec["traffic"] / byts < 1.05
    assert rec["mfma_busy"] == bench.STRIP_COUNTED["mfma_util"]
    assert set(allk) == {"eval", "colblk", "rowblk", "acc_small", "vcinv", "rankk"}


@pytest.mark.parametrize("change", ["variant", "launches", "depth", "freqs"])
def test_roofline_traffic_only_for_the_counted_launches(change):
    """Another tile variant, launch count, factor depth or frequency count than the counted
    launches: no traffic figure."""
    N, K, k, nf, n = 10590, 2184, 24, 9, 183
    if change == "launches":
        n = 150
    if change == "depth":
        K = 1800
    byts = nf * 16.0 * (N * K + 2 * N * k + K * k)
    flops = nf * 8.0 * N * K * k
    h = FakeHandle(N, {1: dict(launches=n, total_ms=n * 0.72, bytes=n * byts, flops=n * flops)},
                   nf=5 if change == "freqs" else 9)
    rec, _ = bench.roofline(h, variant=2 if change == "variant" else 4)
    assert rec["kernel"] == "colblk"
    assert rec["traffic"] is None and "mfma_busy" not in rec


def test_roofline_of_an_hbm_bound_kernel():
    # the PLAIN score streams 16 N (2k + 1 + ne) B for 8 N (k^2 + 2k + ne) flop: 6.3 flop/B
    N, k = 10590, 24
    byts = 16.0 * N * (2 * k + 2)
    flops = 8.0 * N * (k * k + 2 * k + 1)
    h = FakeHandle(N, {0: dict(launches=10, total_ms=10 * 0.002, bytes=10 * byts, flops=10 * flops)})
    rec, _ = bench.roofline(h)
    assert rec["kernel"] == "eval" and rec["bound"] == "hbm" and rec["unit"] == "GB/s"
    assert rec["achieved"] == pytest.approx(byts / 2e-6 / 1e9)
    assert rec["traffic"] is None


def test_no_kernel_timed():
    rec, allk = bench.roofline(FakeHandle(100, {}))
    assert rec is None and allk == {}


def test_oracle_run_check():
    """The bench line's comparison with the oracle's recorded whole run: the record itself passes
    with no objective difference; a run with one accepted flip moved, or a different evaluation
    count, does not."""
    import numpy as np
    z = np.load(bench.ORACLE_RUN)
    r = dict(flip_count=int(z["flip_count"]), acc_tiles=z["acc_tiles"].copy(), acc_eval=z["acc_eval"].copy(),
             F_final=float(z["F_final"]))
    c = bench.oracle_run_check(r)
    assert c["same_accepts"] and c["first_difference"] is None and c["F_rel"] == 0.0
    assert c["accepts"] == len(z["acc_tiles"])
    r2 = dict(r, acc_eval=r["acc_eval"] + np.eye(1, len(r["acc_eval"]), 5, dtype=r["acc_eval"].dtype)[0])
    c2 = bench.oracle_run_check(r2)
    assert not c2["same_accepts"] and c2["first_difference"] == 5
    r3 = dict(r, flip_count=r["flip_count"] + 1)
    c3 = bench.oracle_run_check(r3)
    assert not c3["same_accepts"] and c3["first_difference"] == len(z["acc_tiles"])
