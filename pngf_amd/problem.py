"""Problem definitions: initial patterns, frequency sets and the named workloads.

Initial pattern convention: Bernoulli(0.5) from ``std::mt19937(seed)``, one
``gen() & 1`` per tile in raster order (the reference's ``rand_binary``,
pngf-opt/magichat.hpp:78-81), then the port neighbourhood of main.cpp:213-219
(metal left and right of the port tile, port tile and the tiles above and below
it void).
"""
import numpy as np


class MT19937:
    """std::mt19937 (32-bit Mersenne Twister, init_genrand seeding)."""

    def __init__(self, seed):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.i = 624

    def __call__(self):
        if self.i >= 624:
            mt = self.mt
            for k in range(624):
                y = (mt[k] & 0x80000000) | (mt[(k + 1) % 624] & 0x7FFFFFFF)
                mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def force_port(tm, T):
    """main.cpp:210-219: metal either side of the port tile, port tile and its
    vertical neighbours void."""
    mid = (T - 1) // 2
    tm = np.array(tm, dtype=np.int32).copy()
    tm[mid * T + mid - 1] = 1
    tm[mid * T + mid + 1] = 1
    tm[mid * T + mid] = 0
    tm[(mid - 1) * T + mid] = 0
    tm[(mid + 1) * T + mid] = 0
    return tm


def initial_pattern(T, seed):
    g = MT19937(seed)
    tm = np.array([g() & 1 for _ in range(T * T)], dtype=np.int32)
    return force_port(tm, T)


def dipole_pattern(T):
    """INIT_EMPTY of the reference (main.cpp:155): only the two port neighbours."""
    return force_port(np.zeros(T * T, dtype=np.int32), T)


def freq_list(n, lo, hi):
    return list(np.linspace(lo, hi, n)) if n > 1 else [0.5 * (lo + hi)]


# Named workloads (sizes as used by bench.py and the tests).
WORKLOADS = {
    # one raster DBS pass at 7x7 tiles, one frequency (the CPU-runnable case)
    "antenna7x7_1f": dict(T=7, freqs=[30e9], layout=1, max_evals=49),
    # 21x21 tiles, 9 frequencies 26..34 GHz, DBS to convergence
    "antenna21x21_9f": dict(T=21, freqs=freq_list(9, 26e9, 34e9), layout=1, max_evals=0),
    # 31x31 tiles, 7 frequencies 27..33 GHz
    "antenna31x31_7f": dict(T=31, freqs=freq_list(7, 27e9, 33e9), layout=1, max_evals=0),
}

# 16 concurrent DBS instances: 4 seeds at each of 11, 15, 21 and 25 tiles, 5 frequencies.
# Factor capacity per size, in accepts (runs from these starts accept up to ~50, ~100, ~210 and
# ~230 flips): no instance folds.  W0 of all 16 instances takes 120 GB, the factors ~57 GB.
MULTI16 = dict(sizes=[11, 15, 21, 25], seeds=[0, 1, 2, 3], freqs=freq_list(5, 28e9, 32e9), layout=1,
               max_deferred={11: 64, 15: 128, 21: 256, 25: 256})
