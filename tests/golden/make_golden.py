"""Writes tests/golden/woodbury_small.npz: golden vectors for the oracle.

Independent of the oracle and of the HIP library: everything here is a dense
numpy re-solve of the PNGF system on tiny problems.

  C = I - P + P G               (pngf-opt/pngfopt.hpp:196-233)
  Gobj[0] = sum_port G[p, :],  Einc = sum_port G[:, q]   (pngfopt.hpp:584-595)
  Eobj = Gobj C^-1 (-P o Einc)   (pngfopt.hpp:237-250, 542-552)

For each case: a seeded random G (file layout, with NF2FF rows when ne = 3), a
tile map, a list of flips and, for every flip, the Eobj of the flipped design
(what compute_updated_Eobj must return), plus the Eobj after committing the
flips in order (what update_C_inv + compute_R_matrix(0) must leave).

The layout / port index rules (pngfopt.hpp:86-94, 584-590, 624-664) are restated
here independently so that the fixture does not depend on either implementation.
Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def shared_layout(T, c=3):
    yxx, yxy, yyx, yyy = T * c, T * c + 1, T * c + 1, T * c
    yxs = yxx * yxy
    N = yxs + yyx * yyy

    def comps(t):
        ty, tx = divmod(t, T)
        bx = tx * c + ty * c * yxx
        by = yxs + tx * c + ty * c * yyx
        out = [bx + i + j * yxx for j in range(c + 1) for i in range(c)]
        out += [by + i + j * yyx for j in range(c) for i in range(c + 1)]
        return out

    mid = T // 2
    port = [(mid * c + ii) + (mid * c + jj + 1) * yxx for jj in range(2) for ii in range(c)]
    return N, comps, port


def disjoint_layout(T, c=3):
    per = c * (c + 1) * 2
    N = per * T * T + 2 * c

    def comps(t):
        return list(range(t * per, (t + 1) * per))

    port = list(range(per * T * T, N))
    return N, comps, port


def mask(N, comps, tm):
    P = np.zeros(N, dtype=np.int64)
    for t, v in enumerate(tm):
        if v:
            P[comps(t)] = 1
    return P


def eobj_dense(G, gobj_extra, P, port):
    N = G.shape[0]
    Gobj = np.vstack([G[port, :].sum(0)] + list(gobj_extra))
    Einc = G[:, port].sum(1)
    C = np.eye(N, dtype=complex)
    C[P == 1, :] = G[P == 1, :]
    PE = -P * Einc
    return Gobj @ np.linalg.solve(C, PE)


def make_case(rng, kind, T, ne, flips):
    N, comps, port = (shared_layout if kind == 0 else disjoint_layout)(T)
    M = N + ne - 1
    # diagonally weighted random G: C stays comfortably invertible for any mask
    G = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / np.sqrt(N)
    G += np.diag(2.0 + rng.random(N) + 1j * rng.random(N))
    extra = rng.standard_normal((ne - 1, N)) + 1j * rng.standard_normal((ne - 1, N))
    file_buf = np.zeros((M, M), dtype=complex)
    file_buf[:N, :N] = G
    file_buf[N:, :N] = extra
    tm = rng.integers(0, 2, T * T)
    P0 = mask(N, comps, tm)
    e0 = eobj_dense(G, extra, P0, port)
    per_flip = []
    for fl in flips:
        t2 = tm.copy()
        t2[fl] ^= 1
        per_flip.append(eobj_dense(G, extra, mask(N, comps, t2), port))
    seq = []
    t2 = tm.copy()
    for fl in flips:
        t2[fl] ^= 1
        seq.append(eobj_dense(G, extra, mask(N, comps, t2), port))
    nfl = max(len(f) for f in flips)
    fl_arr = np.full((len(flips), nfl), -1, dtype=np.int32)
    for i, f in enumerate(flips):
        fl_arr[i, :len(f)] = f
    return dict(file=file_buf, tile_map=tm.astype(np.int32), flips=fl_arr, eobj0=e0,
                eobj_flip=np.array(per_flip), eobj_seq=np.array(seq), N=N, ne=ne, kind=kind, T=T)


def main():
    rng = np.random.default_rng(20251203)
    cases = {
        "shared_T2_ne1": make_case(rng, 0, 2, 1, [[0], [1], [2], [3], [0, 3], [1, 2]]),
        "shared_T3_ne3": make_case(rng, 0, 3, 3, [[0], [4], [8], [0, 2, 6, 8], [3, 5], [7]]),
        "disjoint_T2_ne1": make_case(rng, 1, 2, 1, [[0], [1], [2], [3], [1, 2]]),
        "disjoint_T2_ne3": make_case(rng, 1, 2, 3, [[3], [0, 1, 2, 3], [2]]),
    }
    out = {}
    for name, c in cases.items():
        for k, v in c.items():
            out[name + "__" + k] = np.asarray(v)
    np.savez_compressed(os.path.join(HERE, "woodbury_small.npz"), **out)
    print("wrote", os.path.join(HERE, "woodbury_small.npz"))


if __name__ == "__main__":
    main()
