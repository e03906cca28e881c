"""pngf_amd: MI355X (gfx950) implementation of the pngf-opt DBS Woodbury hot path.

Python view of the C ABI in ``include/pngf_gpu.h`` (``libpngf_gpu.so``, built by
``pngf_amd.build``).  ``Optimizer`` mirrors the method names of the reference
``PNGF_Optimizer`` (pngf-opt/pngfopt.hpp:18-679), with one handle covering every
frequency.  There is no CPU fallback: every compute call goes to the HIP library,
and loading fails loudly when the library is missing.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpngf_gpu.so")

LAYOUT_SHARED, LAYOUT_DISJOINT = 0, 1
OBJ_REF, OBJ_S11 = 0, 1
METHOD_PLAIN, METHOD_LOWRANK = 0, 1
DRIVER_REF, DRIVER_RASTER = 0, 1
ORDER_FIRST, ORDER_BEST = 0, 1
SYM_AUTO, SYM_OFF, SYM_PROJECT = 0, 1, 2
FACTORS_AUTO, FACTORS_STORED, FACTORS_REBUILT = 0, 1, 2
LOOP_AUTO, LOOP_HOST = 0, 1
PREC_C128, PREC_C64 = 0, 1
K_EVAL, K_COLBLK, K_ROWBLK, K_ACC_SMALL, K_VCINV, K_RANKK = range(6)
K_NAMES = ["eval", "colblk", "rowblk", "acc_small", "vcinv", "rankk"]

ERRORS = {0: "OK", -1: "ARG", -2: "IO", -3: "SIZE", -4: "SINGULAR", -5: "HIP", -6: "NOMEM",
          -7: "STATE", -8: "NONFINITE"}

# Every function declared in include/pngf_gpu.h.
EXPORTS = [
    "pngf_create", "pngf_destroy", "pngf_last_error", "pngf_layout_size", "pngf_set_frequency",
    "pngf_load_g", "pngf_set_g", "pngf_synth_g", "pngf_synth_g_host", "pngf_set_tile_map",
    "pngf_set_source_port", "pngf_init_state", "pngf_eval", "pngf_eval_batch", "pngf_accept",
    "pngf_get_state", "pngf_get_tile_map", "pngf_get_cinv", "pngf_get_w", "pngf_dbs_run",
    "pngf_profile", "pngf_kernel_stats", "pngf_get_g", "pngf_checkpoint", "pngf_restore",
    "pngf_mlp_create", "pngf_mlp_destroy", "pngf_mlp_last_error", "pngf_mlp_num_params",
    "pngf_mlp_get_params", "pngf_mlp_set_params", "pngf_mlp_predict", "pngf_mlp_grad", "pngf_mlp_train",
    "pngf_ranker_dims", "pngf_features", "pngf_dbs_run_ranked", "pngf_dbs_run_many", "pngf_dbs_run_pool",
    "pngf_autotune_result", "pngf_symmetry", "pngf_factor_capacity", "pngf_sweep", "pngf_mlp_predict_dev",
    "pngf_perturb_g",
]


class PNGFError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pngf error %d (%s): %s" % (code, ERRORS.get(code, "?"), msg))
        self.code = code


class Config(ct.Structure):
    _fields_ = [("layout", ct.c_int32), ("tiles_x", ct.c_int32), ("tiles_y", ct.c_int32),
                ("cells_x", ct.c_int32), ("cells_y", ct.c_int32), ("num_eobj", ct.c_int32),
                ("max_flips", ct.c_int32), ("num_freq", ct.c_int32), ("method", ct.c_int32),
                ("objective", ct.c_int32), ("driver", ct.c_int32), ("max_deferred", ct.c_int32),
                ("device", ct.c_int32), ("tile_variant", ct.c_int32), ("dx", ct.c_double),
                ("dy", ct.c_double), ("dz", ct.c_double), ("z0", ct.c_double),
                ("mem_limit", ct.c_int64), ("synth_soft", ct.c_double), ("symmetry", ct.c_int32),
                ("factors", ct.c_int32), ("precision", ct.c_int32)]


class DbsParams(ct.Structure):
    _fields_ = [("max_evals", ct.c_int64), ("seed", ct.c_int32), ("max_accepts", ct.c_int32),
                ("acc_tiles", ct.c_void_p), ("acc_eval", ct.c_void_p), ("acc_F", ct.c_void_p),
                ("batch_min", ct.c_int32), ("batch_max", ct.c_int32), ("rng_discard", ct.c_int32),
                ("order", ct.c_int32), ("loop", ct.c_int32), ("resume", ct.c_int32),
                ("log_x", ct.c_void_p), ("log_y", ct.c_void_p), ("log_cap", ct.c_int64),
                ("hist_tile", ct.c_void_p), ("hist_dec", ct.c_void_p), ("hist_perf", ct.c_void_p),
                ("hist_cap", ct.c_int64), ("driver_state", ct.c_void_p)]


DRIVER_STATE_WORDS = 2048


class DbsResult(ct.Structure):
    _fields_ = [("flip_count", ct.c_int64), ("accepts", ct.c_int64), ("speculative_evals", ct.c_int64),
                ("batches", ct.c_int64), ("F_start", ct.c_double), ("F_final", ct.c_double),
                ("seconds", ct.c_double), ("accept_seconds", ct.c_double), ("flushes", ct.c_int32),
                ("converged", ct.c_int32), ("logged", ct.c_int64), ("ranked_scored", ct.c_int64),
                ("full_sweeps", ct.c_int64), ("ranked_steps", ct.c_int64), ("ranked_hits", ct.c_int64),
                ("hist_count", ct.c_int64)]


class RankParams(ct.Structure):
    _fields_ = [("top_k", ct.c_int32), ("window", ct.c_int32), ("min_hit_rate", ct.c_float),
                ("reserved", ct.c_int32), ("x_mean", ct.c_void_p), ("x_std", ct.c_void_p),
                ("y_mean", ct.c_void_p), ("y_std", ct.c_void_p)]


class KStat(ct.Structure):
    _fields_ = [("launches", ct.c_int64), ("total_ms", ct.c_double), ("bytes", ct.c_double),
                ("flops", ct.c_double)]


_lib = None


def lib():
    """Load libpngf_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libpngf_gpu.so is not built (run `python -m pngf_amd.build`)")
    L = ct.CDLL(LIB_PATH)
    P, I, D = ct.c_void_p, ct.c_int32, ct.c_double
    sig = {
        "pngf_create": (I, [ct.POINTER(Config), ct.POINTER(P)]),
        "pngf_destroy": (I, [P]),
        "pngf_last_error": (ct.c_char_p, [P]),
        "pngf_layout_size": (I, [I, I, I, I, I]),
        "pngf_set_frequency": (I, [P, I, D]),
        "pngf_load_g": (I, [P, I, ct.c_char_p]),
        "pngf_set_g": (I, [P, I, P]),
        "pngf_synth_g": (I, [P, I]),
        "pngf_perturb_g": (I, [P, I, D, I]),
        "pngf_synth_g_host": (I, [I, I, I, I, I, I, D, D, P]),
        "pngf_set_tile_map": (I, [P, P]),
        "pngf_set_source_port": (I, [P, I, I, I, I]),
        "pngf_init_state": (I, [P]),
        "pngf_eval": (I, [P, P, I, P, P]),
        "pngf_eval_batch": (I, [P, P, P, I, P, P]),
        "pngf_accept": (I, [P, P, I]),
        "pngf_get_state": (I, [P, P, P, P, P, P]),
        "pngf_get_tile_map": (I, [P, P]),
        "pngf_get_cinv": (I, [P, I, P]),
        "pngf_get_w": (I, [P, I, P]),
        "pngf_dbs_run": (I, [P, ct.POINTER(DbsParams), ct.POINTER(DbsResult)]),
        "pngf_profile": (I, [P, I]),
        "pngf_kernel_stats": (I, [P, I, ct.POINTER(KStat)]),
        "pngf_get_g": (I, [P, I, P]),
        "pngf_mlp_create": (I, [I, I, I, I, ct.POINTER(P)]),
        "pngf_mlp_destroy": (I, [P]),
        "pngf_mlp_last_error": (ct.c_char_p, [P]),
        "pngf_mlp_num_params": (I, [P]),
        "pngf_mlp_get_params": (I, [P, P]),
        "pngf_mlp_set_params": (I, [P, P]),
        "pngf_mlp_predict": (I, [P, P, I, P]),
        "pngf_mlp_predict_dev": (I, [P, P, I, P, P]),
        "pngf_mlp_grad": (I, [P, P, P, I, P, P]),
        "pngf_mlp_train": (I, [P, P, P, I, I, I, ct.c_float, I, P]),
        "pngf_ranker_dims": (I, [P, ct.POINTER(I), ct.POINTER(I)]),
        "pngf_dbs_run_many": (I, [P, I, P, P]),
        "pngf_dbs_run_pool": (I, [P, I, P, P, I]),
        "pngf_autotune_result": (I, [P, ct.POINTER(I), ct.POINTER(I), P]),
        "pngf_symmetry": (I, [P, ct.POINTER(I), ct.POINTER(ct.c_double)]),
        "pngf_factor_capacity": (I, [P, ct.POINTER(I), ct.POINTER(I)]),
        "pngf_sweep": (I, [P, ct.c_void_p, ct.c_void_p, I, ct.c_double, ct.POINTER(I), ct.POINTER(ct.c_double),
                           ct.POINTER(ct.c_int64)]),
        "pngf_features": (I, [P, P, I, P]),
        "pngf_dbs_run_ranked": (I, [P, P, ct.POINTER(RankParams), ct.POINTER(DbsParams),
                                    ct.POINTER(DbsResult)]),
        "pngf_checkpoint": (I, [P]),
        "pngf_restore": (I, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def layout_size(layout, tiles_x, tiles_y=None, cells_x=3, cells_y=3):
    return lib().pngf_layout_size(layout, tiles_x, tiles_y or tiles_x, cells_x, cells_y)


def synth_g_host(layout, tiles, num_eobj, hz, cells=3, soft=0.0):
    """Synthetic G in acmefdfd's file layout: (M, M) complex128, column-major (order='F').
    soft: softening radius in cells (0 = default 1.0)."""
    N = layout_size(layout, tiles, tiles, cells, cells)
    M = N + num_eobj - 1
    buf = np.zeros((M, M), dtype=np.complex128, order="F")
    rc = lib().pngf_synth_g_host(layout, tiles, tiles, cells, cells, num_eobj, hz, soft, _p(buf))
    if rc != 0:
        raise PNGFError(rc, "pngf_synth_g_host")
    return buf


def _mix64(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def perturb_block(G, N, rel, seed=1):
    """G (file layout, at least N x N) with its N x N block perturbed as pngf_perturb_g does."""
    G = np.array(G, dtype=np.complex128, order="F")
    blk = G[:N, :N]
    amax = np.max(np.abs(blk.real) + np.abs(blk.imag))
    i, j = np.meshgrid(np.arange(N, dtype=np.uint64), np.arange(N, dtype=np.uint64), indexing="ij")
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) << np.uint64(40)) + np.uint64(2) * (i * np.uint64(N) + j)
        u1 = (_mix64(key) >> np.uint64(11)).astype(np.float64) * 2.0 ** -52 - 1.0
        u2 = (_mix64(key + np.uint64(1)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -52 - 1.0
    amp = rel * amax
    G[:N, :N] = (blk.real + amp * u1) + 1j * (blk.imag + amp * u2)
    return G


def write_history(path, seed, start_map, r, end_map):
    """The reference Logger's text file (logger.hpp:75-107) from dbs_run(..., hist_cap=n):
    "seed records", the start tile map as x/. in index order, one line "i tile decision
    perf..." per record (C++ stream formatting: %g), the final tile map."""
    with open(path, "w") as fh:
        fh.write("%d %d\n" % (seed, len(r["hist_tile"])))
        fh.write("".join("x" if v else "." for v in start_map) + "\n")
        for i, (t, d, pr) in enumerate(zip(r["hist_tile"], r["hist_dec"], r["hist_perf"])):
            fh.write("%d %d %d " % (i, t, d) + "".join("%g " % v for v in pr) + "\n")
        fh.write("".join("x" if v else "." for v in end_map) + "\n")


def save_run(path, seed, tile_map, runs):
    """Resume point of a DBS run, as JSON: the seed, the tile map reached, the flips accepted so far
    (primary tile, evaluation index counted from the start of the whole run, objective after the
    accept), the evaluations done and the driver position.  ``runs`` are the dbs_run results of the
    run so far, in order.  A run continues with load_run + a handle built from the tile map
    (set_tile_map, init_state: C^-1 formed again) + dbs_run(resume_from=...)."""
    import json
    tiles, evals, F, done = [], [], [], 0
    for r in runs:
        tiles += [int(t) for t in r["acc_tiles"]]
        evals += [int(e) + done for e in r["acc_eval"]]
        F += [float(x) for x in r["acc_F"]]
        done += int(r["flip_count"])
    rec = dict(format="pngf-dbs-resume-1", seed=int(seed), tile_map=[int(v) for v in tile_map],
               acc_tiles=tiles, acc_eval=evals, acc_F=F, flip_count=done,
               converged=bool(runs[-1]["converged"]) if runs else False,
               driver_state=[int(v) for v in runs[-1]["driver_state"]])
    with open(path, "w") as fh:
        json.dump(rec, fh)


def load_run(path):
    """The record written by save_run, with numpy arrays for the tile map and driver position."""
    import json
    with open(path) as fh:
        rec = json.load(fh)
    if rec.get("format") != "pngf-dbs-resume-1":
        raise ValueError("%s is not a DBS resume record" % path)
    rec["tile_map"] = np.array(rec["tile_map"], dtype=np.int32)
    rec["driver_state"] = np.array(rec["driver_state"], dtype=np.uint32)
    return rec


class Optimizer:
    """All frequencies of one design problem on one GPU (reference: one PNGF_Optimizer per frequency)."""

    def __init__(self, layout, tiles, num_freq=1, num_eobj=1, max_flips=1, method=METHOD_LOWRANK,
                 objective=OBJ_S11, driver=DRIVER_RASTER, max_deferred=0, device=0, mem_limit=0,
                 dx=0.0, dy=0.0, dz=0.0, z0=50.0, cells=3, synth_soft=0.0, tile_variant=0, symmetry=0,
                 factors=0, precision=PREC_C128):
        self.cfg = Config(layout, tiles, tiles, cells, cells, num_eobj, max_flips, num_freq, method,
                          objective, driver, max_deferred, device, tile_variant, dx, dy, dz, z0, mem_limit,
                          synth_soft, symmetry, factors, precision)
        self.h = ct.c_void_p()
        rc = lib().pngf_create(ct.byref(self.cfg), ct.byref(self.h))
        if rc != 0:
            msg = lib().pngf_last_error(self.h).decode() if self.h else "invalid configuration"
            if self.h:
                lib().pngf_destroy(self.h)
                self.h = None
            raise PNGFError(rc, msg)
        self.N = layout_size(layout, tiles, tiles, cells, cells)
        self.T, self.nf, self.ne, self.max_flips = tiles, num_freq, num_eobj, max_flips

    def close(self):
        if getattr(self, "h", None):
            lib().pngf_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ck(self, rc, what):
        if rc != 0:
            raise PNGFError(rc, "%s: %s" % (what, lib().pngf_last_error(self.h).decode()))

    # ---- inputs (pngfopt.hpp:66-191, 559-618, 669-674)
    def set_frequency(self, f, hz):
        self._ck(lib().pngf_set_frequency(self.h, f, hz), "set_frequency")

    def load_g(self, f, path):
        self._ck(lib().pngf_load_g(self.h, f, path.encode()), "load_g")

    def set_g(self, f, file_buf):
        fb = np.asfortranarray(file_buf, dtype=np.complex128)
        M = self.N + self.ne - 1
        if fb.shape != (M, M):
            raise ValueError("G buffer must be %dx%d" % (M, M))
        self._ck(lib().pngf_set_g(self.h, f, _p(fb)), "set_g")

    def synth_g(self, f):
        self._ck(lib().pngf_synth_g(self.h, f), "synth_g")

    def perturb_g(self, f, rel, seed=1):
        """Non-reciprocal perturbation of slot f's G: rel * max|G| * (u1 + i u2); see perturb_block."""
        self._ck(lib().pngf_perturb_g(self.h, f, float(rel), int(seed)), "perturb_g")

    def set_tile_map(self, tm):
        a = np.ascontiguousarray(tm, dtype=np.int32)
        self._ck(lib().pngf_set_tile_map(self.h, _p(a)), "set_tile_map")

    def set_source_port(self, tx, ty, yoff=1, ysz=2):
        self._ck(lib().pngf_set_source_port(self.h, tx, ty, yoff, ysz), "set_source_port")

    def init_state(self):
        """compute_C_matrix + invertMatrix + compute_PEinc + compute_R_matrix(1), every frequency."""
        self._ck(lib().pngf_init_state(self.h), "init_state")

    # ---- hot path (pngfopt.hpp:379-539, 256-367)
    def compute_updated_Eobj(self, tiles):
        t = np.ascontiguousarray(tiles, dtype=np.int32)
        e = np.zeros((self.nf, self.ne), dtype=np.complex128)
        F = ct.c_double()
        self._ck(lib().pngf_eval(self.h, _p(t), len(t), _p(e), ct.byref(F)), "eval")
        return e, F.value

    def _pack(self, cands):
        """[n][max_flips] tile array and per-candidate counts (a one-row buffer at n = 0, so the
        library gets valid pointers and scores nothing); a flip of more tiles than max_flips is
        refused as the library refuses it."""
        n = len(cands)
        t = np.full((max(n, 1), self.max_flips), -1, dtype=np.int32)
        nt = np.zeros(max(n, 1), dtype=np.int32)
        for i, c in enumerate(cands):
            if not 1 <= len(c) <= self.max_flips:
                raise PNGFError(-1, "candidate %d has %d tiles (1..%d)" % (i, len(c), self.max_flips))
            t[i, :len(c)] = c
            nt[i] = len(c)
        return n, t, nt

    def eval_batch(self, cands):
        n, t, nt = self._pack(cands)
        F = np.zeros(n)
        e = np.zeros((n, self.nf, self.ne), dtype=np.complex128)
        self._ck(lib().pngf_eval_batch(self.h, _p(t), _p(nt), n, _p(F), _p(e)), "eval_batch")
        return F, e

    def sweep(self, cands, F_best):
        """First improving flip of a visit sequence (not committed): (index or -1, F, flips scored)."""
        n, t, nt = self._pack(cands)
        idx, F, ev = ct.c_int32(), ct.c_double(), ct.c_int64()
        self._ck(lib().pngf_sweep(self.h, _p(t), _p(nt), n, float(F_best), ct.byref(idx), ct.byref(F), ct.byref(ev)),
                 "sweep")
        return idx.value, F.value, ev.value

    def accept(self, tiles):
        """update_C_inv + compute_PEinc + compute_R_matrix(0) for the named flip."""
        t = np.ascontiguousarray(tiles, dtype=np.int32)
        self._ck(lib().pngf_accept(self.h, _p(t), len(t)), "accept")

    update_C_inv = accept

    # ---- state
    def state(self):
        e = np.zeros((self.nf, self.ne), dtype=np.complex128)
        v = np.zeros(self.nf, dtype=np.complex128)
        ln = np.zeros((self.nf, 2), dtype=np.complex128)
        s = np.zeros(self.nf, dtype=np.complex128)
        F = ct.c_double()
        self._ck(lib().pngf_get_state(self.h, _p(e), _p(v), _p(ln), _p(s), ct.byref(F)), "get_state")
        return dict(Eobj=e, Vinc=v, LN=ln, S11=s, F=F.value)

    def tile_map(self):
        a = np.zeros(self.T * self.T, dtype=np.int32)
        self._ck(lib().pngf_get_tile_map(self.h, _p(a)), "get_tile_map")
        return a

    def cinv(self, f=0):
        out = np.zeros((self.N, self.N), dtype=np.complex128, order="F")
        self._ck(lib().pngf_get_cinv(self.h, f, _p(out)), "get_cinv")
        return out

    def g(self, f=0):
        out = np.zeros((self.N, self.N), dtype=np.complex128, order="F")
        self._ck(lib().pngf_get_g(self.h, f, _p(out)), "get_g")
        return out

    def checkpoint(self):
        self._ck(lib().pngf_checkpoint(self.h), "checkpoint")

    def restore(self):
        self._ck(lib().pngf_restore(self.h), "restore")

    def w(self, f=0):
        out = np.zeros((self.N, self.N), dtype=np.complex128)
        self._ck(lib().pngf_get_w(self.h, f, _p(out)), "get_w")
        return out

    # ---- DBS loop (main.cpp:338-459)
    def _dbs_params(self, max_evals, seed, max_accepts, batch_min, batch_max, rng_discard, log_cap, order=0,
                    loop=0, hist_cap=0, resume_from=None):
        acc = (np.zeros(max_accepts, dtype=np.int32), np.zeros(max_accepts, dtype=np.int64),
               np.zeros(max_accepts))
        log = None
        if log_cap:
            din, dout = self.ranker_dims()
            log = (np.zeros((log_cap, din), dtype=np.float32), np.zeros((log_cap, dout), dtype=np.float32))
        hist = None
        if hist_cap:
            hist = (np.zeros(hist_cap, dtype=np.int32), np.zeros(hist_cap, dtype=np.int32),
                    np.zeros((hist_cap, 2 + 3 * self.nf)))
        ds = np.zeros(DRIVER_STATE_WORDS, dtype=np.uint32)
        if resume_from is not None:
            ds[:] = np.asarray(resume_from, dtype=np.uint32)
        p = DbsParams(max_evals, seed, max_accepts, acc[0].ctypes.data, acc[1].ctypes.data, acc[2].ctypes.data,
                      batch_min, batch_max, rng_discard, order, loop, int(resume_from is not None),
                      log[0].ctypes.data if log else None,
                      log[1].ctypes.data if log else None, log_cap,
                      hist[0].ctypes.data if hist else None, hist[1].ctypes.data if hist else None,
                      hist[2].ctypes.data if hist else None, hist_cap, ds.ctypes.data)
        return p, acc + (ds,), (log, hist)

    @staticmethod
    def _dbs_result(r, acc, logs):
        log, hist = logs
        na = r.accepts
        out = dict(flip_count=r.flip_count, accepts=na, speculative_evals=r.speculative_evals,
                   batches=r.batches, F_start=r.F_start, F_final=r.F_final, seconds=r.seconds,
                   accept_seconds=r.accept_seconds, flushes=r.flushes, converged=bool(r.converged),
                   acc_tiles=acc[0][:na].copy(), acc_eval=acc[1][:na].copy(), acc_F=acc[2][:na].copy(),
                   ranked_scored=r.ranked_scored, full_sweeps=r.full_sweeps, ranked_steps=r.ranked_steps,
                   ranked_hits=r.ranked_hits)
        if len(acc) > 3:
            out["driver_state"] = acc[3].copy()
        if log is not None:
            out["log_x"] = log[0][:r.logged].copy()
            out["log_y"] = log[1][:r.logged].copy()
        if hist is not None:
            n = r.hist_count
            out["hist_tile"] = hist[0][:n].copy()
            out["hist_dec"] = hist[1][:n].copy()
            out["hist_perf"] = hist[2][:n].copy()
        return out

    def dbs_run(self, max_evals=0, seed=-514213645, max_accepts=100000, batch_min=0, batch_max=0,
                rng_discard=0, log_cap=0, order=ORDER_FIRST, loop=LOOP_AUTO, hist_cap=0, resume_from=None):
        """DBS to convergence (or max_evals).  hist_cap > 0 also returns the reference Logger's
        records (hist_tile, hist_dec, hist_perf; see write_history).  The result's driver_state is
        the driver position at return; passing it back as resume_from (on this handle or on one
        rebuilt from the tile map) continues the visit order from there."""
        p, acc, log = self._dbs_params(max_evals, seed, max_accepts, batch_min, batch_max, rng_discard, log_cap,
                                       order, loop, hist_cap, resume_from)
        r = DbsResult()
        self._ck(lib().pngf_dbs_run(self.h, ct.byref(p), ct.byref(r)), "dbs_run")
        return self._dbs_result(r, acc, log)

    # ---- learned ranker
    def ranker_dims(self):
        din, dout = ct.c_int32(), ct.c_int32()
        self._ck(lib().pngf_ranker_dims(self.h, ct.byref(din), ct.byref(dout)), "ranker_dims")
        return din.value, dout.value

    def features(self, tiles):
        t = np.ascontiguousarray(tiles, dtype=np.int32)
        din, _ = self.ranker_dims()
        out = np.zeros((len(t), din), dtype=np.float32)
        self._ck(lib().pngf_features(self.h, _p(t), len(t), _p(out)), "features")
        return out

    def dbs_run_ranked(self, mlp, norm, top_k=32, window=8, min_hit_rate=0.25, max_evals=0,
                       max_accepts=100000, log_cap=0):
        xm, xs, ym, ys = [np.ascontiguousarray(a, dtype=np.float32) for a in norm]
        rp = RankParams(top_k, window, min_hit_rate, 0, xm.ctypes.data, xs.ctypes.data, ym.ctypes.data,
                        ys.ctypes.data)
        p, acc, log = self._dbs_params(max_evals, 0, max_accepts, 0, 0, 0, log_cap)
        r = DbsResult()
        self._ck(lib().pngf_dbs_run_ranked(self.h, mlp.h, ct.byref(rp), ct.byref(p), ct.byref(r)),
                 "dbs_run_ranked")
        return self._dbs_result(r, acc, log)

    def autotune_result(self):
        ch, nv = ct.c_int32(), ct.c_int32()
        ms = np.zeros(16)
        self._ck(lib().pngf_autotune_result(self.h, ct.byref(ch), ct.byref(nv), _p(ms)), "autotune_result")
        return dict(chosen=ch.value, ms=ms[:nv.value].tolist())

    def symmetry(self):
        """(used, rel_asym): whether the LOWRANK accept takes W[t,:] = W[:,t]^T, and the measured
        max|G_ij - G_ji| / max|G_ij| over the frequencies."""
        u, a = ct.c_int32(), ct.c_double()
        self._ck(lib().pngf_symmetry(self.h, ct.byref(u), ct.byref(a)), "symmetry")
        return bool(u.value), a.value

    def factor_capacity(self):
        """(columns, rebuilt): deferred factor columns held before a fold, and whether M rows are
        rebuilt from A and Lam' instead of stored."""
        c, r = ct.c_int32(), ct.c_int32()
        self._ck(lib().pngf_factor_capacity(self.h, ct.byref(c), ct.byref(r)), "factor_capacity")
        return c.value, bool(r.value)

    def profile(self, enable=True):
        self._ck(lib().pngf_profile(self.h, int(enable)), "profile")

    def kernel_stats(self, which):
        k = KStat()
        self._ck(lib().pngf_kernel_stats(self.h, which, ct.byref(k)), "kernel_stats")
        return dict(launches=k.launches, total_ms=k.total_ms, bytes=k.bytes, flops=k.flops)


class MLP:
    """The learned flip ranker (din -> dh -> dh -> dout, GELU, BF16 MFMA / FP32 accumulation)."""

    def __init__(self, din, dh, dout, seed=0):
        self.h = ct.c_void_p()
        rc = lib().pngf_mlp_create(din, dh, dout, seed, ct.byref(self.h))
        if rc != 0:
            msg = lib().pngf_mlp_last_error(self.h).decode() if self.h else "bad dimensions"
            raise PNGFError(rc, msg)
        self.din, self.dh, self.dout = din, dh, dout
        self.nparam = lib().pngf_mlp_num_params(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().pngf_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ck(self, rc, what):
        if rc != 0:
            raise PNGFError(rc, "%s: %s" % (what, lib().pngf_mlp_last_error(self.h).decode()))

    def params(self):
        out = np.zeros(self.nparam, dtype=np.float32)
        self._ck(lib().pngf_mlp_get_params(self.h, _p(out)), "get_params")
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._ck(lib().pngf_mlp_set_params(self.h, _p(w)), "set_params")

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros((len(X), self.dout), dtype=np.float32)
        self._ck(lib().pngf_mlp_predict(self.h, _p(X), len(X), _p(out)), "predict")
        return out

    def grad(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = ct.c_float()
        self._ck(lib().pngf_mlp_grad(self.h, _p(X), _p(Y), len(X), _p(g), ct.byref(loss)), "grad")
        return g, loss.value

    def train(self, X, Y, epochs=10, batch=512, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        loss = ct.c_float()
        self._ck(lib().pngf_mlp_train(self.h, _p(X), _p(Y), len(X), epochs, batch, lr, seed, ct.byref(loss)),
                 "train")
        return loss.value

    def split(self, w=None):
        """Parameter vector -> (W1, b1, W2, b2, W3, b3) views."""
        w = self.params() if w is None else w
        d, h, o = self.din, self.dh, self.dout
        sizes = [d * h, h, h * h, h, h * o, o]
        shapes = [(d, h), (h,), (h, h), (h,), (h, o), (o,)]
        out, off = [], 0
        for n, s in zip(sizes, shapes):
            out.append(w[off:off + n].reshape(s))
            off += n
        return out


def dbs_run_many(opts, max_evals=0, max_accepts=100000, workers=0):
    """DBS on several independent optimizers concurrently (host worker threads, largest
    problems first; workers = 0: the library default)."""
    n = len(opts)
    hs = (ct.c_void_p * n)(*[o.h for o in opts])
    ps = (DbsParams * n)()
    rs = (DbsResult * n)()
    keep = []
    for i, o in enumerate(opts):
        p, acc, log = o._dbs_params(max_evals, 0, max_accepts, 0, 0, 0, 0)
        ps[i] = p
        keep.append(acc)
    rc = lib().pngf_dbs_run_pool(hs, n, ps, rs, workers) if workers else lib().pngf_dbs_run_many(hs, n, ps, rs)
    if rc != 0:
        msgs = [lib().pngf_last_error(o.h).decode() for o in opts]
        raise PNGFError(rc, "dbs_run_many: " + "; ".join(m for m in msgs if m))
    return [Optimizer._dbs_result(rs[i], keep[i], (None, None)) for i in range(n)]
