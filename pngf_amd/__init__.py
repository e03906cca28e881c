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
                ("device", ct.c_int32), ("reserved", ct.c_int32), ("dx", ct.c_double),
                ("dy", ct.c_double), ("dz", ct.c_double), ("z0", ct.c_double),
                ("mem_limit", ct.c_int64)]


class DbsParams(ct.Structure):
    _fields_ = [("max_evals", ct.c_int64), ("seed", ct.c_int32), ("max_accepts", ct.c_int32),
                ("acc_tiles", ct.c_void_p), ("acc_eval", ct.c_void_p), ("acc_F", ct.c_void_p),
                ("batch_min", ct.c_int32), ("batch_max", ct.c_int32), ("rng_discard", ct.c_int32)]


class DbsResult(ct.Structure):
    _fields_ = [("flip_count", ct.c_int64), ("accepts", ct.c_int64), ("speculative_evals", ct.c_int64),
                ("batches", ct.c_int64), ("F_start", ct.c_double), ("F_final", ct.c_double),
                ("seconds", ct.c_double), ("accept_seconds", ct.c_double), ("flushes", ct.c_int32),
                ("converged", ct.c_int32)]


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
        "pngf_synth_g_host": (I, [I, I, I, I, I, I, D, P]),
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


def synth_g_host(layout, tiles, num_eobj, hz, cells=3):
    """Synthetic G in acmefdfd's file layout: (M, M) complex128, column-major (order='F')."""
    N = layout_size(layout, tiles, tiles, cells, cells)
    M = N + num_eobj - 1
    buf = np.zeros((M, M), dtype=np.complex128, order="F")
    rc = lib().pngf_synth_g_host(layout, tiles, tiles, cells, cells, num_eobj, hz, _p(buf))
    if rc != 0:
        raise PNGFError(rc, "pngf_synth_g_host")
    return buf


class Optimizer:
    """All frequencies of one design problem on one GPU (reference: one PNGF_Optimizer per frequency)."""

    def __init__(self, layout, tiles, num_freq=1, num_eobj=1, max_flips=1, method=METHOD_LOWRANK,
                 objective=OBJ_S11, driver=DRIVER_RASTER, max_deferred=0, device=0, mem_limit=0,
                 dx=0.0, dy=0.0, dz=0.0, z0=50.0, cells=3):
        self.cfg = Config(layout, tiles, tiles, cells, cells, num_eobj, max_flips, num_freq, method,
                          objective, driver, max_deferred, device, 0, dx, dy, dz, z0, mem_limit)
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

    def eval_batch(self, cands):
        n = len(cands)
        t = np.full((n, self.max_flips), -1, dtype=np.int32)
        nt = np.zeros(n, dtype=np.int32)
        for i, c in enumerate(cands):
            t[i, :len(c)] = c
            nt[i] = len(c)
        F = np.zeros(n)
        e = np.zeros((n, self.nf, self.ne), dtype=np.complex128)
        self._ck(lib().pngf_eval_batch(self.h, _p(t), _p(nt), n, _p(F), _p(e)), "eval_batch")
        return F, e

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
    def dbs_run(self, max_evals=0, seed=-514213645, max_accepts=100000, batch_min=0, batch_max=0,
                rng_discard=0):
        acc_t = np.zeros(max_accepts, dtype=np.int32)
        acc_e = np.zeros(max_accepts, dtype=np.int64)
        acc_F = np.zeros(max_accepts)
        p = DbsParams(max_evals, seed, max_accepts, acc_t.ctypes.data, acc_e.ctypes.data,
                      acc_F.ctypes.data, batch_min, batch_max, rng_discard)
        r = DbsResult()
        self._ck(lib().pngf_dbs_run(self.h, ct.byref(p), ct.byref(r)), "dbs_run")
        na = r.accepts
        return dict(flip_count=r.flip_count, accepts=na, speculative_evals=r.speculative_evals,
                    batches=r.batches, F_start=r.F_start, F_final=r.F_final, seconds=r.seconds,
                    accept_seconds=r.accept_seconds, flushes=r.flushes, converged=bool(r.converged),
                    acc_tiles=acc_t[:na].copy(), acc_eval=acc_e[:na].copy(), acc_F=acc_F[:na].copy())

    def profile(self, enable=True):
        self._ck(lib().pngf_profile(self.h, int(enable)), "profile")

    def kernel_stats(self, which):
        k = KStat()
        self._ck(lib().pngf_kernel_stats(self.h, which, ct.byref(k)), "kernel_stats")
        return dict(launches=k.launches, total_ms=k.total_ms, bytes=k.bytes, flops=k.flops)
