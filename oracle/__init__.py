"""CPU oracle for the pngf-opt DBS Woodbury path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module.  It wraps ``liboracle_pngf.so`` (built from
``pngf_oracle.cpp`` by ``oracle/Makefile``), a plain C++ restatement of
pngf-opt/pngfopt.hpp, pngf-opt/linalg.hpp, pngf-opt/main.cpp:271-459 and
pngf-opt/magichat.hpp.  The product library never loads it.
"""
import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_pngf.so")
# the complex64 build (the reference's USE_SINGLE_PRECISION, pngf-opt/main.cpp:58-72)
_LIB_PATH_C64 = os.path.join(_HERE, "liboracle_pngf_c64.so")
_libs = {}

_P = ct.c_void_p
_I = ct.c_int
_D = ct.c_double


def build():
    """Compile the oracle (complex128 and complex64 builds) with its Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(c64=False):
    """The complex128 oracle, or with c64=True its complex64 build (same source and C ABI, whose
    arrays stay complex128: the build converts at the boundary)."""
    if c64 not in _libs:
        path = _LIB_PATH_C64 if c64 else _LIB_PATH
        if not os.path.exists(path):
            build()
        L = ct.CDLL(path)
        sig = {
            "or_layout_size": (_I, [_I, _I, _I, _I]),
            "or_layout_mask": (None, [_I, _I, _I, _I, _P, _P]),
            "or_layout_port": (_I, [_I, _I, _I, _I, _I, _I, _I, _I, _P]),
            "or_zgetrf": (_I, [_I, _P, _P]),
            "or_zgetrs": (None, [_I, _P, _P, _P]),
            "or_zgetri": (None, [_I, _P, _P]),
            "or_invert": (_I, [_I, _P]),
            "or_opt_create": (_P, [_I, _I, _I, _I, _I, _I, _P]),
            "or_opt_destroy": (None, [_P]),
            "or_opt_N": (_I, [_P]),
            "or_opt_set_tile_map": (None, [_P, _P]),
            "or_opt_set_source_port": (_I, [_P, _I, _I, _I, _I]),
            "or_opt_compute_C": (None, [_P]),
            "or_opt_invert": (_I, [_P]),
            "or_opt_set_cinv": (None, [_P, _P]),
            "or_opt_get_cinv": (None, [_P, _P]),
            "or_opt_compute_PEinc": (None, [_P]),
            "or_opt_compute_R": (None, [_P, _I]),
            "or_opt_eval": (None, [_P, _P, _I, _P]),
            "or_opt_last_k": (_I, [_P]),
            "or_opt_last_info": (_I, [_P]),
            "or_opt_update": (None, [_P]),
            "or_opt_scalars": (None, [_P, _P]),
            "or_opt_get_vec": (None, [_P, _I, _P]),
            "or_opt_get_tile_map": (None, [_P, _P]),
            "or_objective": (None, [_I, _I, _P, _I, _P, _P, _P, _D, _D, _D, _P, _P, _P]),
            "or_hat_create": (_P, [_I, _I]),
            "or_hat_destroy": (None, [_P]),
            "or_hat_draw": (_I, [_P]),
            "or_hat_reset": (None, [_P]),
            "or_hat_rand_binary": (_I, [_P]),
            "or_mt19937_nth": (ct.c_uint, [ct.c_uint, _I]),
            "or_dbs_run": (_I, [_P, _I, _I, _P, _D, _D, _D, _I, _I, _I,
                                _P, _P, _P, _P, _I, _P, _P, _D, _P, _P, _P, _I, _P]),
            "or_num_threads": (_I, []),
            "or_is_c64": (_I, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _libs[c64] = L
    return _libs[c64]


def _ptr(a):
    return a.ctypes.data_as(ct.c_void_p)


LAYOUT_SHARED, LAYOUT_DISJOINT = 0, 1
OBJ_REF, OBJ_S11 = 0, 1


def layout_size(kind, T, cx=3, cy=3):
    return lib().or_layout_size(kind, T, cx, cy)


def layout_mask(kind, T, tile_map, cx=3, cy=3):
    tm = np.ascontiguousarray(tile_map, dtype=np.int32)
    P = np.zeros(layout_size(kind, T, cx, cy), dtype=np.int32)
    lib().or_layout_mask(kind, T, cx, cy, _ptr(tm), _ptr(P))
    return P


def layout_port(kind, T, tx, ty, yoff=1, ysz=2, cx=3, cy=3):
    out = np.zeros(64, dtype=np.int32)
    n = lib().or_layout_port(kind, T, cx, cy, tx, ty, yoff, ysz, _ptr(out))
    return out[:n].copy()


def zgetrf(a):
    """LU of a square complex matrix (returns column-major factors, 0-based ipiv, info)."""
    a = np.asfortranarray(a, dtype=np.complex128).copy(order="F")
    n = a.shape[0]
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().or_zgetrf(n, _ptr(a), _ptr(ipiv))
    return a, ipiv, info


def zgetrs(lu, ipiv, b):
    b = np.array(b, dtype=np.complex128)
    lib().or_zgetrs(lu.shape[0], _ptr(lu), _ptr(ipiv), _ptr(b))
    return b


def zgetri(lu, ipiv):
    a = np.asfortranarray(lu).copy(order="F")
    lib().or_zgetri(a.shape[0], _ptr(a), _ptr(ipiv))
    return a


def invert(a):
    a = np.asfortranarray(a, dtype=np.complex128).copy(order="F")
    info = lib().or_invert(a.shape[0], _ptr(a))
    return a, info


class Optimizer:
    """One frequency's PNGF_Optimizer (pngf-opt/pngfopt.hpp:18-679), on the CPU; c64=True runs
    it in complex64 like the reference's USE_SINGLE_PRECISION build."""

    def __init__(self, kind, T, file_buf, ne=1, max_flips=1, cx=3, cy=3, c64=False):
        fb = np.asfortranarray(file_buf, dtype=np.complex128)
        self.c64 = c64
        self.L = lib(c64)
        self.h = self.L.or_opt_create(kind, T, cx, cy, ne, max_flips, _ptr(fb))
        self.N = self.L.or_opt_N(self.h)
        self.ne = ne
        self.T = T

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_opt_destroy(self.h)
            self.h = None

    def set_tile_map(self, tm):
        tm = np.ascontiguousarray(tm, dtype=np.int32)
        self.L.or_opt_set_tile_map(self.h, _ptr(tm))

    def set_source_port(self, tx, ty, yoff=1, ysz=2):
        return self.L.or_opt_set_source_port(self.h, tx, ty, yoff, ysz)

    def compute_C_matrix(self):
        self.L.or_opt_compute_C(self.h)

    def invert(self):
        return self.L.or_opt_invert(self.h)

    def set_cinv(self, cinv_colmajor):
        c = np.asfortranarray(cinv_colmajor, dtype=np.complex128)
        self.L.or_opt_set_cinv(self.h, _ptr(c))

    def cinv(self):
        out = np.zeros((self.N, self.N), dtype=np.complex128, order="F")
        self.L.or_opt_get_cinv(self.h, _ptr(out))
        return out

    def compute_PEinc(self):
        self.L.or_opt_compute_PEinc(self.h)

    def compute_R_matrix(self, flag):
        self.L.or_opt_compute_R(self.h, flag)

    def init_state(self, tile_map, port):
        self.set_tile_map(tile_map)
        self.set_source_port(*port)
        self.compute_C_matrix()
        info = self.invert()
        self.compute_PEinc()
        self.compute_R_matrix(1)
        return info

    def compute_updated_Eobj(self, tiles):
        t = np.ascontiguousarray(tiles, dtype=np.int32)
        out = np.zeros(self.ne, dtype=np.complex128)
        self.L.or_opt_eval(self.h, _ptr(t), len(t), _ptr(out))
        return out

    def last_k(self):
        return self.L.or_opt_last_k(self.h)

    def update_C_inv(self):
        self.L.or_opt_update(self.h)

    def accept(self):
        """update_C_inv + compute_PEinc + compute_R_matrix(0) (main.cpp:449-451)."""
        self.update_C_inv()
        self.compute_PEinc()
        self.compute_R_matrix(0)

    def scalars(self):
        out = np.zeros(self.ne + 3, dtype=np.complex128)
        self.L.or_opt_scalars(self.h, _ptr(out))
        return dict(Eobj=out[: self.ne].copy(), Vinc=out[self.ne], LN1inc=out[self.ne + 1],
                    LN2inc=out[self.ne + 2])

    def vec(self, which):
        n = self.N * (self.ne if which == 3 else 1)
        out = np.zeros(n, dtype=np.complex128)
        self.L.or_opt_get_vec(self.h, {"S": 0, "Einc": 1, "PEinc": 2, "R": 3}[which], _ptr(out))
        return out.reshape(self.N, self.ne).T.copy() if which == 3 else out

    def tile_map(self):
        out = np.zeros(self.T * self.T, dtype=np.int32)
        self.L.or_opt_get_tile_map(self.h, _ptr(out))
        return out


def objective(mode, eobj, vinc, ln, freqs, dx, dy, dz):
    """Objective of main.cpp:277-289 (mode OBJ_REF) or mean |S11|^2 (OBJ_S11)."""
    eobj = np.ascontiguousarray(eobj, dtype=np.complex128)
    nf, ne = eobj.shape
    vinc = np.ascontiguousarray(vinc, dtype=np.complex128)
    ln = np.ascontiguousarray(ln if ln is not None else np.zeros((nf, 2)), dtype=np.complex128)
    fr = np.ascontiguousarray(freqs, dtype=np.float64)
    F = ct.c_double()
    s11 = np.zeros(nf, dtype=np.complex128)
    d = np.zeros(nf)
    lib().or_objective(mode, nf, _ptr(eobj), ne, _ptr(vinc), _ptr(ln), _ptr(fr), dx, dy, dz,
                       ct.byref(F), _ptr(s11), _ptr(d))
    return F.value, s11, d


class Hat:
    """MagicHat (pngf-opt/magichat.hpp:32-111), seed != 0 branch."""

    def __init__(self, seed, n):
        self.h = lib().or_hat_create(seed, n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_hat_destroy(self.h)

    def draw(self):
        return lib().or_hat_draw(self.h)

    def reset(self):
        lib().or_hat_reset(self.h)

    def rand_binary(self):
        return lib().or_hat_rand_binary(self.h)


def mt19937_nth(seed, n):
    return lib().or_mt19937_nth(seed & 0xFFFFFFFF, n)


def dbs_run(opts, mode, freqs, dx, dy, dz, seed=0, max_evals=0, max_acc=100000, trace_cap=0,
            max_seconds=0.0, hist_cap=0):
    """Full DBS loop over all frequencies (main.cpp:338-459 or the raster S11 driver), in the
    precision of the optimizers' build (all complex128 or all complex64)."""
    nf = len(opts)
    assert len({o.c64 for o in opts}) == 1
    arr = (ct.c_void_p * nf)(*[o.h for o in opts])
    T = opts[0].T
    acc_t = np.zeros(max_acc, dtype=np.int32)
    acc_e = np.zeros(max_acc, dtype=np.int32)
    acc_F = np.zeros(max_acc)
    tr = np.zeros(max(trace_cap, 1))
    fm = np.zeros(T * T, dtype=np.int32)
    st = np.zeros(6)
    fr = np.ascontiguousarray(freqs, dtype=np.float64)
    ht = np.zeros(max(hist_cap, 1), dtype=np.int32)
    hd = np.zeros(max(hist_cap, 1), dtype=np.int32)
    hp = np.zeros((max(hist_cap, 1), 2 + 3 * nf))
    hc = np.zeros(1, dtype=np.int32)
    h_on = hist_cap > 0
    rc = opts[0].L.or_dbs_run(arr, nf, mode, _ptr(fr), dx, dy, dz, seed, max_evals, max_acc,
                          _ptr(acc_t), _ptr(acc_e), _ptr(acc_F), _ptr(tr) if trace_cap else None,
                          trace_cap, _ptr(fm), _ptr(st), max_seconds,
                          _ptr(ht) if h_on else None, _ptr(hd) if h_on else None, _ptr(hp) if h_on else None,
                          hist_cap, _ptr(hc))
    na = int(st[1])
    nh = int(hc[0])
    return dict(rc=rc, flip_count=int(st[0]), accepts=na, F_start=st[2], F_final=st[3],
                loop_s=st[4], eval_s=st[5], acc_tiles=acc_t[:na].copy(), acc_eval=acc_e[:na].copy(),
                acc_F=acc_F[:na].copy(), F_trace=tr[: min(trace_cap, int(st[0]))].copy(),
                final_map=fm, hist_tile=ht[:nh].copy(), hist_dec=hd[:nh].copy(), hist_perf=hp[:nh].copy())


def num_threads():
    return lib().or_num_threads()
