"""CPU checks of the C-ABI library: it loads, exports every declared symbol, and the
host-only entry points (layout sizes, synthetic G) behave; no GPU compute calls."""
import os
import re

import numpy as np
import pytest

import oracle as O
import pngf_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "pngf_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pngf_\w+)\s*\(", txt, re.M)))


def test_exports_match_header():
    L = P.lib()
    decl = header_functions()
    assert decl, "no declarations parsed"
    assert sorted(P.EXPORTS) == decl
    for name in decl:
        assert hasattr(L, name), name


@pytest.mark.parametrize("layout,T,N", [
    (P.LAYOUT_SHARED, 7, 924), (P.LAYOUT_SHARED, 21, 8064), (P.LAYOUT_SHARED, 31, 17484),
    (P.LAYOUT_DISJOINT, 7, 1182), (P.LAYOUT_DISJOINT, 11, 2910), (P.LAYOUT_DISJOINT, 15, 5406),
    (P.LAYOUT_DISJOINT, 21, 10590), (P.LAYOUT_DISJOINT, 25, 15006), (P.LAYOUT_DISJOINT, 31, 23070),
])
def test_layout_size(layout, T, N):
    assert P.layout_size(layout, T) == N
    assert O.layout_size(layout, T) == N


def test_bad_layout_args():
    assert P.lib().pngf_layout_size(7, 3, 3, 3, 3) == -1


def test_synth_g_host_structure():
    T = 3
    for layout in (P.LAYOUT_SHARED, P.LAYOUT_DISJOINT):
        G = P.synth_g_host(layout, T, 3, 30e9)
        N = P.layout_size(layout, T)
        assert G.shape == (N + 2, N + 2)
        assert np.isfinite(G).all()
        g = G[:N, :N]
        # reciprocity of the dyadic kernel: G is complex symmetric
        assert np.array_equal(g, g.T)
        # normalisation: port block sums to -(50+30i)/zfac at 30 GHz
        port = O.layout_port(layout, T, 1, 1)
        zfac = (0.01 / 60) * 0.5 / (2 * (0.01 / 60) * (0.01 / 180))
        s = g[np.ix_(port, port)].sum()
        assert abs(s - (-(50 + 30j) / zfac)) < 1e-12 * abs(s)
        # NF2FF rows: x / y orientation, zero padding columns
        assert np.all(G[:, N:] == 0)
        assert np.all((G[N, :N] != 0) | (G[N + 1, :N] != 0))


def test_synth_g_host_bad_args():
    buf = np.zeros(4, dtype=np.complex128)
    assert P.lib().pngf_synth_g_host(3, 3, 3, 3, 3, 1, 30e9, 0.0, buf.ctypes.data) == -1
    assert P.lib().pngf_synth_g_host(1, 3, 3, 3, 3, 1, -1.0, 0.0, buf.ctypes.data) == -1


def test_create_rejects_bad_config():
    with pytest.raises(P.PNGFError) as e:
        P.Optimizer(P.LAYOUT_DISJOINT, 7, num_eobj=1, objective=P.OBJ_REF)
    assert e.value.code == -1
    with pytest.raises(P.PNGFError):
        P.Optimizer(P.LAYOUT_SHARED, 8, num_eobj=3, max_flips=4, driver=P.DRIVER_REF, objective=P.OBJ_REF)
    # enumerations outside their range are refused before any device is touched
    for kw in (dict(symmetry=3), dict(symmetry=P.SYM_PROJECT, method=P.METHOD_PLAIN), dict(factors=3),
               dict(factors=-1), dict(tile_variant=7)):
        with pytest.raises(P.PNGFError) as e:
            P.Optimizer(P.LAYOUT_DISJOINT, 7, **kw)
        assert e.value.code == -1, kw


def test_initial_pattern_convention():
    from pngf_amd.problem import MT19937, initial_pattern
    g = MT19937(5489)
    v = [g() for _ in range(10000)]
    assert v[-1] == 4123659995  # C++ [rand.predef] known answer
    h = O.Hat(-514213645, 120)
    g2 = MT19937(-514213645 & 0xFFFFFFFF)
    assert [h.rand_binary() for _ in range(64)] == [g2() & 1 for _ in range(64)]
    tm = initial_pattern(21, 0)
    assert tm[10 * 21 + 9] == 1 and tm[10 * 21 + 11] == 1 and tm[10 * 21 + 10] == 0
    assert tm[9 * 21 + 10] == 0 and tm[11 * 21 + 10] == 0


def header_struct(name):
    """(field, C type) pairs of `typedef struct { ... } name;` in the header."""
    txt = open(os.path.join(ROOT, "include", "pngf_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    body = re.search(r"typedef struct \{([^}]*)\}\s*" + name + ";", txt).group(1)
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"((?:const\s+)?\w+\s*\**)\s*(.*)", decl)
        ctype = m.group(1).replace(" ", "")
        for var in m.group(2).split(","):
            var = var.strip()
            stars = len(var) - len(var.lstrip("*"))
            out.append((var.lstrip("*"), ctype + "*" * stars))
    return out


CT_OF = {"int32_t": "c_int", "int64_t": "c_long", "double": "c_double", "float": "c_float"}


@pytest.mark.parametrize("cname,py", [("pngf_config", P.Config), ("pngf_dbs_params", P.DbsParams),
                                      ("pngf_dbs_result", P.DbsResult), ("pngf_rank_params", P.RankParams),
                                      ("pngf_kstat", P.KStat)])
def test_struct_mirrors_match_header(cname, py):
    fields = header_struct(cname)
    assert [f for f, _ in fields] == [f for f, _ in py._fields_]
    for (f, ct_), (_, pt) in zip(fields, py._fields_):
        if ct_.endswith("*"):
            assert pt.__name__ in ("c_void_p",) or pt.__name__.startswith("LP_"), (f, pt)
        else:
            assert pt.__name__ == CT_OF[ct_], (f, ct_, pt)


def test_write_history_layout(tmp_path):
    """write_history follows the Logger's text layout (logger.hpp:75-107) with C++ stream
    formatting of the values."""
    r = dict(hist_tile=np.array([-1, 4, 9], dtype=np.int32), hist_dec=np.array([-1, 0, 1], dtype=np.int32),
             hist_perf=np.array([[0.5, 0.5, 0.1, -0.2, 0.0], [0.6, 0.5, 0.125, 1e-5, 0.0],
                                 [0.25, 0.5, -1.5, 2.0, 12.345678]]))
    P.write_history(str(tmp_path / "log.txt"), -514213645, [1, 0, 1, 1], r, [0, 0, 1, 1])
    assert (tmp_path / "log.txt").read_text() == (
        "-514213645 3\n"
        "x.xx\n"
        "0 -1 -1 0.5 0.5 0.1 -0.2 0 \n"
        "1 4 0 0.6 0.5 0.125 1e-05 0 \n"
        "2 9 1 0.25 0.5 -1.5 2 12.3457 \n"
        "..xx\n")


def test_headers_compile_as_c_and_cpp():
    """The C ABI header is plain C99 (a cgo or FFI binding includes it as C); the C++ view
    compiles on top of it."""
    import shutil
    import subprocess
    inc = os.path.join(ROOT, "include")
    if shutil.which("gcc") is None or shutil.which("g++") is None:
        pytest.skip("no host compiler")
    r = subprocess.run(["gcc", "-fsyntax-only", "-std=c99", "-Wall", "-Werror", "-x", "c",
                        os.path.join(inc, "pngf_gpu.h")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run(["g++", "-fsyntax-only", "-std=c++17", "-Wall", "-I", inc, "-x", "c++", "-"],
                       input='#include "pngf_optimizer.hpp"\nint main() { return 0; }\n', capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr

