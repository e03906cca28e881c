"""Builds libpngf_gpu.so (the C-ABI library of include/pngf_gpu.h) for gfx950.

hipcc cross-compiles without a GPU, so this runs anywhere ROCm is installed.
Usage: python -m pngf_amd.build   (or pngf_amd.build.build() from Python).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpngf_gpu.so")
SOURCES = ["kernels.hip", "runtime.hip", "mlp.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions stay off: the compiler is not allowed to form them
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]
LIBS = ["-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-Wl,-rpath,/opt/rocm/lib"]


def build(verbose=False):
    """Compile every HIP source of the library and link the shared object."""
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        r = subprocess.run(cmd, capture_output=True, text=True)
        err = "\n".join(line for line in r.stderr.splitlines() if "packed-fp32-ops" not in line)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed for %s:\n%s" % (src, err))
        if err.strip() and verbose:
            print(err)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + LIBS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    # main.cpp-shaped command-line driver over the C ABI (host code only)
    cli_src = os.path.join(HERE, "cli", "pngf_opt.cpp")
    cli = os.path.join(HERE, "cli", "pngf_opt")
    cmd = [HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), cli_src, "-o", cli,
           "-L" + HERE, "-lpngf_gpu", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("cli build failed:\n%s" % r.stderr)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
