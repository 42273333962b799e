"""In-tree build of the gfx950 library (libo2f.so) with hipcc.

    python -m optimisation2forme_amd.build

compiles optimisation2forme_amd/csrc/*.{hip,cpp} for gfx950 and links
optimisation2forme_amd/libo2f.so next to this file.  hipcc cross-compiles,
so this runs on a machine without a GPU.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libo2f.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["host_mesh.cpp", "host_reorder.cpp", "host_tangent.cpp", "host_tiles.cpp", "o2f_kernels.hip",
           "o2f_adjoint.hip", "o2f_tangent.hip", "o2f_tiles.hip", "o2f_capi.cpp", "o2f_batch.cpp", "o2f_mlp.hip",
           "o2f_noalloc.hip"]

# -ffp-contract=off: no fused multiply-add anywhere (host or device), so the
# results round exactly like the reference's C++ expressions.  Packed FP32
# instructions are disabled for all device code.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return out


HOST = os.path.join(HERE, "host")
DRIVER = os.path.join(HERE, "bin", "navier_stokes")


def build_driver(verbose=False):
    """The reference's `navier_stokes [config]` driver and operator headers
    (host/include) over the C ABI, linked against libo2f.so."""
    srcs = sorted(os.path.join(HOST, "src", f) for f in os.listdir(os.path.join(HOST, "src"))
                  if f.endswith(".cpp"))
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(HOST, "include")] + srcs + [
           "-L" + HERE, "-lo2f", "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print("built", DRIVER)
    return DRIVER


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_driver(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
