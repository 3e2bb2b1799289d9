"""Build recipe for libwtv_gpupr.so (HIP, gfx950) and the wtivo_gpupr Python module.

    python wtivo-watertightvoxel_amd/wtv_build.py    # builds in-tree

Outputs (git-ignored, shipped with the tree):
    wtivo-watertightvoxel_amd/libwtv_gpupr.so
    wtivo-watertightvoxel_amd/wtivo_gpupr<EXT_SUFFIX>
    wtivo-watertightvoxel_amd/libwtv_delaunay.so  (host tetrahedraliser of the
        benchmark and test meshes, not part of the solver)
Objects go to wtivo-watertightvoxel_amd/_build/.
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys
import sysconfig
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

LIB = os.path.join(HERE, "libwtv_gpupr.so")
# the tetrahedraliser of the benchmark and test meshes: host C++, not part of
# the solver library
DELAUNAY_LIB = os.path.join(HERE, "libwtv_delaunay.so")
EXT = os.path.join(HERE, "wtivo_gpupr" + sysconfig.get_config_var("EXT_SUFFIX"))

HIP_SOURCES = ["wtv_solver.hip", "wtv_order.hip", "wtv_flow.hip", "wtv_graph_build.hip", "wtv_topology.hip", "wtv_udf.hip"]
HOST_SOURCES = ["wtv_server.cpp", "wtv_upload.cpp", "wtv_capi.cpp"]

# -packed-fp32-ops: no v_pk_*_f32 generation (the kernels are integer work;
# this keeps hipcc from forming packed fp32 ops anywhere in the library).
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-I", INCLUDE, "-I", CSRC,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
# -ffp-contract=off: the FP64 capacity formula must not be FMA-contracted.
HOST_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math",
    "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include"), "-I", INCLUDE, "-I", CSRC,
    "-Wall", "-Wextra", "-Wno-unused-parameter",
]


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    if verbose and (r.stderr.strip()):
        # hipcc's host pass warns about the packed-fp32 feature; keep other output
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


def build(verbose: bool = False) -> dict:
    """Compile every source (always; no up-to-date checks) and link both libraries."""
    t0 = time.time()
    os.makedirs(BUILD, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-Wl,-soname,libwtv_gpupr.so"], verbose)
    os.replace(tmp, LIB)

    tmp = DELAUNAY_LIB + ".tmp"
    _run(["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall", "-Wextra",
          os.path.join(CSRC, "wtv_delaunay.cpp"), "-o", tmp], verbose)
    os.replace(tmp, DELAUNAY_LIB)

    import pybind11

    py_inc = sysconfig.get_paths()["include"]
    tmp = EXT + ".tmp"
    _run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-fvisibility=hidden",
          "-I", pybind11.get_include(), "-I", py_inc, "-I", INCLUDE,
          os.path.join(CSRC, "pybind_wtivo_gpupr.cpp"), "-o", tmp,
          "-L", HERE, "-lwtv_gpupr", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, EXT)
    return {"lib": LIB, "ext": EXT, "delaunay": DELAUNAY_LIB, "seconds": time.time() - t0}


if __name__ == "__main__":
    info = build(verbose="-q" not in sys.argv)
    print(f"built {info['lib']} and {info['ext']} in {info['seconds']:.1f}s")
