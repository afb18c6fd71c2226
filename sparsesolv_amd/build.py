"""Build recipes of the native parts (all in-tree, under sparsesolv_amd/lib).

  libsparsesolv_amd.so  HIP kernels + solver drivers + C-ABI  (hipcc, gfx950)
  libssvoxel.so         curl-curl voxel generator            (g++, OpenMP)
  libSparseSolv.so      SparseMat / MatSolverICCG / MatSolverMRTR C++ classes (g++)
  SparseSolvPy*.so      pybind11 module with the reference's Python names

Run as `python -m sparsesolv_amd.build` or through __graft_entry__.build().
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INC = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off keeps every product and sum separately rounded (the
# reference's x86-64 arithmetic); packed FP32 instructions are disabled.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


HIP_SOURCES = ("ssgpu_kernels.hip", "ssgpu_sweep.hip", "ssgpu_icfact.hip", "ssgpu_block.hip", "ssgpu_batch.hip",
               "ssgpu_order.hip", "ssgpu_assemble.hip", "ssgpu_capi.hip")


def build_hip():
    """Each source compiled on its own (in parallel), then linked."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(LIB, "obj")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    jobs = []
    for f in HIP_SOURCES:
        src = os.path.join(CSRC, f)
        if not os.path.exists(src):
            continue
        obj = os.path.join(objdir, f.replace(".hip", ".o"))
        jobs.append(([HIPCC, *compile_flags, f"-I{INC}", f"-I{CSRC}", "-c", src, "-o", obj], obj))
    with ThreadPoolExecutor(max_workers=min(len(jobs), max(1, min(8, os.cpu_count() or 1)))) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    out = os.path.join(LIB, "libsparsesolv_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *[o for _, o in jobs], "-o", tmp])
    os.replace(tmp, out)
    return out


def build_voxel():
    out = os.path.join(LIB, "libssvoxel.so")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", f"-I{INC}",
          os.path.join(CSRC, "voxel_gen.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cpp():
    """libSparseSolv.so: the reference's C++ classes over the C-ABI."""
    out = os.path.join(LIB, "libSparseSolv.so")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", f"-I{INC}", os.path.join(CSRC, "sparsesolv_cpp.cpp"),
          os.path.join(CSRC, "sprs_trans.cpp"),
          f"-L{LIB}", "-lsparsesolv_amd", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_pybind():
    """SparseSolvPy extension module (pybind11) over libSparseSolv.so."""
    import sysconfig
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(LIB, "SparseSolvPy" + suffix)
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", f"-I{INC}",
          f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
          os.path.join(CSRC, "pybind_module.cpp"), f"-L{LIB}", "-lSparseSolv", "-lsparsesolv_amd",
          "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all():
    os.makedirs(LIB, exist_ok=True)
    build_voxel()
    build_hip()
    build_cpp()
    build_pybind()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
