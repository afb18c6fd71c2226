"""The C-ABI libraries load and export every entry point their headers
declare (CPU only: no compute call needs a GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBS = {
    "sparsesolv_amd.h": os.path.join(ROOT, "sparsesolv_amd", "lib", "libsparsesolv_amd.so"),
    "sparsesolv_voxel.h": os.path.join(ROOT, "sparsesolv_amd", "lib", "libssvoxel.so"),
}


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:ssgpu|ssv)_\w+)\s*\(", text)))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.mark.parametrize("header", sorted(LIBS))
def test_every_declared_symbol_is_exported(header):
    lib = LIBS[header]
    assert os.path.exists(lib), f"{lib} not built"
    ctypes.CDLL(lib)                     # loads without a GPU
    names = _declared(header)
    assert names
    missing = [n for n in names if n not in _exported(lib)]
    assert not missing, missing


def test_status_strings_and_defaults_without_gpu():
    from sparsesolv_amd import capi
    assert capi.lib().ssgpu_status_string(capi.NO_DIAGONAL) == b"missing diagonal entry"
    o = capi.default_options()
    # MatSolverBase.cpp:32-47 and MatSolverICCG.hpp:40 defaults
    assert (o.is_init, o.is_diag_scale, o.is_save_best, o.is_save_log) == (1, 0, 0, 0)
    assert (o.diverge_judge_type, o.bad_div_val, o.bad_div_count_thres) == (0, 1000.0, 1000)
    assert (o.conv_normalize_type, o.conv_normalize_const, o.accel_factor) == (0, 1.0, 1.05)


def test_bad_arguments_rejected_without_gpu():
    from sparsesolv_amd import capi
    L = capi.lib()
    assert L.ssgpu_context_create(0, None) == capi.BAD_ARG
    assert L.ssgpu_solve(None, 0, None, 1e-8, 10, 1.0, None, None, None, 0, None) == capi.BAD_ARG
