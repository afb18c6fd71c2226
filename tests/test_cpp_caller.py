"""A C++ caller written against the reference's API compiles against
include/, links libSparseSolv.so and runs its host-side part (CPU only).

The caller (tests/cpp/eigen_caller_main.cpp) uses the VoxelFEM example's
calls (examples/VoxelFEM/main.cpp:70-95) and the Eigen-facing SparseMat
members (the Eigen::SparseMatrix / Eigen::Triplet constructors, setMatrix,
mat(), getTrueSize: SparseMat.hpp:63-78 of the reference).  Eigen is the
caller's dependency, not the library's: the test compiles against the Eigen
copy vendored in the reference tree and is skipped where that is absent.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EIGEN = "/root/reference/000_thirdparty"
LIB = os.path.join(ROOT, "sparsesolv_amd", "lib")


@pytest.mark.skipif(not os.path.isdir(os.path.join(EIGEN, "Eigen")), reason="no Eigen headers here")
def test_reference_style_caller_builds_and_runs(tmp_path):
    assert os.path.exists(os.path.join(LIB, "libSparseSolv.so")), "libSparseSolv.so not built"
    exe = str(tmp_path / "caller")
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{ROOT}/include", f"-I{EIGEN}",
                    os.path.join(ROOT, "tests", "cpp", "eigen_caller_main.cpp"), f"-L{LIB}", "-lSparseSolv",
                    f"-Wl,-rpath,{LIB}", "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "OK" in out.stdout
