"""Write the golden fixtures of the oracle tests from the compiled reference.

Runs the reference's own solver classes (built from its sources by
oracle/Makefile into oracle/_ref/libssref.so, in the container that holds the
reference tree) on small generated systems and stores inputs and outputs as
.npz data: CSR, right-hand side, solver settings, returned x, residual log,
convergence flag, IC factor and kernel outputs.  Regenerate with
    make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from sparsesolv_amd import voxel  # noqa: E402

# (name, system factory)
SYSTEMS = [
    ("real_n4", lambda: voxel.curlcurl(4, seed=0, iron_block=1)),
    ("real_n6", lambda: voxel.curlcurl(6, seed=1, iron_block=2)),
    ("cplx_n5", lambda: voxel.eddy_current(5, seed=2, block=2)),
]

# (case, solver, options, normb_mode, max_ite, x0 kind)
CASES = [
    ("iccg_a105", "iccg", dict(accel_factor=1.05), 2, 500, None),
    ("iccg_a12", "iccg", dict(accel_factor=1.2), 2, 500, None),
    ("iccg_auto", "iccg", dict(accel_factor=-1.3), 2, 500, None),
    ("iccg_diag", "iccg", dict(accel_factor=1.2, is_diag_scale=1), 2, 500, None),
    ("iccg_ptr_naive", "iccg", dict(accel_factor=1.1), 1, 500, None),
    ("iccg_warm", "iccg", dict(accel_factor=1.2, is_init=0), 2, 500, "rand"),
    ("iccg_best", "iccg", dict(accel_factor=1.2, is_save_best=1), 2, 3, None),
    ("iccg_diverge", "iccg", dict(accel_factor=1.2, diverge_judge_type=1, bad_div_val=0.5, bad_div_count_thres=2),
     2, 500, None),
    ("iccg_norm_r0", "iccg", dict(accel_factor=1.2, conv_normalize_type=1), 2, 500, None),
    ("iccg_norm_const", "iccg", dict(accel_factor=1.2, conv_normalize_type=2, conv_normalize_const=1e3), 2, 500,
     None),
    ("mrtr_ic", "mrtr", dict(accel_factor=1.2, precondition_type=0), 2, 500, None),
    ("mrtr_ic_diag", "mrtr", dict(accel_factor=1.1, precondition_type=0, is_diag_scale=1), 2, 500, None),
    ("mrtr_ic_warm", "mrtr", dict(accel_factor=1.2, precondition_type=0, is_init=0), 2, 500, "rand"),
    ("mrtr_sgs", "mrtr", dict(precondition_type=1), 2, 500, None),
    ("mrtr_sgs_warm", "mrtr", dict(precondition_type=1, is_init=0), 2, 500, "rand"),
    ("mrtr_sgs_best", "mrtr", dict(precondition_type=1, is_save_best=1), 2, 2, None),
    ("iccg_b0", "iccg", dict(accel_factor=1.2), 2, 500, "zero_b"),
]


def main():
    if not O.reference_available():
        raise SystemExit("oracle/_ref/libssref.so missing: run make -C oracle in the reference container")
    for name, make in SYSTEMS:
        s = make()
        rng = np.random.default_rng(7)
        cplx = s.is_complex
        def rvec():
            v = rng.standard_normal(s.n)
            return v + 1j * rng.standard_normal(s.n) if cplx else v
        out = dict(row_ptr=s.row_ptr, col_idx=s.col_idx, values=s.values, rhs=s.rhs)
        xs, a2 = rvec(), rvec()
        out["k_x"] = xs
        out["k_a"] = a2
        out["k_spmv"] = O.ref_spmv(s.row_ptr, s.col_idx, s.values, xs)
        out["k_dot"] = np.array([O.ref_dot(xs, a2)])
        out["k_norm"] = np.array([O.ref_norm(xs)])
        for acc in (1.0, 1.2):
            L = O.ref_ic_decomp(s.row_ptr, s.col_idx, s.values, acc)
            tag = f"ic{int(acc * 100)}"
            out[f"{tag}_rp"], out[f"{tag}_ci"], out[f"{tag}_v"], out[f"{tag}_d"] = L
        names = []
        for case, solver, opts, nbm, max_ite, x0kind in CASES:
            b = s.rhs
            x0 = None
            if x0kind == "rand":
                x0 = 1e-3 * rvec() * np.abs(s.rhs).max()
            elif x0kind == "zero_b":
                b = np.zeros_like(s.rhs)
            conv, x, log = O.ref_solve(solver, s.row_ptr, s.col_idx, s.values, b, x0, max_ite=max_ite,
                                       normb_mode=nbm, **opts)
            out[f"{case}_x0"] = x0 if x0 is not None else np.zeros(s.n, s.values.dtype)
            out[f"{case}_b"] = b
            out[f"{case}_x"] = x
            out[f"{case}_log"] = log
            out[f"{case}_conv"] = np.array([conv])
            names.append(case)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"wrote {path}: N={s.n} nnz={s.nnz} cases={len(names)} ({os.path.getsize(path) // 1024} KiB)")


if __name__ == "__main__":
    main()
