"""Batched ICCG / COCG and IC-MRTR (ssgpu_batch_*, ssgpu_solve_batch) through the C-ABI.

Many independent systems of uneven size, and several right-hand sides of one
matrix, advance together: one launch per iteration phase for all of them,
scalars and convergence bookkeeping on the device, the iteration replayed
from a hipGraph.  Bars:

- every system's x, iteration count, final and true residual are
  bit-identical to its solo solve against the same factor
  (ssgpu_solve_factored), whatever else is in the batch, including systems
  that join while others run;
- a sample is bit-identical to the oracle run in the GPU's reduction order
  (the oracle refactors at the same shift; the factor is bit-exact);
- the hipGraph replay gives the same bits as eager launches, over many
  replays with host synchronizations between them;
- every system of a batch that converges meets the tolerance in true residual.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from sparsesolv_amd import capi, voxel

pytestmark = pytest.mark.gpu

CONV = 1e-8


def _systems(cplx=False):
    """(system, shift) of uneven sizes; one matrix appears with three right-hand
    sides (real: the n=7 system, coil moved; complex: the n=6 system with
    A u right-hand sides), so the batch sweeps them as one group."""
    if cplx:
        specs = [(4, 1, 1.1), (6, 2, 1.2), (5, 3, 1.15), (7, 4, 1.3)]
        out = [(voxel.eddy_current(n, seed=sd), a) for n, sd, a in specs]
        s6 = out[1][0]
        A = s6.scipy()
        for k in range(2):
            u = np.random.default_rng(60 + k).uniform(-1.0, 1.0, s6.n) * (1.0 + 0.5j)
            out.append((dataclasses.replace(s6, rhs=A @ u), 1.2))
    else:
        specs = [(4, 1, 1.05), (9, 2, 1.2), (6, 3, 1.3), (12, 4, 1.1), (5, 5, 1.25), (8, 6, 1.15)]
        out = [(voxel.curlcurl(n, seed=sd), a) for n, sd, a in specs]
        base = voxel.curlcurl(7, seed=7)
        for k in range(3):
            out.append((voxel.curlcurl(7, seed=7, coil_seed=k + 1), 1.2))
        assert all(np.array_equal(s.values, base.values) for s, _ in out[-3:])
    return out


class _Set:
    """Device matrices (one per distinct CSR), factorized at their shift, and the systems' vectors."""

    def __init__(self, torch, ctx, systems, precision=capi.FACTOR_F64, abmc_block=None):
        self.torch = torch
        self.mats, self.items = [], []
        cache = {}
        for s, a in systems:
            key = (s.values.tobytes(), a)
            if key not in cache:
                m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values, abmc_block=abmc_block)
                m.factorize(capi.ICCG, a, precision)
                cache[key] = m
                self.mats.append(m)
            self.items.append((s, cache[key]))

    def vectors(self, x0s=None, bs=None):
        d_b, d_x = [], []
        for k, (s, m) in enumerate(self.items):
            b = s.rhs if bs is None or bs[k] is None else bs[k]
            x0 = np.zeros_like(b) if x0s is None or x0s[k] is None else x0s[k]
            d_b.append(self.torch.from_numpy(np.ascontiguousarray(b)).cuda())
            d_x.append(self.torch.from_numpy(np.ascontiguousarray(x0)).cuda())
        self.torch.cuda.synchronize()
        return d_b, d_x

    def close(self):
        for m in self.mats:
            m.close()


def _solo(torch, ctx, m, d_b, x0, normB, opts, max_ite, solver=capi.ICCG):
    d_x = torch.from_numpy(np.ascontiguousarray(x0)).cuda()
    torch.cuda.synchronize()
    r = m.solve_factored(ctx, solver, d_b.data_ptr(), d_x.data_ptr(), conv_cri=CONV, max_ite=max_ite,
                         normB=normB, opts=opts)
    ctx.synchronize()
    return d_x.cpu().numpy(), r


def _same_as_solo(torch, ctx, st, d_b, x_b, r_b, x0, normB, opts, max_ite, label, solver=capi.ICCG):
    s, m = st
    x_s, r_s = _solo(torch, ctx, m, d_b, x0, normB, opts, max_ite, solver)
    assert r_b.converged == r_s.converged, label
    assert r_b.iterations == r_s.iterations, (label, r_b.iterations, r_s.iterations)
    assert r_b.final_residual == r_s.final_residual, label
    assert r_b.true_residual == r_s.true_residual, label
    assert r_b.refinements == r_s.refinements, label
    assert np.array_equal(x_b, x_s), label


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
@pytest.mark.parametrize("precision", [capi.FACTOR_F64, capi.FACTOR_F32], ids=["f64", "f32"])
def test_batch_bit_identical_to_solo(torch_cuda, ctx, cplx, precision):
    torch = torch_cuda
    S = _Set(torch, ctx, _systems(cplx), precision)
    try:
        d_b, d_x = S.vectors()
        normB = [float(np.linalg.norm(s.rhs)) for s, _ in S.items]
        opts = capi.default_options(factor_precision=precision)
        print(f"batch cplx={cplx} precision={precision}: {len(S.items)} systems, "
              f"N={[s.n for s, _ in S.items]}", flush=True)
        res = capi.solve_batch(ctx, [m for _, m in S.items], [t.data_ptr() for t in d_b],
                               [t.data_ptr() for t in d_x], normB, conv_cri=CONV, max_ite=3000, opts=opts)
        c2 = capi.Context(0)
        try:
            for k, (st, r) in enumerate(zip(S.items, res)):
                s, m = st
                assert r.converged, k
                assert r.true_residual <= 2 * CONV, (k, r.true_residual)
                _same_as_solo(torch, c2, st, d_b[k], d_x[k].cpu().numpy(), r, np.zeros_like(s.rhs), normB[k],
                              opts, 3000, k)
        finally:
            c2.close()
        print(f"  iterations {[r.iterations for r in res]}", flush=True)
    finally:
        S.close()


def test_batch_sample_bit_identical_to_tree_order_oracle(torch_cuda, ctx):
    torch = torch_cuda
    systems = _systems(False)[:3] + _systems(True)[:1]
    for cplx in (False, True):
        sel = [(s, a) for s, a in systems if s.is_complex == cplx]
        S = _Set(torch, ctx, sel)
        try:
            d_b, d_x = S.vectors()
            normB = [O.norm(s.rhs) for s, _ in S.items]
            res = capi.solve_batch(ctx, [m for _, m in S.items], [t.data_ptr() for t in d_b],
                                   [t.data_ptr() for t in d_x], normB, conv_cri=CONV, max_ite=3000)
            for k, ((s, a), r) in enumerate(zip(sel, res)):
                with O.reduction_order(1):
                    c_o, x_o, log_o = O.solve("iccg", s.row_ptr, s.col_idx, s.values, s.rhs, conv_cri=CONV,
                                              max_ite=3000, normB=normB[k], accel_factor=a, precondition_type=1)
                assert r.converged == c_o
                assert r.iterations == len(log_o) - 1
                assert r.final_residual == log_o[-1]
                assert np.array_equal(d_x[k].cpu().numpy(), x_o), (cplx, k)
        finally:
            S.close()


def test_batch_per_system_options_match_solo(torch_cuda, ctx):
    """Each system carries its own settings: save-best with an early stop,
    the divergence stop, normalizer types 1 and 2, warm start and b = 0,
    side by side in one batch."""
    torch = torch_cuda
    s0 = voxel.curlcurl(6, seed=11)
    systems = [(s0, 1.2)] * 7
    S = _Set(torch, ctx, systems)
    rng = np.random.default_rng(3)
    cases = [
        (dict(is_save_best=1), 3, None, None),
        (dict(diverge_judge_type=1, bad_div_val=0.5, bad_div_count_thres=2), 500, None, None),
        (dict(conv_normalize_type=1), 500, None, None),
        (dict(conv_normalize_type=2, conv_normalize_const=1e3), 500, None, None),
        (dict(is_init=0), 500, 1e-3 * rng.standard_normal(s0.n), None),
        (dict(), 500, None, np.zeros(s0.n)),
        (dict(is_save_best=1), 500, None, None),
    ]
    try:
        x0s = [c[2] for c in cases]
        bs = [c[3] for c in cases]
        d_b, d_x = S.vectors(x0s, bs)
        B = capi.Batch(ctx)
        try:
            for k, (o, max_ite, _, _) in enumerate(cases):
                b = s0.rhs if bs[k] is None else bs[k]
                B.add(S.items[k][1], d_b[k].data_ptr(), d_x[k].data_ptr(), k, normB=float(np.linalg.norm(b)),
                      conv_cri=CONV, max_ite=max_ite, opts=capi.default_options(accel_factor=1.2, **o))
            got = {}
            while len(got) < len(cases):
                running, _ = B.step(4)
                got.update(dict(B.poll()))
                assert running > 0 or len(got) == len(cases)
        finally:
            B.close()
        c2 = capi.Context(0)
        try:
            for k, (o, max_ite, x0, b) in enumerate(cases):
                bb = s0.rhs if b is None else b
                _same_as_solo(torch, c2, S.items[k], d_b[k], d_x[k].cpu().numpy(), got[k],
                              np.zeros(s0.n) if x0 is None else x0, float(np.linalg.norm(bb)),
                              capi.default_options(accel_factor=1.2, **o), max_ite, k)
        finally:
            c2.close()
        print(f"batch options: iterations {[got[k].iterations for k in range(len(cases))]}, "
              f"converged {[got[k].converged for k in range(len(cases))]}", flush=True)
        assert not got[0].converged and got[0].iterations == 3
        assert got[5].converged and got[5].iterations == 0       # b = 0: the prologue's early exit
    finally:
        S.close()


def test_continuous_batching_joins_and_leaves(torch_cuda, ctx):
    """Systems added between steps while others are iterating, finished ones
    leaving as they converge: every result is still the solo one."""
    torch = torch_cuda
    S = _Set(torch, ctx, _systems(False))
    try:
        d_b, d_x = S.vectors()
        normB = [float(np.linalg.norm(s.rhs)) for s, _ in S.items]
        opts = capi.default_options()
        B = capi.Batch(ctx)
        got = {}
        try:
            waves = [[0, 1, 2], [3, 4], [5, 6, 7, 8]]
            for w in waves:
                for k in w:
                    B.add(S.items[k][1], d_b[k].data_ptr(), d_x[k].data_ptr(), k, normB=normB[k], conv_cri=CONV,
                          max_ite=3000, opts=opts)
                B.step(3)
                got.update(dict(B.poll()))
            while len(got) < len(S.items):
                B.step(5)
                got.update(dict(B.poll()))
        finally:
            B.close()
        c2 = capi.Context(0)
        try:
            for k, st in enumerate(S.items):
                _same_as_solo(torch, c2, st, d_b[k], d_x[k].cpu().numpy(), got[k], np.zeros(st[0].n), normB[k],
                              opts, 3000, k)
        finally:
            c2.close()
    finally:
        S.close()


def test_graph_replay_equals_eager(torch_cuda, ctx, monkeypatch):
    torch = torch_cuda
    S = _Set(torch, ctx, _systems(False)[:5])
    try:
        normB = [float(np.linalg.norm(s.rhs)) for s, _ in S.items]
        outs = []
        for graph in ("1", "0"):
            monkeypatch.setenv("SSGPU_BATCH_GRAPH", graph)
            monkeypatch.setenv("SSGPU_BATCH_CHECK", "3")    # many replays, a host check between each three
            d_b, d_x = S.vectors()
            res = capi.solve_batch(ctx, [m for _, m in S.items], [t.data_ptr() for t in d_b],
                                   [t.data_ptr() for t in d_x], normB, conv_cri=CONV, max_ite=3000)
            outs.append(([t.cpu().numpy() for t in d_x], [(r.iterations, r.final_residual) for r in res]))
        assert outs[0][1] == outs[1][1]
        for a, b in zip(outs[0][0], outs[1][0]):
            assert np.array_equal(a, b)
    finally:
        S.close()


def test_batch_with_abmc_ordered_matrix(torch_cuda, ctx):
    """A system held in ABMC order takes b and x in the caller's order, as
    its solo solve does, and gives the solo bits."""
    torch = torch_cuda
    systems = _systems(False)[:3]
    S = _Set(torch, ctx, systems, abmc_block=2)
    try:
        d_b, d_x = S.vectors()
        normB = [float(np.linalg.norm(s.rhs)) for s, _ in S.items]
        res = capi.solve_batch(ctx, [m for _, m in S.items], [t.data_ptr() for t in d_b],
                               [t.data_ptr() for t in d_x], normB, conv_cri=CONV, max_ite=3000)
        c2 = capi.Context(0)
        try:
            for k, st in enumerate(S.items):
                assert res[k].converged and res[k].true_residual <= 2 * CONV
                _same_as_solo(torch, c2, st, d_b[k], d_x[k].cpu().numpy(), res[k], np.zeros(st[0].n), normB[k],
                              capi.default_options(), 3000, k)
        finally:
            c2.close()
    finally:
        S.close()


def test_batch_sweep_spin_limit_then_recovery(torch_cuda, ctx, monkeypatch):
    """A batched sweep forced to give up at once (spin limit 0, set from the
    host) fails the step with HIP_ERROR and drops the batch's systems; a new
    batch on the same context and matrices then solves as before."""
    torch = torch_cuda
    S = _Set(torch, ctx, _systems(False)[:3])
    try:
        normB = [float(np.linalg.norm(s.rhs)) for s, _ in S.items]
        d_b, d_x = S.vectors()
        ref = capi.solve_batch(ctx, [m for _, m in S.items], [t.data_ptr() for t in d_b],
                               [t.data_ptr() for t in d_x], normB, conv_cri=CONV, max_ite=3000)
        x_ref = [t.cpu().numpy() for t in d_x]
        monkeypatch.setenv("SSGPU_SPIN_LIMIT", "0")
        d_b, d_x = S.vectors()
        print("batched sweep with spin limit 0", flush=True)
        with pytest.raises(capi.SolverError) as e:
            capi.solve_batch(ctx, [m for _, m in S.items], [t.data_ptr() for t in d_b],
                             [t.data_ptr() for t in d_x], normB, conv_cri=CONV, max_ite=3000)
        assert e.value.status == capi.HIP_ERROR and "spin limit" in str(e.value)
        monkeypatch.delenv("SSGPU_SPIN_LIMIT")
        d_b, d_x = S.vectors()
        res = capi.solve_batch(ctx, [m for _, m in S.items], [t.data_ptr() for t in d_b],
                               [t.data_ptr() for t in d_x], normB, conv_cri=CONV, max_ite=3000)
        for k in range(len(S.items)):
            assert res[k].iterations == ref[k].iterations
            assert np.array_equal(d_x[k].cpu().numpy(), x_ref[k])
    finally:
        S.close()


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
@pytest.mark.parametrize("fix", [0, 1], ids=["ref_direction", "std_direction"])
def test_batched_mrtr_bit_identical_to_solo(torch_cuda, ctx, cplx, fix):
    """IC-MRTR batched (MatSolverMRTR_ICver.cpp:183-373, with the reference's
    direction p = u + t u or the standard one): every system as its solo
    solve, including the FP32 factor with its refinement pass (standard
    direction) and per-system options."""
    torch = torch_cuda
    allsys = _systems(cplx)
    systems = allsys[:4] + allsys[-2:]   # the last two share a matrix (one sweep group)
    precision = capi.FACTOR_F32 if fix else capi.FACTOR_F64
    S = _Set(torch, ctx, systems, precision)
    try:
        rng = np.random.default_rng(8)
        cases = [dict(), dict(is_save_best=1), dict(conv_normalize_type=1), dict(is_init=0), dict(),
                 dict(is_init=0)]
        x0s = [1e-3 * rng.standard_normal(s.n) * (1 + 1j if cplx else 1) if "is_init" in c else None
               for c, (s, _) in zip(cases, S.items)]
        d_b, d_x = S.vectors(x0s)
        normB = [float(np.linalg.norm(s.rhs)) for s, _ in S.items]
        B = capi.Batch(ctx, cplx, precision, solver=capi.MRTR_IC)
        got = {}
        try:
            for k, (st, c) in enumerate(zip(S.items, cases)):
                o = capi.default_options(factor_precision=precision, mrtr_direction_fix=fix, **c)
                B.add(st[1], d_b[k].data_ptr(), d_x[k].data_ptr(), k, normB=normB[k], conv_cri=CONV,
                      max_ite=3000 if k != 1 else 5, opts=o)
            while len(got) < len(S.items):
                B.step(6)
                got.update(dict(B.poll()))
        finally:
            B.close()
        c2 = capi.Context(0)
        try:
            for k, (st, c) in enumerate(zip(S.items, cases)):
                o = capi.default_options(factor_precision=precision, mrtr_direction_fix=fix, **c)
                x0 = np.zeros_like(st[0].rhs) if x0s[k] is None else x0s[k]
                _same_as_solo(torch, c2, st, d_b[k], d_x[k].cpu().numpy(), got[k], x0, normB[k], o,
                              3000 if k != 1 else 5, k, solver=capi.MRTR_IC)
        finally:
            c2.close()
        print(f"batched MRTR cplx={cplx} fix={fix}: iterations {[got[k].iterations for k in sorted(got)]}",
              flush=True)
    finally:
        S.close()


def test_batched_mrtr_tree_order_oracle(torch_cuda, ctx):
    """A batched IC-MRTR system bit-identical to the oracle's IC-MRTR in the
    GPU's reduction order."""
    torch = torch_cuda
    sel = _systems(False)[:2]
    S = _Set(torch, ctx, sel)
    try:
        d_b, d_x = S.vectors()
        normB = [O.norm(s.rhs) for s, _ in S.items]
        res = capi.solve_batch(ctx, [m for _, m in S.items], [t.data_ptr() for t in d_b],
                               [t.data_ptr() for t in d_x], normB, conv_cri=CONV, max_ite=3000, solver=capi.MRTR_IC)
        for k, ((s, a), r) in enumerate(zip(sel, res)):
            with O.reduction_order(1):
                c_o, x_o, log_o = O.solve("mrtr", s.row_ptr, s.col_idx, s.values, s.rhs, conv_cri=CONV, max_ite=3000,
                                          normB=normB[k], accel_factor=a, precondition_type=0)
            assert r.converged == c_o
            assert r.iterations == len(log_o) - 1
            assert np.array_equal(d_x[k].cpu().numpy(), x_o), k
    finally:
        S.close()
