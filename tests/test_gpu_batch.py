"""Batched ICCG / COCG, IC-MRTR and SGS-MRTR (ssgpu_batch_*, ssgpu_solve_batch) through the C-ABI.

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
    """Device matrices (one per distinct CSR), factorized at their shift (SGS-MRTR: S A S and its
    triangles), and the systems' vectors."""

    def __init__(self, torch, ctx, systems, precision=capi.FACTOR_F64, abmc_block=None, solver=capi.ICCG):
        self.torch = torch
        self.mats, self.items = [], []
        cache = {}
        for s, a in systems:
            key = (s.values.tobytes(), a)
            if key not in cache:
                m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values, abmc_block=abmc_block)
                m.factorize(capi.ICCG if solver == capi.MRTR_IC else solver, a, precision)
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


@pytest.mark.parametrize("solver", [capi.ICCG, capi.MRTR_SGS], ids=["iccg", "mrtr_sgs"])
def test_graph_replay_equals_eager(torch_cuda, ctx, monkeypatch, solver):
    torch = torch_cuda
    S = _Set(torch, ctx, _systems(False)[:5], solver=solver)
    try:
        normB = [float(np.linalg.norm(s.rhs)) for s, _ in S.items]
        outs = []
        for graph in ("1", "0"):
            monkeypatch.setenv("SSGPU_BATCH_GRAPH", graph)
            monkeypatch.setenv("SSGPU_BATCH_CHECK", "3")    # many replays, a host check between each three
            d_b, d_x = S.vectors()
            res = capi.solve_batch(ctx, [m for _, m in S.items], [t.data_ptr() for t in d_b],
                                   [t.data_ptr() for t in d_x], normB, conv_cri=CONV, max_ite=3000, solver=solver)
            outs.append(([t.cpu().numpy() for t in d_x], [(r.iterations, r.final_residual) for r in res]))
        assert outs[0][1] == outs[1][1]
        for a, b in zip(outs[0][0], outs[1][0]):
            assert np.array_equal(a, b)
    finally:
        S.close()


@pytest.mark.parametrize("solver", [capi.ICCG, capi.MRTR_SGS], ids=["iccg", "mrtr_sgs"])
def test_batch_with_abmc_ordered_matrix(torch_cuda, ctx, solver):
    """A system held in ABMC order takes b and x in the caller's order, as
    its solo solve does, and gives the solo bits."""
    torch = torch_cuda
    systems = _systems(False)[:3]
    S = _Set(torch, ctx, systems, abmc_block=2, solver=solver)
    try:
        d_b, d_x = S.vectors()
        normB = [float(np.linalg.norm(s.rhs)) for s, _ in S.items]
        res = capi.solve_batch(ctx, [m for _, m in S.items], [t.data_ptr() for t in d_b],
                               [t.data_ptr() for t in d_x], normB, conv_cri=CONV, max_ite=3000, solver=solver)
        c2 = capi.Context(0)
        try:
            for k, st in enumerate(S.items):
                assert res[k].converged
                if solver == capi.ICCG:
                    assert res[k].true_residual <= 2 * CONV
                _same_as_solo(torch, c2, st, d_b[k], d_x[k].cpu().numpy(), res[k], np.zeros(st[0].n), normB[k],
                              capi.default_options(), 3000, k, solver=solver)
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


@pytest.mark.parametrize("cplx,precision", [(False, capi.FACTOR_F64), (False, capi.FACTOR_F32),
                                            (True, capi.FACTOR_F64)], ids=["real_f64", "real_f32", "complex_f64"])
def test_batched_sgs_bit_identical_to_solo(torch_cuda, ctx, cplx, precision):
    """SGS-MRTR batched (MatSolverMRTR_SGSver.cpp:42-282 behind the scaling of
    the cached solve): every system as its solo solve against the same S A S,
    with per-system options (save-best stop, normalizer type 1, warm start,
    b = 0 with a warm start, the divergence stop), systems joining in two
    waves, and one system through the host-array entry.  FP32 triangles take
    the explicit S A S product between the sweeps, as the solo solve does.
    On the complex systems the reference's operator is degenerate (it assumes
    a unit diagonal of S A S), so only the bits are compared there, over a
    bounded run."""
    torch = torch_cuda
    allsys = _systems(cplx)
    systems = allsys[:4] + allsys[-2:]   # the last two share a matrix (one sweep group)
    S = _Set(torch, ctx, systems, precision, solver=capi.MRTR_SGS)
    max_ite = 300 if cplx else 3000
    try:
        rng = np.random.default_rng(9)
        cases = [dict(), dict(is_save_best=1), dict(conv_normalize_type=1), dict(is_init=0), dict(is_init=0),
                 dict(diverge_judge_type=1, bad_div_val=0.5, bad_div_count_thres=2)]
        lims = [max_ite, 5, max_ite, max_ite, max_ite, max_ite]
        x0s = [1e-3 * rng.standard_normal(s.n) * (1 + 1j if cplx else 1) if "is_init" in c else None
               for c, (s, _) in zip(cases, S.items)]
        bs = [None] * 4 + [np.zeros_like(S.items[4][0].rhs), None]
        d_b, d_x = S.vectors(x0s, bs)
        normB = [float(np.linalg.norm(s.rhs if b is None else b)) for (s, _), b in zip(S.items, bs)]
        h_s = S.items[0][0]
        h_b = np.ascontiguousarray(h_s.rhs)
        h_x0 = 1e-3 * rng.standard_normal(h_s.n) * (1 + 1j if cplx else 1)
        h_x = h_x0.copy()
        B = capi.Batch(ctx, cplx, precision, solver=capi.MRTR_SGS)
        got = {}
        print(f"batched SGS cplx={cplx} precision={precision}: N={[s.n for s, _ in S.items]}", flush=True)
        try:
            for wave in ([0, 1, 2], [3, 4, 5]):
                for k in wave:
                    o = capi.default_options(factor_precision=precision, **cases[k])
                    B.add(S.items[k][1], d_b[k].data_ptr(), d_x[k].data_ptr(), k, normB=normB[k], conv_cri=CONV,
                          max_ite=lims[k], opts=o)
                B.step(3)
                got.update(dict(B.poll()))
            B.add_host(S.items[0][1], h_b, h_x, 100, normB=normB[0], conv_cri=CONV, max_ite=max_ite,
                       opts=capi.default_options(factor_precision=precision, is_init=0))
            while len(got) < len(S.items) + 1:
                B.step(6)
                got.update(dict(B.poll()))
        finally:
            B.close()
        c2 = capi.Context(0)
        try:
            for k, (st, c) in enumerate(zip(S.items, cases)):
                o = capi.default_options(factor_precision=precision, **c)
                x0 = np.zeros_like(st[0].rhs) if x0s[k] is None else x0s[k]
                _same_as_solo(torch, c2, st, d_b[k], d_x[k].cpu().numpy(), got[k], x0, normB[k], o, lims[k], k,
                              solver=capi.MRTR_SGS)
            d_hb = torch.from_numpy(h_b.copy()).cuda()
            _same_as_solo(torch, c2, S.items[0], d_hb, h_x, got[100], h_x0, normB[0],
                          capi.default_options(factor_precision=precision, is_init=0), max_ite, "host",
                          solver=capi.MRTR_SGS)
        finally:
            c2.close()
        its = [got[k].iterations for k in sorted(got)]
        print(f"  iterations {its}, converged {[got[k].converged for k in sorted(got)]}", flush=True)
        assert not got[1].converged and got[1].iterations == 5
        assert got[4].converged and got[4].iterations == 0      # b = 0: the prologue's early exit
        if not cplx:
            assert got[0].converged and got[3].converged and got[100].converged
    finally:
        S.close()


def test_batched_sgs_tree_order_oracle(torch_cuda, ctx):
    """A batched SGS-MRTR system bit-identical to the oracle's SGS-MRTR in the
    GPU's reduction order (the oracle scales and splits as the reference)."""
    torch = torch_cuda
    sel = _systems(False)[:2]
    S = _Set(torch, ctx, sel, solver=capi.MRTR_SGS)
    try:
        d_b, d_x = S.vectors()
        normB = [O.norm(s.rhs) for s, _ in S.items]
        res = capi.solve_batch(ctx, [m for _, m in S.items], [t.data_ptr() for t in d_b],
                               [t.data_ptr() for t in d_x], normB, conv_cri=CONV, max_ite=3000,
                               solver=capi.MRTR_SGS)
        for k, ((s, a), r) in enumerate(zip(sel, res)):
            with O.reduction_order(1):
                c_o, x_o, log_o = O.solve("mrtr", s.row_ptr, s.col_idx, s.values, s.rhs, conv_cri=CONV, max_ite=3000,
                                          normB=normB[k], accel_factor=a, precondition_type=1)
            assert r.converged == c_o
            assert r.iterations == len(log_o) - 1
            assert r.final_residual == log_o[-1]
            assert np.array_equal(d_x[k].cpu().numpy(), x_o), k
    finally:
        S.close()


def test_batched_sgs_refuses_without_sgs_preparation(torch_cuda, ctx):
    s = voxel.curlcurl(5, seed=1)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    B = capi.Batch(ctx, False, capi.FACTOR_F64, solver=capi.MRTR_SGS)
    try:
        d_b = torch_cuda.from_numpy(s.rhs.copy()).cuda()
        d_x = torch_cuda.zeros_like(d_b)
        m.factorize(capi.ICCG, 1.1)
        with pytest.raises(capi.SolverError) as e:
            B.add(m, d_b.data_ptr(), d_x.data_ptr(), 0, normB=1.0)
        assert e.value.status == capi.BAD_ARG
        m.factorize(capi.MRTR_SGS, 1.0)
        B32 = capi.Batch(ctx, False, capi.FACTOR_F32, solver=capi.MRTR_SGS)
        try:
            with pytest.raises(capi.SolverError) as e:
                B32.add(m, d_b.data_ptr(), d_x.data_ptr(), 0, normB=1.0)
            assert e.value.status == capi.BAD_ARG
        finally:
            B32.close()
    finally:
        B.close()
        m.close()
