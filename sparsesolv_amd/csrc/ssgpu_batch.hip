// Batched ICCG / COCG, IC-MRTR and SGS-MRTR: many independent systems (uneven sizes,
// any mix of matrices, several right-hand sides of one matrix) advanced
// together, one launch per phase of the iteration for all of them.
//
// The reference solves one system per call and keeps the Krylov loop on the
// host (MatSolverICCG.cpp:257-318).  Here the loop's scalars live on the
// device, one BatchState per system: the finalize kernels evaluate alpha,
// beta, the residual norm and the reference's convergence bookkeeping for
// every system at once, and a system that converges, diverges or reaches its
// iteration limit turns every later launch into a no-op for itself.  The
// launches of one iteration are built once into a hipGraph and replayed;
// the host looks at the batch only between replays (one small read-back),
// where finished systems leave and newly added ones join (continuous
// batching).  The sweeps of all systems run in one persistent launch
// (k_batch_sweep), so the dependency chains of the systems advance side by
// side instead of one after another.
//
// Every per-system operation is the one-system kernel's, on the same blocks
// and reduction tree, so each system's x, iteration count and residuals are
// bit-identical to its solo solve against the same factor
// (ssgpu_solve_factored).
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <string>

#include "ssgpu_host.h"

namespace ssgpu {

namespace {

using Clock = std::chrono::steady_clock;

enum SlotState { SLOT_FREE = 0, SLOT_NEW = 1, SLOT_RUN = 2 };

struct Slot {
    int state = SLOT_FREE;
    Matrix* m = nullptr;
    int64_t tag = 0;
    ssgpu_options opt{};
    double conv_cri = 0.0, normB = 0.0;
    int max_ite = 0;
    const void* user_b = nullptr;
    void* user_x = nullptr;
    bool own_bx = false;   // b and x live in private copies (ABMC-ordered matrix, or host arrays)
    bool host_io = false;  // b and x are host arrays: copied in at add, x copied out at finish
    std::unique_ptr<DevBuf> vecs, part;
    Clock::time_point t0;
};

struct Batch {
    Context* c = nullptr;
    int solver = SSGPU_ICCG;   // SSGPU_ICCG, SSGPU_MRTR_IC or SSGPU_MRTR_SGS
    int dtype = SSGPU_R64;
    int precision = SSGPU_FACTOR_F64;
    long spin_limit = kSweepSpinLimit;
    int sweep_grid = 0;       // fixed sweep grid (SSGPU_BATCH_GRID); 0: sized to the systems at each rebuild
    int grid_f = 128, grid_b = 128;   // persistent waves of the forward / backward batched sweep
    bool use_graph = true;
    std::vector<Slot> slots;
    std::vector<BatchSys> h_sys;
    std::vector<BatchState> h_state;
    DevBuf d_sys, d_state, d_items_mv, d_items_red, d_fin, d_ford, d_bord, d_flags;
    int n_mv = 0, n_red = 0, n_fin = 0, n_ford = 0, n_bord = 0, kmax_f = 1, kmax_b = 1;
    int cap = 0;
    bool dirty = false;
    bool has_new = false;
    int* h_flags = nullptr;   // pinned: [0] active count, [1] sweep error flag
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::deque<std::pair<int64_t, ssgpu_result>> done;
    int64_t iterations_run = 0;
    int capture_failures = 0;   // rebuilds whose graph capture fell back to eager launches

    int* d_active() { return d_flags.as<int>(); }
    int* d_err() { return d_flags.as<int>() + 1; }
    int* d_tickets() { return d_flags.as<int>() + 2; }

    ~Batch() { drop_graph(); if (h_flags) (void)hipHostFree(h_flags); }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

template <typename T> constexpr int parts() { return sizeof(T) / sizeof(double); }

// the backward sweep's structure: the IC factor's transpose, or for SGS-MRTR
// the upper triangle of S A S
const SfCsr& backward_sf(int solver, const Matrix& m) { return solver == SSGPU_MRTR_SGS ? m.sfUs : m.sfUic; }

// chunk levels of a sweep structure: the level of each chunk's first row
std::vector<int> chunk_levels(const LevelCsr& lv, int n) {
    const int nch = (n + kSfRows - 1) / kSfRows;
    std::vector<int> out(nch);
    int l = 0;
    for (int c = 0; c < nch; ++c) {
        const int p = c * kSfRows;
        while (l + 1 < (int)lv.levptr.size() - 1 && lv.levptr[l + 1] <= p) ++l;
        out[c] = l;
    }
    return out;
}

// ticket order of one sweep direction: every system's chunks in its own
// level order, systems interleaved by chunk level (counting sort on the
// level, stable in (slot, chunk))
std::vector<int2> merged_order(const std::vector<std::pair<int, const std::vector<int>*>>& sys_levels) {
    int maxl = 0;
    size_t total = 0;
    for (const auto& sl : sys_levels) {
        total += sl.second->size();
        if (!sl.second->empty()) maxl = std::max(maxl, sl.second->back());
    }
    std::vector<int> cnt(maxl + 2, 0);
    for (const auto& sl : sys_levels)
        for (int l : *sl.second) cnt[l + 1]++;
    for (int l = 0; l <= maxl; ++l) cnt[l + 1] += cnt[l];
    std::vector<int2> out(total);
    for (const auto& sl : sys_levels) {
        const std::vector<int>& lv = *sl.second;
        for (int ch = 0; ch < (int)lv.size(); ++ch) out[cnt[lv[ch]]++] = make_int2(sl.first, ch);
    }
    return out;
}

template <typename T> void upload_vec(Context* c, DevBuf& d, const std::vector<T>& h) {
    const size_t need = sizeof(T) * std::max<size_t>(h.size(), 1);
    if (d.bytes < need) d.alloc(c, need + need / 2);
    if (!h.empty()) SSGPU_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c->stream));
}

template <typename T> struct Runner {
    using F = typename std::conditional<std::is_same<T, double>::value, float, fc>::type;
    Batch& B;
    hipStream_t s;
    explicit Runner(Batch& b) : B(b), s(b.c->stream) {}

    const BatchSys* sys() { return B.d_sys.as<BatchSys>(); }
    BatchState* st() { return B.d_state.as<BatchState>(); }
    const int2* mv() { return B.d_items_mv.as<int2>(); }
    const int2* red() { return B.d_items_red.as<int2>(); }
    const int* fin() { return B.d_fin.as<int>(); }

    // launches go to the stream, or into the graph being built when rec is set
    GraphRec* rec = nullptr;
    void vec(int op, int want, int a, int b = -1, int c = -1, bool reset_tickets = false) {
        launch_batch_vec<T>(s, op, B.n_red, red(), sys(), st(), want, a, b, c, reset_tickets ? B.d_tickets() : nullptr,
                            rec);
    }
    void finalize(int phase, int want) {
        launch_batch_finalize<T>(s, phase, B.n_fin, fin(), sys(), st(), want, B.d_active(), B.d_tickets(), rec);
    }
    // one sweep of every system with status `want`: forward (ticket 0) or
    // backward (ticket 1; IC: the factor's transpose with D, SGS: upper(S A S))
    void sweep(bool fwd, int want, int in, int out) {
        int* tk = B.d_tickets() + (fwd ? 0 : 1);
        const int mode = fwd ? FWD_DIV_LAST : B.solver == SSGPU_MRTR_SGS ? BWD_DIV_FIRST : BWD_IC;
        const int kmax = fwd ? B.kmax_f : B.kmax_b, grid = fwd ? B.grid_f : B.grid_b;
        const int norder = fwd ? B.n_ford : B.n_bord;
        const int2* order = (fwd ? B.d_ford : B.d_bord).as<int2>();
        if (B.precision == SSGPU_FACTOR_F32)
            launch_batch_sweep<T, F>(s, mode, kmax, grid, norder, order, tk, B.d_err(), B.spin_limit, sys(), st(),
                                     want, in, out, 2, rec);
        else
            launch_batch_sweep<T, T>(s, mode, kmax, grid, norder, order, tk, B.d_err(), B.spin_limit, sys(), st(),
                                     want, in, out, 2, rec);
    }
    // the IC preconditioner: forward, then backward
    void sweeps(int want, int in, int mid, int out) {
        sweep(true, want, in, mid);
        sweep(false, want, mid, out);
    }
    void spmv(int want, int in, int out, int with_dot) {
        launch_batch_spmv<T>(s, B.n_mv, mv(), sys(), st(), want, in, out, with_dot, rec);
    }

    // the prologue of the systems added since the last step: ||r0|| and the
    // reference's early exits, then
    //   ICCG: z0 = M^-1 r0, p = z0, (r, z)   (MatSolverICCG.cpp:212-251)
    //   IC-MRTR: u = M^-1 r0, y = -r0, z = M^-1 y   (MatSolverMRTR_ICver.cpp:224-262)
    //   SGS-MRTR: ||b~|| first (the normB of the reference's checks), then
    //   rd = C^-1 tmp before ||rd||, y = -rd   (MatSolverMRTR_SGSver.cpp:127-168)
    void prologue() {
        if (B.solver == SSGPU_MRTR_SGS) {
            vec(BO_NORM2, BS_PENDING, BV_B);
            finalize(BP_NORMB, BS_PENDING);
            vec(BO_SENT, BS_PENDING, BV_R, BV_U);
            sweep(true, BS_PENDING, BV_T, BV_R);
            vec(BO_NORM2, BS_PENDING, BV_R);
            finalize(BP_PROLOGUE, BS_PENDING);
            vec(BO_NEG, BS_PRIMED, BV_R, BV_Y);
            finalize(BP_ACTIVATE, BS_PRIMED);
            return;
        }
        vec(BO_NORM2, BS_PENDING, BV_R);
        finalize(BP_PROLOGUE, BS_PENDING);
        if (B.solver == SSGPU_ICCG) {
            vec(BO_SENT, BS_PRIMED, BV_T, BV_P);
            sweeps(BS_PRIMED, BV_R, BV_T, BV_P);
            vec(BO_COPY, BS_PRIMED, BV_P, BV_Z);
            vec(BO_DOT, BS_PRIMED, BV_R, BV_Z);
            finalize(BP_RZ0, BS_PRIMED);
        } else {
            vec(BO_SENT, BS_PRIMED, BV_T, BV_U);
            sweeps(BS_PRIMED, BV_R, BV_T, BV_U);
            vec(BO_NEG, BS_PRIMED, BV_R, BV_Y);
            vec(BO_SENT, BS_PRIMED, BV_T, BV_Z, -1, true);
            sweeps(BS_PRIMED, BV_Y, BV_T, BV_Z);
            vec(BO_SENT, BS_PRIMED, BV_T, BV_W);
            finalize(BP_ACTIVATE, BS_PRIMED);
        }
    }
    // one iteration for every active system: MatSolverICCG.cpp:257-318,
    // MatSolverMRTR_ICver.cpp:270-360, or MatSolverMRTR_SGSver.cpp:182-262
    void iteration() {
        if (B.solver == SSGPU_MRTR_SGS) {
            iteration_sgs();
        } else if (B.solver == SSGPU_ICCG) {
            spmv(BS_ACTIVE, BV_P, BV_Q, 1);
            finalize(BP_ALPHA, BS_ACTIVE);
            vec(BO_CG_UPDATE, BS_ACTIVE, 0);
            finalize(BP_MONITOR, BS_ACTIVE);
            sweeps(BS_ACTIVE, BV_R, BV_T, BV_Z);
            vec(BO_DOT, BS_ACTIVE, BV_R, BV_Z);
            finalize(BP_BETA, BS_ACTIVE);
            vec(BO_PUPDATE, BS_ACTIVE, 0);
        } else {
            spmv(BS_ACTIVE, BV_U, BV_Q, 0);
            sweeps(BS_ACTIVE, BV_Q, BV_T, BV_W);
            vec(BO_DOT3, BS_ACTIVE, 0);
            finalize(BP_MRTR_COEF, BS_ACTIVE);
            vec(BO_MRTR_UPDATE, BS_ACTIVE, 0);
            finalize(BP_MONITOR, BS_ACTIVE);
            vec(BO_MRTR_UPDATE2, BS_ACTIVE, 0);
        }
    }
    // SGS-MRTR, as Solver::sgs_main: u = bc(rd); A rd = u + fr(rd - u) with
    // FP64 triangles (Eisenstat's splitting), fr(S A S u) with FP32 ones
    void iteration_sgs() {
        const int split = B.precision == SSGPU_FACTOR_F32 ? 0 : 1;
        sweep(false, BS_ACTIVE, BV_R, BV_U);
        vec(BO_SGS_SUB, BS_ACTIVE, split);
        if (!split) spmv(BS_ACTIVE, BV_U, BV_T, 0);
        sweep(true, BS_ACTIVE, BV_T, BV_Q);
        vec(BO_SGS_DOT3, BS_ACTIVE, split);
        finalize(BP_MRTR_COEF, BS_ACTIVE);
        vec(BO_SGS_UPDATE, BS_ACTIVE, 0);
        finalize(BP_MONITOR, BS_ACTIVE);
    }

    // Build the graph of one iteration node by node (hipGraphAddKernelNode in
    // launch order, each node depending on the one before).  No stream is put
    // in capture mode, so calls other threads make meanwhile (allocations,
    // frees, synchronizations) cannot meet a capturing stream.  A graph the
    // runtime refuses leaves the batch on eager launches of the same kernels
    // (the same bits), counted in capture_failures.
    void capture() {
        B.drop_graph();
        if (!B.use_graph || B.n_fin == 0) return;
        hipGraph_t g = nullptr;
        hipError_t e = hipGraphCreate(&g, 0);
        if (e == hipSuccess) {
            GraphRec r;
            r.graph = g;
            rec = &r;
            iteration();
            rec = nullptr;
            e = r.err;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&B.exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            if (g) (void)hipGraphDestroy(g);
            B.exec = nullptr;
            ++B.capture_failures;
            return;
        }
        B.graph = g;
    }

    // rebuild the device tables for the systems in the batch, run the
    // prologue of the new ones and capture the iteration
    void rebuild() {
        Context* c = B.c;
        const int ns = (int)B.slots.size();
        if (B.cap < ns) {   // grow the descriptor and state arrays (states of running systems kept)
            const int ncap = std::max(ns, 2 * B.cap);
            DevBuf nsys, nst;
            nsys.alloc(c, sizeof(BatchSys) * ncap);
            nst.alloc(c, sizeof(BatchState) * ncap);
            if (B.cap > 0)
                SSGPU_HIP(hipMemcpyAsync(nst.p, B.d_state.p, sizeof(BatchState) * B.cap, hipMemcpyDeviceToDevice, s));
            std::swap(B.d_sys.p, nsys.p);
            std::swap(B.d_sys.bytes, nsys.bytes);
            std::swap(B.d_state.p, nst.p);
            std::swap(B.d_state.bytes, nst.bytes);
            B.d_sys.ctx = B.d_state.ctx = c;
            B.cap = ncap;
            SSGPU_HIP(hipStreamSynchronize(s));
        }
        std::vector<int2> imv, ired;
        std::vector<int> fin;
        std::vector<std::pair<int, const std::vector<int>*>> fl, bl;
        std::vector<std::pair<Matrix*, std::pair<std::vector<int>, std::vector<int>>>> lev_cache;
        auto levels_of = [&](Matrix* m) -> std::pair<std::vector<int>, std::vector<int>>& {
            for (auto& e : lev_cache)
                if (e.first == m) return e.second;
            lev_cache.emplace_back(m, std::make_pair(chunk_levels(*m->sfL.lv, m->n),
                                                     chunk_levels(*backward_sf(B.solver, *m).lv, m->n)));
            return lev_cache.back().second;
        };
        lev_cache.reserve(ns);
        int kf = 1, kb = 1, active = 0;
        for (int i = 0; i < ns; ++i) {
            Slot& q = B.slots[i];
            if (q.state == SLOT_FREE) continue;
            const BatchSys& d = B.h_sys[i];
            for (int b = 0; b < d.nblk; ++b) imv.push_back(make_int2(i, b));
            for (int b = 0; b < d.rblk; ++b) ired.push_back(make_int2(i, b));
            fin.push_back(i);
            auto& lv = levels_of(q.m);
            fl.emplace_back(i, &lv.first);
            bl.emplace_back(i, &lv.second);
            kf = std::max(kf, q.m->sfL.kmax);
            kb = std::max(kb, backward_sf(B.solver, *q.m).kmax);
            ++active;
        }
        const std::vector<int2> ford = merged_order(fl), bord = merged_order(bl);
        // Sweep grid: a system keeps about 6 levels of chunks in flight (the
        // one-system sweep's tuned 128-160 waves at n=95, 24.5 chunks per
        // level); waves beyond what the systems' level fronts can use only
        // add polling traffic, which slows every level.
        auto grid_for = [&](const std::vector<std::pair<int, const std::vector<int>*>>& sl, bool fwd) {
            if (B.sweep_grid > 0) return B.sweep_grid;
            double want = 0.0;
            for (const auto& e : sl) {
                const Matrix& mm = *B.slots[e.first].m;
                const int nlev = std::max(1, (fwd ? mm.sfL.lv : backward_sf(B.solver, mm).lv)->nlev);
                want += std::ceil(6.0 * (double)e.second->size() / nlev);
            }
            return std::max(64, std::min(1024, (int)want));
        };
        B.grid_f = grid_for(fl, true);
        B.grid_b = grid_for(bl, false);
        upload_vec(c, B.d_items_mv, imv);
        upload_vec(c, B.d_items_red, ired);
        upload_vec(c, B.d_fin, fin);
        upload_vec(c, B.d_ford, ford);
        upload_vec(c, B.d_bord, bord);
        SSGPU_HIP(hipMemcpyAsync(B.d_sys.p, B.h_sys.data(), sizeof(BatchSys) * ns, hipMemcpyHostToDevice, s));
        for (int i = 0; i < ns; ++i)
            if (B.slots[i].state == SLOT_NEW)
                SSGPU_HIP(hipMemcpyAsync(B.d_state.as<BatchState>() + i, &B.h_state[i], sizeof(BatchState),
                                         hipMemcpyHostToDevice, s));
        B.h_flags[0] = active;
        B.h_flags[1] = 0;
        SSGPU_HIP(hipMemcpyAsync(B.d_active(), B.h_flags, sizeof(int), hipMemcpyHostToDevice, s));
        B.n_mv = (int)imv.size();
        B.n_red = (int)ired.size();
        B.n_fin = (int)fin.size();
        B.n_ford = (int)ford.size();
        B.n_bord = (int)bord.size();
        B.kmax_f = kf;
        B.kmax_b = kb;
        SSGPU_HIP(hipStreamSynchronize(s));   // the host tables above are temporaries
        if (B.has_new) prologue();
        for (auto& q : B.slots)
            if (q.state == SLOT_NEW) q.state = SLOT_RUN;
        B.has_new = false;
        capture();
        B.dirty = false;
    }

    // a system that left the loop: best-iterate restore, x back to the
    // caller's order, true residual, result
    void finish(int i, const BatchState& q) {
        Slot& sl = B.slots[i];
        Matrix& m = *sl.m;
        const BatchSys& d = B.h_sys[i];
        T* x = static_cast<T*>(d.vec[BV_X]);
        const T* b = static_cast<const T*>(d.vec[BV_B]);
        const bool conv = q.status == BS_CONVERGED || q.status == BS_EARLY;
        // on failure with save-best, the best iterate (the current one when the
        // last check had just saved it): MatSolverICCG.cpp:315-317
        const bool restore = !conv && sl.opt.is_save_best && !q.copy_best;
        if (B.solver == SSGPU_MRTR_SGS) {
            // solve_factored's SGS epilogue: x = s o xi (the scaled iterate,
            // untouched when the prologue returned early); the true residual
            // is taken against the caller's b
            launch_copy<T>(s, m.n, static_cast<const T*>(d.vec[restore ? BV_BEST : BV_XI]), x);
            launch_mul_scale<T>(s, m.n, m.sc.as<T>(), x);
            b = static_cast<const T*>(d.vec[BV_Z]);
        } else {
            // IC-MRTR keeps its iterate apart and returns it at the end, except
            // from the prologue's early exit (MatSolverMRTR_ICver.cpp:230-240, 362-367)
            if (B.solver == SSGPU_MRTR_IC && q.status != BS_EARLY)
                launch_copy<T>(s, m.n, static_cast<const T*>(d.vec[BV_XI]), x);
            if (restore) launch_copy<T>(s, m.n, static_cast<const T*>(d.vec[BV_BEST]), x);
        }
        int refinements = 0, refine_its = 0;
        const double true_res = batch_finish<T>(B.c, m, B.solver, sl.opt, conv, sl.conv_cri, sl.max_ite, sl.normB, b,
                                                x, refinements, refine_its);
        if (sl.own_bx) {   // x back in the caller's order, to the caller's device or host array
            T* dst = sl.host_io ? static_cast<T*>(d.vec[BV_T]) : static_cast<T*>(sl.user_x);
            if (m.ordered) launch_gather<T, T>(s, m.n, m.iperm.as<int>(), x, dst);
            else if (!sl.host_io) launch_copy<T>(s, m.n, x, dst);
            if (sl.host_io)
                SSGPU_HIP(hipMemcpyAsync(sl.user_x, m.ordered ? dst : x, sizeof(T) * m.n, hipMemcpyDeviceToHost, s));
        }
        SSGPU_HIP(hipStreamSynchronize(s));
        ssgpu_result r{};
        r.converged = conv ? 1 : 0;
        r.status = conv ? OK : NOT_CONVERGED;
        r.iterations = q.iterations;
        r.log_len = 0;
        r.final_residual = q.final_res;
        r.accel_used = m.factor_accel;
        r.pivot_fixes = m.factor_pivot_fixes;
        r.setup_ms = 0.0;
        r.iterate_ms = std::chrono::duration<double, std::milli>(Clock::now() - sl.t0).count();
        r.true_residual = true_res;
        r.refinements = refinements;
        r.refine_iterations = refine_its;
        B.done.emplace_back(sl.tag, r);
        sl.state = SLOT_FREE;
        sl.m = nullptr;
        B.dirty = true;
    }

    // run `iters` iterations (fewer if every system finishes), then finish
    // the systems that left the loop
    void step(int iters) {
        if (B.dirty) rebuild();
        if (B.n_fin == 0) return;
        for (int k = 0; k < iters; ++k) {
            if (B.exec) SSGPU_HIP(hipGraphLaunch(B.exec, s));
            else iteration();
        }
        B.iterations_run += iters;
        SSGPU_HIP(hipMemcpyAsync(B.h_flags, B.d_flags.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        SSGPU_HIP(hipMemcpyAsync(B.h_state.data(), B.d_state.p, sizeof(BatchState) * B.slots.size(),
                                 hipMemcpyDeviceToHost, s));
        SSGPU_HIP(hipStreamSynchronize(s));
        if (B.h_flags[1] != 0) {
            // a sweep gave up: every system of the batch is dropped, the flag cleared
            launch_fill<int>(s, 1, B.d_err(), 0);
            for (auto& q : B.slots) {
                if (q.state == SLOT_FREE) continue;
                ssgpu_result r{};
                r.status = HIP_ERROR;
                B.done.emplace_back(q.tag, r);
                q.state = SLOT_FREE;
                q.m = nullptr;
            }
            B.dirty = true;
            SSGPU_HIP(hipStreamSynchronize(s));
            throw Error{HIP_ERROR, "a batched sweep exceeded its spin limit (a dependency was never produced)"};
        }
        for (int i = 0; i < (int)B.slots.size(); ++i) {
            if (B.slots[i].state != SLOT_RUN) continue;
            const int stt = B.h_state[i].status;
            if (stt == BS_CONVERGED || stt == BS_STOPPED || stt == BS_EARLY) finish(i, B.h_state[i]);
        }
    }

    // queue one system; it joins the loop at the next step
    void add(Matrix& m, const ssgpu_options& o, double conv_cri, int max_ite, double normB, const void* ub, void* ux,
             int64_t tag, bool host_io) {
        Context* c = B.c;
        int i = 0;
        while (i < (int)B.slots.size() && B.slots[i].state != SLOT_FREE) ++i;
        if (i == (int)B.slots.size()) {
            B.slots.emplace_back();
            B.h_sys.emplace_back();
            B.h_state.emplace_back();
        }
        Slot& sl = B.slots[i];
        const int n = m.n;
        const size_t vbytes = sizeof(T) * (size_t)n;
        if (!sl.vecs) sl.vecs.reset(new DevBuf());
        if (!sl.part) sl.part.reset(new DevBuf());
        const int nvec = B.solver == SSGPU_ICCG ? kBatchVecsIccg : kBatchVecs;
        if (sl.vecs->bytes < vbytes * nvec) sl.vecs->alloc(c, vbytes * nvec);
        const int rblk = reduce_blocks(n);
        // partials: the SpMV's dot writes 1-2 doubles per SpMV block, the
        // three-dot phases of the MRTR solvers 3-6 per reduction block
        const size_t pbytes = sizeof(double) * std::max<size_t>(2 * (size_t)m.nblk, 6 * (size_t)rblk);
        if (sl.part->bytes < pbytes) sl.part->alloc(c, pbytes);
        BatchSys d{};
        d.n = n;
        d.nblk = m.nblk;
        d.rblk = rblk;
        d.blk = m.blk.as<int>();
        d.rp = m.rp.as<int>();
        d.ci = m.ci.as<int>();
        const bool sgs = B.solver == SSGPU_MRTR_SGS;
        d.val = sgs ? m.scaledA.p : m.val.p;
        const bool f32 = B.precision == SSGPU_FACTOR_F32;
        d.f_srow = m.sfL.srow;
        d.f_soff = m.sfL.slice_off.as<int>();
        d.f_rlen = m.sfL.rlen.as<int>();
        d.f_ecol = m.sfL.ecol.as<int>();
        d.f_eval = f32 ? m.sfL.eval32.p : m.sfL.eval.p;
        d.f_dval = f32 ? m.sfL.dval32.p : m.sfL.dval.p;
        const SfCsr& bs = backward_sf(B.solver, m);
        d.b_srow = bs.srow;
        d.b_soff = bs.slice_off.as<int>();
        d.b_rlen = bs.rlen.as<int>();
        d.b_ecol = bs.ecol.as<int>();
        d.b_eval = f32 ? bs.eval32.p : bs.eval.p;
        d.b_dval = sgs ? (f32 ? bs.dval32.p : bs.dval.p) : nullptr;
        d.diagD = sgs ? nullptr : m.diagD.p;
        char* base = static_cast<char*>(sl.vecs->p);
        for (int k = 0; k < kBatchVecs; ++k) d.vec[k] = k < nvec ? base + vbytes * k : nullptr;
        d.part = sl.part->as<double>();
        sl.own_bx = m.ordered || host_io;
        sl.host_io = host_io;
        const int bv = sgs ? BV_Z : BV_B;   // the caller's b, in the matrix's order
        if (sl.own_bx) {
            // private b and x in the matrix's order (PermutedIO in the one-system
            // path); host arrays come in through two scratch vectors
            const T* sb = static_cast<const T*>(ub);
            const T* sx = static_cast<const T*>(ux);
            if (host_io) {
                T* tb = static_cast<T*>(m.ordered ? d.vec[BV_T] : d.vec[bv]);
                T* tx = static_cast<T*>(m.ordered ? d.vec[BV_Q] : d.vec[BV_X]);
                SSGPU_HIP(hipMemcpyAsync(tb, ub, vbytes, hipMemcpyHostToDevice, s));
                SSGPU_HIP(hipMemcpyAsync(tx, ux, vbytes, hipMemcpyHostToDevice, s));
                sb = tb;
                sx = tx;
            }
            if (m.ordered) {
                launch_gather<T, T>(s, n, m.perm.as<int>(), sb, static_cast<T*>(d.vec[bv]));
                launch_gather<T, T>(s, n, m.perm.as<int>(), sx, static_cast<T*>(d.vec[BV_X]));
            }
        } else {
            d.vec[bv] = const_cast<void*>(ub);
            d.vec[BV_X] = ux;
        }
        // the first lines of iccg_main / mrtr_main / sgs_main, on this system alone
        T* x = static_cast<T*>(d.vec[BV_X]);
        const T* b = static_cast<const T*>(d.vec[BV_B]);
        T* r = static_cast<T*>(d.vec[BV_R]);
        T* q = static_cast<T*>(d.vec[BV_Q]);
        if (o.is_save_best) launch_fill<T>(s, n, static_cast<T*>(d.vec[BV_BEST]), T{});
        if (sgs) {
            // solve_factored's scaling: b~ = s o b, xi = x0 / s (MatSolverMRTR_SGSver.cpp:42-69);
            // then tmp = b~, less S A S xi only when xi is the zero start
            // (MatSolverMRTR_SGSver.cpp:113-117); ||b~|| and rd = C^-1 tmp come
            // in the prologue
            T* bt = static_cast<T*>(d.vec[BV_B]);
            T* xi = static_cast<T*>(d.vec[BV_XI]);
            T* t = static_cast<T*>(d.vec[BV_T]);
            launch_copy<T>(s, n, static_cast<const T*>(d.vec[BV_Z]), bt);
            launch_mul_scale<T>(s, n, m.sc.as<T>(), bt);
            launch_fill<T>(s, n, static_cast<T*>(d.vec[BV_P]), T{});
            if (o.is_init) {
                launch_fill<T>(s, n, xi, T{});
                launch_spmv<T>(s, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), m.scaledA.as<T>(), xi, q,
                               nullptr, nullptr, m.spmv_stage_kb);
                launch_residual<T>(s, n, bt, q, t);
            } else {
                launch_copy<T>(s, n, x, xi);
                launch_div_scale<T>(s, n, m.sc.as<T>(), xi);
                launch_residual<T>(s, n, bt, nullptr, t);
            }
        } else if (B.solver == SSGPU_ICCG) {
            if (o.is_init) launch_fill<T>(s, n, x, T{});
            if (!o.is_init) {
                launch_spmv<T>(s, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), m.val.as<T>(), x, q,
                               nullptr, nullptr, m.spmv_stage_kb);
                launch_residual<T>(s, n, b, q, r);
            } else {
                launch_residual<T>(s, n, b, nullptr, r);
            }
        } else {
            T* xi = static_cast<T*>(d.vec[BV_XI]);
            launch_fill<T>(s, n, static_cast<T*>(d.vec[BV_P]), T{});
            if (o.is_init) {
                launch_fill<T>(s, n, x, T{});
                launch_fill<T>(s, n, xi, T{});
                // MatSolverMRTR_ICver.cpp:213-216 subtracts A x only here (x is zero)
                launch_spmv<T>(s, m.nblk, m.blk.as<int>(), m.rp.as<int>(), m.ci.as<int>(), m.val.as<T>(), x, q,
                               nullptr, nullptr, m.spmv_stage_kb);
                launch_residual<T>(s, n, b, q, r);
            } else {
                launch_copy<T>(s, n, x, xi);
                launch_residual<T>(s, n, b, nullptr, r);
            }
        }
        BatchState st{};
        st.normB = normB;
        st.normalizer = 1.0;
        st.best = 1.0e+6;
        st.conv_cri = conv_cri;
        st.bad_div_val = o.bad_div_val;
        st.normalize_const = o.conv_normalize_const;
        st.max_ite = max_ite;
        st.status = BS_PENDING;
        st.normalize_type = o.conv_normalize_type;
        st.save_best = o.is_save_best;
        st.diverge_type = o.diverge_judge_type;
        st.bad_div_thres = o.bad_div_count_thres;
        st.mrtr_fix = o.mrtr_direction_fix;
        st.zeta[0] = st.zeta_old[0] = st.nu[0] = 1.0;   // zeta = zeta_old = nu = 1, eta = 0
        B.h_sys[i] = d;
        B.h_state[i] = st;
        sl.state = SLOT_NEW;
        sl.m = &m;
        sl.tag = tag;
        sl.opt = o;
        sl.conv_cri = conv_cri;
        sl.max_ite = max_ite;
        sl.normB = normB;
        sl.user_b = ub;
        sl.user_x = ux;
        sl.t0 = Clock::now();
        B.dirty = true;
        B.has_new = true;
    }
};

}  // namespace

}  // namespace ssgpu

using namespace ssgpu;

struct ssgpu_batch_s {
    Batch b;
};

namespace {

void check_batch_matrix(const Batch& B, const Matrix& m) {
    if (m.ctx->device != B.c->device) throw Error{BAD_ARG, "matrix and batch are on different devices"};
    if (m.dtype != B.dtype) throw Error{BAD_ARG, "matrix value type differs from the batch's"};
    if (B.solver == SSGPU_MRTR_SGS) {
        if (m.prep_kind != 2) throw Error{BAD_ARG, "no SGS preparation: call ssgpu_factorize with SSGPU_MRTR_SGS"};
        if (B.precision == SSGPU_FACTOR_F32 && m.prep_precision != SSGPU_FACTOR_F32)
            throw Error{BAD_ARG, "the SGS preparation was built without its FP32 copy"};
    } else {
        if (m.prep_kind != 1) throw Error{BAD_ARG, "no IC factor: call ssgpu_factorize first"};
        if (B.precision == SSGPU_FACTOR_F32 && m.prep_precision != SSGPU_FACTOR_F32)
            throw Error{BAD_ARG, "the factor was built without its FP32 copy"};
    }
    if (m.sfL.kmax <= 0 || backward_sf(B.solver, m).kmax <= 0)
        throw Error{UNSUPPORTED, "batched solves need rows of at most 64 off-diagonal entries"};
}

template <typename F> int with_runner(Batch& B, F&& f) {
    if (B.dtype == SSGPU_R64) {
        Runner<double> r(B);
        f(r);
    } else {
        Runner<zc> r(B);
        f(r);
    }
    return SSGPU_OK;
}

}  // namespace

extern "C" {

int ssgpu_batch_create(ssgpu_context ctx, int solver, int dtype, int factor_precision, ssgpu_batch* out) {
    if (!ctx || !out || (dtype != SSGPU_R64 && dtype != SSGPU_C128)) return SSGPU_BAD_ARG;
    if (factor_precision != SSGPU_FACTOR_F64 && factor_precision != SSGPU_FACTOR_F32) return SSGPU_BAD_ARG;
    if (solver != SSGPU_ICCG && solver != SSGPU_MRTR_IC && solver != SSGPU_MRTR_SGS) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    std::unique_ptr<ssgpu_batch_s> h(new ssgpu_batch_s());
    Batch& B = h->b;
    B.c = &ctx->c;
    B.solver = solver;
    B.dtype = dtype;
    B.precision = factor_precision;
    if (const char* e = std::getenv("SSGPU_SPIN_LIMIT")) B.spin_limit = std::max(0L, std::atol(e));
    if (const char* e = std::getenv("SSGPU_BATCH_GRAPH")) B.use_graph = std::atoi(e) != 0;
    if (const char* e = std::getenv("SSGPU_BATCH_GRID")) B.sweep_grid = std::max(0, std::atoi(e));
    SSGPU_HIP(hipSetDevice(B.c->device));
    std::lock_guard<std::mutex> lk(B.c->mu);
    B.d_flags.alloc(B.c, 64);
    launch_fill<int>(B.c->stream, 16, B.d_flags.as<int>(), 0);
    SSGPU_HIP(hipHostMalloc((void**)&B.h_flags, 64, hipHostMallocDefault));
    SSGPU_HIP(hipStreamSynchronize(B.c->stream));
    *out = h.release();
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_batch_add(ssgpu_batch bh, ssgpu_matrix mh, const ssgpu_options* opt, double conv_cri, int max_ite,
                    double normB, const void* d_b, void* d_x, int64_t tag) {
    if (!bh || !mh || !opt || !d_b || !d_x) return SSGPU_BAD_ARG;
    if (opt->is_diag_scale && bh->b.solver != SSGPU_MRTR_SGS) return SSGPU_UNSUPPORTED;
    SSGPU_GUARD_BEGIN
    Batch& B = bh->b;
    Matrix& m = mh->m;
    check_batch_matrix(B, m);
    std::lock_guard<std::mutex> lk(B.c->mu);
    SSGPU_HIP(hipSetDevice(B.c->device));
    return with_runner(B, [&](auto& r) { r.add(m, *opt, conv_cri, max_ite, normB, d_b, d_x, tag, false); });
    SSGPU_GUARD_END
}

int ssgpu_batch_add_host(ssgpu_batch bh, ssgpu_matrix mh, const ssgpu_options* opt, double conv_cri, int max_ite,
                         double normB, const void* b, void* x, int64_t tag) {
    if (!bh || !mh || !opt || !b || !x) return SSGPU_BAD_ARG;
    if (opt->is_diag_scale && bh->b.solver != SSGPU_MRTR_SGS) return SSGPU_UNSUPPORTED;
    SSGPU_GUARD_BEGIN
    Batch& B = bh->b;
    Matrix& m = mh->m;
    check_batch_matrix(B, m);
    std::lock_guard<std::mutex> lk(B.c->mu);
    SSGPU_HIP(hipSetDevice(B.c->device));
    return with_runner(B, [&](auto& r) { r.add(m, *opt, conv_cri, max_ite, normB, b, x, tag, true); });
    SSGPU_GUARD_END
}

int ssgpu_batch_step(ssgpu_batch bh, int iterations, int* running, int* finished) {
    if (!bh || iterations < 0) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Batch& B = bh->b;
    std::lock_guard<std::mutex> lk(B.c->mu);
    SSGPU_HIP(hipSetDevice(B.c->device));
    with_runner(B, [&](auto& r) { r.step(iterations); });
    if (running) {
        int k = 0;
        for (const auto& q : B.slots) k += q.state != SLOT_FREE;
        *running = k;
    }
    if (finished) *finished = (int)B.done.size();
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_batch_poll(ssgpu_batch bh, int* got, int64_t* tag, ssgpu_result* res) {
    if (!bh || !got) return SSGPU_BAD_ARG;
    Batch& B = bh->b;
    std::lock_guard<std::mutex> lk(B.c->mu);
    *got = 0;
    if (B.done.empty()) return SSGPU_OK;
    if (tag) *tag = B.done.front().first;
    if (res) *res = B.done.front().second;
    B.done.pop_front();
    *got = 1;
    return SSGPU_OK;
}

int ssgpu_batch_info(ssgpu_batch bh, int* systems, int64_t* iterations_run, int* graph_fallbacks) {
    if (!bh) return SSGPU_BAD_ARG;
    Batch& B = bh->b;
    std::lock_guard<std::mutex> lk(B.c->mu);
    if (systems) {
        int k = 0;
        for (const auto& q : B.slots) k += q.state != SLOT_FREE;
        *systems = k;
    }
    if (iterations_run) *iterations_run = B.iterations_run;
    if (graph_fallbacks) *graph_fallbacks = B.capture_failures;
    return SSGPU_OK;
}

int ssgpu_batch_trim(ssgpu_batch bh) {
    if (!bh) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    Batch& B = bh->b;
    std::lock_guard<std::mutex> lk(B.c->mu);
    SSGPU_HIP(hipSetDevice(B.c->device));
    SSGPU_HIP(hipStreamSynchronize(B.c->stream));   // no queued launch may still use a slot's vectors
    bool freed = false;
    for (auto& q : B.slots) {
        if (q.state != SLOT_FREE) continue;
        freed |= (q.vecs && q.vecs->bytes) || (q.part && q.part->bytes);
        q.vecs.reset();
        q.part.reset();
    }
    if (freed) B.dirty = true;   // the device tables and the graph are rebuilt before the next step
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_batch_destroy(ssgpu_batch bh) {
    if (!bh) return SSGPU_BAD_ARG;
    SSGPU_GUARD_BEGIN
    {
        std::lock_guard<std::mutex> lk(bh->b.c->mu);
        (void)hipSetDevice(bh->b.c->device);
        (void)hipStreamSynchronize(bh->b.c->stream);
    }
    delete bh;
    return SSGPU_OK;
    SSGPU_GUARD_END
}

int ssgpu_solve_batch(ssgpu_context ctx, int solver, int nsys, const ssgpu_matrix* mats, const ssgpu_options* opt,
                      double conv_cri, int max_ite, const double* normB, const void* const* d_b, void* const* d_x,
                      ssgpu_result* res) {
    if (!ctx || nsys < 1 || !mats || !opt || !normB || !d_b || !d_x || !res) return SSGPU_BAD_ARG;
    if (opt->is_diag_scale && solver != SSGPU_MRTR_SGS) return SSGPU_UNSUPPORTED;
    for (int i = 0; i < nsys; ++i)
        if (!mats[i] || !d_b[i] || !d_x[i]) return SSGPU_BAD_ARG;
    const int dtype = mats[0]->m.dtype;
    ssgpu_batch bh = nullptr;
    int st = ssgpu_batch_create(ctx, solver, dtype, opt->factor_precision, &bh);
    if (st != SSGPU_OK) return st;
    for (int i = 0; i < nsys && st == SSGPU_OK; ++i)
        st = ssgpu_batch_add(bh, mats[i], opt, conv_cri, max_ite, normB[i], d_b[i], d_x[i], i);
    int running = nsys;
    int chunk = 8;
    if (const char* e = std::getenv("SSGPU_BATCH_CHECK")) chunk = std::max(1, std::atoi(e));
    while (st == SSGPU_OK && running > 0) st = ssgpu_batch_step(bh, chunk, &running, nullptr);
    int all_conv = 1;
    if (st == SSGPU_OK) {
        int got = 1;
        int64_t tag = 0;
        ssgpu_result r{};
        while (ssgpu_batch_poll(bh, &got, &tag, &r) == SSGPU_OK && got) {
            res[tag] = r;
            all_conv &= r.converged;
        }
    }
    const std::string err = ssgpu_last_error();
    ssgpu_batch_destroy(bh);
    if (st != SSGPU_OK) {
        set_last_error(err);
        return st;
    }
    return all_conv ? SSGPU_OK : SSGPU_NOT_CONVERGED;
}

}  // extern "C"
