// solve.hip — orchestration of the pressure solve (BuildPreconditioner +
// SolveConjugateGradient, src/sim/solver.inl:104-135, 289-379) on a handle's
// stream: preconditioner set-up, the PCG iterations queued in chunks with one
// readback of the solver scalars per chunk, reliable updates, and flexible PCG
// with the learned preconditioner.
#include "ariel_handle.h"

using namespace ariel;

namespace ariel {
namespace host {

int read_scalars(ariel_sim* h) {
    HIPCHK(h, hipMemcpyAsync(h->sc_host, h->sc, sizeof(SolverScalars), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

// PCG iterations are queued in chunks with one readback of the solver scalars
// after each. A readback idles the GPU for a few tens of microseconds; an
// iteration queued past convergence still costs the dispatch of each of its
// kernels (they return at once), about 80 us. The first chunk is two thirds of
// the previous solve's count; later ones extrapolate the residual's decay
// since the previous readback (or since the initial residual) to the stop
// threshold and queue half of that estimate: CG's decay speeds up as it
// proceeds, so the estimate runs high.
class ChunkPlan {
  public:
    explicit ChunkPlan(int last_iters) : first_(std::max(4, 2 * last_iters / 3)) {}
    int first() const { return first_; }
    int next(const SolverScalars& sc) {
        const double r0 = have_ ? r2_ : sc.rnorm2_0;
        const int dn = sc.it - (have_ ? it_ : 0);
        have_ = true;
        it_ = sc.it;
        r2_ = sc.rnorm2;
        if (dn <= 0 || !(sc.rnorm2 > 0.0) || !(r0 > sc.rnorm2) || !(sc.thr > 0.0)) return 4;
        const double need = std::log(sc.thr / sc.rnorm2);   // < 0 while not converged
        if (!(need < 0.0)) return 2;
        const double per_it = std::log(sc.rnorm2 / r0) / dn;  // < 0
        const double est = 0.5 * need / per_it;
        return est < 2.0 ? 2 : est > 64.0 ? 64 : (int)est;
    }

  private:
    int first_;
    bool have_ = false;
    int it_ = 0;
    double r2_ = 0.0;
};

// Active-set descriptor of a handle: counts from the last readback, or -1
// (read on the device) for batched launches; caps are the largest possible.
void set_active(ariel_sim* h, bool device_counts) {
    const int64_t TN = h->N / 8;
    const int64_t nt = TN * TN * TN;
    const int64_t per_colour = (TN & 1) == 0 ? nt / 2 : nt;
    h->act.segs = h->seg_list;
    h->act.nseg = device_counts ? -1 : h->sc_host->nseg;
    h->act.seg_cap = h->nc / 8;
    for (int c = 0; c < 3; ++c) {
        h->act.tiles[c] = h->tile_lists + c * nt;
        h->act.ntiles[c] = device_counts ? -1 : (int64_t)h->sc_host->ntile[c];
        h->act.tile_cap[c] = c == 2 ? nt : per_colour;
    }
}

// Preconditioner set-up on the current cell classes: diagonal, fluid segment
// bits, the active segment and tile lists (one host round trip for their
// counts), then the factorisation over the active tiles. Leaves the solver
// scalars initialised.
// Iterations between reliable updates during a solve. Multigrid solves take
// none: they converge in a few tens of iterations, and a residual replaced near
// convergence (by the true one, which has drifted from the recursive one by the
// FP32 rounding of A s at free-surface cells) stalls them; their iterate still
// accumulates in FP64 when refinement is on, and the end of the solve flushes it.
int refine_interval(const ariel_sim* h) {
    return h->d.precond == ARIEL_PRECOND_MG ? 0 : h->d.refine_every;
}

int prepare_solve(ariel_sim* h) {
    const Params p = h->params();
    hipStream_t s = h->stream;
    launch_precond_setup(s, p, h->A, h->L, h->diag, h->nup, h->pc);
    launch_fluid_segments(s, h->N, h->A, h->fl8);
    launch_ghost_levelset(s, h->N, h->A, h->L, h->Gt, h->sb);
    launch_solver_init(s, h->sc, h->d.max_iters, refine_interval(h),
                       h->d.precond == ARIEL_PRECOND_LEARNED ? 1 : 0);
    launch_active_sets(s, h->N, h->fl8, h->seg_flag, h->seg_scan, h->seg_list, h->tile_any,
                       h->tile_lists, h->scan_sums, h->sc);
    TRY(check_launch(h, "active sets"));
    TRY(read_scalars(h));
    set_active(h, false);
    if (h->d.precond == ARIEL_PRECOND_MG) launch_mg_build(s, h->N, h->A, h->mg);
    else launch_precond_build(s, p, h->A, h->diag, h->nup, h->pc, h->act, h->sc);
    return check_launch(h, "preconditioner build");
}

void precond_apply(ariel_sim* h, const Reduce& red) {
    if (h->d.precond == ARIEL_PRECOND_MG) {
        launch_mg_apply(h->stream, h->N, h->A, h->diag, h->mg, h->Rt, h->Zt, red, h->act);
        return;
    }
    launch_precond_apply(h->stream, h->params(), h->fl8, h->pc, h->Rt, h->Q, h->Zt, red, h->act);
}

void learned_precond(ariel_sim* h, const Params& p, const Reduce& red, bool mic_only);

// The launches of a solve after the preconditioner build, shared by the single
// tank path below and the batched tanks' recorded phases.
//
// Set-up: r = b - A p from the warm start into the tile-major Rt, the iterate
// into Pt (with reliable updates x64 = p and Pt = 0), then z = M^-1 r. The
// first iteration's operator pass takes s = z.
void enqueue_solve_setup(ariel_sim* h, bool learned) {
    const Params p = h->params();
    const Reduce red = h->reduce();
    hipStream_t s = h->stream;
    const Active& act = h->act;
    const int refine = h->d.refine_every > 0;
    launch_residual0(s, p, h->A, h->L, h->P, h->D, h->Rt, red, (double)h->d.tol, h->d.stop_rule,
                     h->P64, refine, h->Pt);
    if (refine) launch_fill_f32(s, h->Pt, h->nc, 0.0f);
    if (!learned) {
        precond_apply(h, red);
        return;
    }
    // the network reads r and writes z in the natural layout; both are zero
    // outside the active tiles
    launch_fill_f32(s, h->R, h->nc, 0.0f);
    launch_fill_f32(s, h->Z, h->nc, 0.0f);
    launch_fill_f32(s, h->R_old, h->nc, 0.0f);
    launch_diag_weight(s, h->N, h->A, h->diag, h->net.wd);
    learned_precond(h, p, red, false);
}

// One PCG iteration: s and y = A s (alpha), p += alpha s and r -= alpha y (the
// stop test), the reliable update when due, z = M^-1 r (beta). `it` is the
// iteration's number (1-based) for sampled profiling; prof = 0 disables it.
void enqueue_iteration(ariel_sim* h, int it, bool refine, bool learned, bool mic_only, bool prof) {
    const Params p = h->params();
    const Reduce red = h->reduce();
    hipStream_t s = h->stream;
    const Active& act = h->act;
    const bool sampled = prof && prof_sampled(it);
    // the tiled red-black MIC(0) carries the AXPY in its first pass over each
    // colour, except where the reliable update must come between the two; that
    // first pass also sums the SpMV's partials for alpha (no finalize launch)
    const bool fused = !learned && !refine && p.precond == ARIEL_PRECOND_MIC0_TILED;
    {
        Prof pr(h, ARIEL_KERNEL_SPMV, sampled);
        launch_spmv_dot(s, p, h->Gt, h->sb, h->fl8, h->Zt, h->St, h->Yt, red, act, fused);
    }
    if (fused) {
        const IterAxpy ax{h->Pt, h->St, h->Yt, h->nc};
        Prof pr(h, ARIEL_KERNEL_PRECOND, sampled);
        launch_precond_apply(s, p, h->fl8, h->pc, h->Rt, h->Q, h->Zt, red, act, &ax);
        return;
    }
    {
        Prof pr(h, ARIEL_KERNEL_AXPY, sampled);
        if (learned) launch_copy_list(s, h->Rt, h->R_old, act, h->sc);
        launch_axpy2(s, h->Pt, h->St, h->Rt, h->Yt, h->nc, red, act);
    }
    if (refine) launch_reliable_update(s, p, h->A, h->L, h->Pt, h->P64, h->D, h->Rt, red, act);
    {
        Prof pr(h, ARIEL_KERNEL_PRECOND, sampled);
        if (learned) learned_precond(h, p, red, mic_only);
        else precond_apply(h, red);
    }
}

// After the iterations: the pressure (natural) from the iterate; with reliable
// updates first x64 += p and the FP64 solution's true residual into sc->tmp64.
void enqueue_solve_finish(ariel_sim* h) {
    const Params p = h->params();
    const Reduce red = h->reduce();
    hipStream_t s = h->stream;
    const Active& act = h->act;
    const bool refine = h->d.refine_every > 0;
    if (refine) launch_true_residual_norm64(s, p, h->A, h->L, h->Pt, h->P64, h->D, red, act);
    launch_refine_finish(s, h->N, h->A, h->P, h->Pt, refine ? h->P64 : nullptr, act, h->sc);
}

// BuildPreconditioner + SolveConjugateGradient (solver.inl:104-135, 289-354) with
// the relative stop rule, FP64 reductions and residual replacement, or flexible
// PCG on the learned network. The iterations touch the active tiles only.
int stage_solve(ariel_sim* h) {
    const bool learned = h->d.precond == ARIEL_PRECOND_LEARNED;
    if (learned && !h->net_ready)
        return fail(h, ARIEL_ERR_STATE, "learned preconditioner: no weights (ariel_set_learned_weights)");
    const Params p = h->params();
    const Reduce red = h->reduce();
    hipStream_t s = h->stream;
    HIPCHK(h, hipEventRecord(h->solve_t0, s));
    TRY(prepare_solve(h));
    enqueue_solve_setup(h, learned);
    TRY(check_launch(h, "solve setup"));
    bool on_mic = false;
    int queued = 0;
    const int k = refine_interval(h);
    ChunkPlan plan(h->last_iters);
    int chunk = plan.first();
    for (;;) {
        const int n = std::min(chunk, h->d.max_iters - queued);
        for (int j = 0; j < n; ++j) {
            const int it = queued + j + 1;
            enqueue_iteration(h, it, k > 0 && it % k == 0, learned, on_mic, true);
        }
        TRY(check_launch(h, "pcg"));
        queued += n;
        TRY(read_scalars(h));
        on_mic = on_mic || h->sc_host->use_mic;
        if (h->sc_host->done || queued >= h->d.max_iters) break;
        chunk = plan.next(*h->sc_host);
    }
    enqueue_solve_finish(h);
    HIPCHK(h, hipEventRecord(h->solve_t1, s));
    h->last_iters = h->sc_host->it;
    h->stats.iterations = h->sc_host->it;
    h->stats.solve_status = h->sc_host->status;
    h->stats.rhs_norm = std::sqrt(h->sc_host->bnorm2);
    h->stats.residual_norm = std::sqrt(h->sc_host->rnorm2);
    h->stats.precond_fallback = learned && h->sc_host->use_mic ? h->sc_host->fallback_reason : 0;
    h->stats.fallback_iter = learned && h->sc_host->use_mic ? h->sc_host->fallback_it : 0;
    // true residual of the final pressure, for the report only
    launch_true_residual_norm(s, p, h->A, h->L, h->P, h->D, red, h->act);
    TRY(check_launch(h, "true residual"));
    TRY(read_scalars(h));
    const bool refine = h->d.refine_every > 0;
    h->stats.true_residual_norm = std::sqrt(h->sc_host->tmp);
    h->stats.true_residual_norm64 = refine ? std::sqrt(h->sc_host->tmp64) : h->stats.true_residual_norm;
    return ARIEL_OK;
}

// Flexible PCG (Notay 2000; Polak-Ribiere beta) with the learned network as
// preconditioner and the tiled MIC(0) as fallback: over every window of
// kFlexWindow iterations the residual must fall by 5% per iteration on average,
// and every z of the network must be a descent direction (z.r > 0); otherwise
// the solve switches to MIC(0) for good and restarts its search direction. The
// network and MIC(0) kernels are both queued each iteration; the one not in use
// returns at once. Once the host has seen the switch (at a readback), it
// queues the MIC(0) kernels only. The network reads r and writes z in the
// natural layout, converted from and to the tile-major vectors around it.
void learned_precond(ariel_sim* h, const Params& p, const Reduce& red, bool mic_only) {
    hipStream_t s = h->stream;
    if (!mic_only) {
        launch_tile_convert(s, h->N, h->Rt, h->R, h->act, h->sc, false, true);
        launch_learned_apply(s, h->net, h->N, h->R, h->A, h->diag, h->Z, h->sc, h->act.nseg);
        launch_tile_convert(s, h->N, h->Z, h->Zt, h->act, h->sc, true, true);
        launch_pr_dot(s, h->Zt, h->Rt, h->R_old, red, h->act, true);
    }
    launch_precond_apply(s, p, h->fl8, h->pc, h->Rt, h->Q, h->Zt, red, h->act);
    launch_pr_dot(s, h->Zt, h->Rt, h->R_old, red, h->act, false);
}

}  // namespace host
}  // namespace ariel
