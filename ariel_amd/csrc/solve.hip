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
int prepare_solve(ariel_sim* h) {
    const Params p = h->params();
    hipStream_t s = h->stream;
    launch_precond_setup(s, p, h->A, h->L, h->diag, h->nup, h->pc);
    launch_fluid_segments(s, h->N, h->A, h->fl8);
    launch_solver_init(s, h->sc, h->d.max_iters, h->d.refine_every,
                       h->d.precond == ARIEL_PRECOND_LEARNED ? 1 : 0);
    launch_active_sets(s, h->N, h->fl8, h->seg_flag, h->seg_scan, h->seg_list, h->tile_any,
                       h->tile_lists, h->scan_sums, h->sc);
    TRY(check_launch(h, "active sets"));
    TRY(read_scalars(h));
    set_active(h, false);
    launch_precond_build(s, p, h->A, h->diag, h->nup, h->pc, h->act, h->sc);
    return check_launch(h, "preconditioner build");
}

// BuildPreconditioner + SolveConjugateGradient (solver.inl:104-135, 289-354) with
// the relative stop rule, FP64 reductions and residual replacement. The
// iterations touch the active segments and tiles only. The full-grid set-up
// below leaves P and R zero off-FLUID and Z zero everywhere; S and Z are then
// written on every active segment before they are read, and values off the
// active sets are never read as FLUID values.
int stage_solve_learned(ariel_sim* h);

int stage_solve(ariel_sim* h) {
    if (h->d.precond == ARIEL_PRECOND_LEARNED) return stage_solve_learned(h);
    const Params p = h->params();
    const Reduce red = h->reduce();
    hipStream_t s = h->stream;
    HIPCHK(h, hipEventRecord(h->solve_t0, s));
    TRY(prepare_solve(h));
    const Active& act = h->act;
    const int refine = h->d.refine_every > 0;
    launch_residual0(s, p, h->A, h->L, h->P, h->D, h->R, red, (double)h->d.tol, h->d.stop_rule,
                     h->P64, refine);
    if (refine) launch_fill_f32(s, h->P, h->nc, 0.0f);
    launch_fill_f32(s, h->Z, h->nc, 0.0f);
    launch_precond_apply(s, p, h->fl8, h->pc, h->R, h->Q, h->Z, red, act);
    launch_update_s(s, h->Z, h->S, red, 1, act);
    TRY(check_launch(h, "solve setup"));
    int queued = 0;
    const int k = h->d.refine_every;
    ChunkPlan plan(h->last_iters);
    int chunk = plan.first();
    for (;;) {
        const int n = std::min(chunk, h->d.max_iters - queued);
        for (int j = 0; j < n; ++j) {
            const int it = queued + j + 1;
            {
                Prof prof(h, ARIEL_KERNEL_SPMV, prof_sampled(it));
                launch_spmv_dot(s, p, h->A, h->L, h->S, h->Z, red, act);
            }
            {
                Prof prof(h, ARIEL_KERNEL_AXPY, prof_sampled(it));
                launch_axpy2(s, h->P, h->S, h->R, h->Z, red, act);
            }
            if (k > 0 && it % k == 0)
                launch_reliable_update(s, p, h->A, h->L, h->P, h->P64, h->D, h->R, red, act);
            {
                Prof prof(h, ARIEL_KERNEL_PRECOND, prof_sampled(it));
                launch_precond_apply(s, p, h->fl8, h->pc, h->R, h->Q, h->Z, red, act);
            }
            launch_update_s(s, h->Z, h->S, red, 0, act);
        }
        TRY(check_launch(h, "pcg"));
        queued += n;
        TRY(read_scalars(h));
        if (h->sc_host->done || queued >= h->d.max_iters) break;
        chunk = plan.next(*h->sc_host);
    }
    if (refine) {
        launch_true_residual_norm64(s, p, h->A, h->L, h->P, h->P64, h->D, red, act);
        launch_refine_finish(s, h->A, h->P, h->P64, act, h->sc);
    }
    HIPCHK(h, hipEventRecord(h->solve_t1, s));
    h->last_iters = h->sc_host->it;
    h->stats.iterations = h->sc_host->it;
    h->stats.solve_status = h->sc_host->status;
    h->stats.rhs_norm = std::sqrt(h->sc_host->bnorm2);
    h->stats.residual_norm = std::sqrt(h->sc_host->rnorm2);
    // true residual of the final pressure, for the report only
    launch_true_residual_norm(s, p, h->A, h->L, h->P, h->D, red, act);
    TRY(check_launch(h, "true residual"));
    TRY(read_scalars(h));
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
// queues the MIC(0) kernels only.
void learned_precond(ariel_sim* h, const Params& p, const Reduce& red, bool mic_only) {
    hipStream_t s = h->stream;
    if (!mic_only) {
        launch_learned_apply(s, h->net, h->N, h->R, h->A, h->diag, h->Z, h->sc, h->act.nseg);
        launch_pr_dot(s, h->Z, h->R, h->R_old, red, h->act, true);
    }
    launch_precond_apply(s, p, h->fl8, h->pc, h->R, h->Q, h->Z, red, h->act);
    launch_pr_dot(s, h->Z, h->R, h->R_old, red, h->act, false);
}

int stage_solve_learned(ariel_sim* h) {
    if (!h->net_ready)
        return fail(h, ARIEL_ERR_STATE, "learned preconditioner: no weights (ariel_set_learned_weights)");
    const Params p = h->params();
    const Reduce red = h->reduce();
    hipStream_t s = h->stream;
    HIPCHK(h, hipEventRecord(h->solve_t0, s));
    TRY(prepare_solve(h));
    const Active& act = h->act;
    const int refine = h->d.refine_every > 0;
    launch_residual0(s, p, h->A, h->L, h->P, h->D, h->R, red, (double)h->d.tol, h->d.stop_rule,
                     h->P64, refine);
    if (refine) launch_fill_f32(s, h->P, h->nc, 0.0f);
    launch_fill_f32(s, h->Z, h->nc, 0.0f);
    launch_fill_f32(s, h->R_old, h->nc, 0.0f);
    launch_diag_weight(s, h->N, h->A, h->diag, h->net.wd);
    learned_precond(h, p, red, false);
    launch_update_s(s, h->Z, h->S, red, 1, act);
    TRY(check_launch(h, "learned solve setup"));
    bool on_mic = false;
    int queued = 0;
    const int k = h->d.refine_every;
    ChunkPlan plan(h->last_iters);
    int chunk = plan.first();
    for (;;) {
        const int n = std::min(chunk, h->d.max_iters - queued);
        for (int j = 0; j < n; ++j) {
            const int it = queued + j + 1;
            {
                Prof prof(h, ARIEL_KERNEL_SPMV, prof_sampled(it));
                launch_spmv_dot(s, p, h->A, h->L, h->S, h->Z, red, act);
            }
            {
                Prof prof(h, ARIEL_KERNEL_AXPY, prof_sampled(it));
                launch_copy_list(s, h->R, h->R_old, act, h->sc);
                launch_axpy2(s, h->P, h->S, h->R, h->Z, red, act);
            }
            if (k > 0 && it % k == 0)
                launch_reliable_update(s, p, h->A, h->L, h->P, h->P64, h->D, h->R, red, act);
            {
                Prof prof(h, ARIEL_KERNEL_PRECOND, prof_sampled(it));
                learned_precond(h, p, red, on_mic);
            }
            launch_update_s(s, h->Z, h->S, red, 0, act);
        }
        TRY(check_launch(h, "learned pcg"));
        queued += n;
        TRY(read_scalars(h));
        on_mic = on_mic || h->sc_host->use_mic;
        if (h->sc_host->done || queued >= h->d.max_iters) break;
        chunk = plan.next(*h->sc_host);
    }
    if (refine) {
        launch_true_residual_norm64(s, p, h->A, h->L, h->P, h->P64, h->D, red, act);
        launch_refine_finish(s, h->A, h->P, h->P64, act, h->sc);
    }
    HIPCHK(h, hipEventRecord(h->solve_t1, s));
    h->last_iters = h->sc_host->it;
    h->stats.iterations = h->sc_host->it;
    h->stats.solve_status = h->sc_host->status;
    h->stats.rhs_norm = std::sqrt(h->sc_host->bnorm2);
    h->stats.residual_norm = std::sqrt(h->sc_host->rnorm2);
    h->stats.precond_fallback = h->sc_host->use_mic ? h->sc_host->fallback_reason : 0;
    h->stats.fallback_iter = h->sc_host->use_mic ? h->sc_host->fallback_it : 0;
    launch_true_residual_norm(s, p, h->A, h->L, h->P, h->D, red, act);
    TRY(check_launch(h, "true residual"));
    TRY(read_scalars(h));
    h->stats.true_residual_norm = std::sqrt(h->sc_host->tmp);
    h->stats.true_residual_norm64 = refine ? std::sqrt(h->sc_host->tmp64) : h->stats.true_residual_norm;
    return ARIEL_OK;
}

}  // namespace host
}  // namespace ariel
