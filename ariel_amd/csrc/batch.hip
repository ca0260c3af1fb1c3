// batch.hip — independent tanks stepped together (ariel_batch_* in
// include/ariel.h; no counterpart in the reference, which runs one FlipSim per
// process).
#include "ariel_handle.h"

using namespace ariel;
using namespace ariel::host;

// ================================================================ batches
// Independent tanks stepped together: every stage is one launch over all tanks
// (blockIdx.y = tank). The launch sequence of a step is recorded per tank with
// batch sizing (solver counts read on the device), the tanks' Args are packed
// into device tables, and each launch position is issued once with the largest
// tank's grid. Kernels size their work by their own tank, so each tank's
// results equal those of stepping it alone. The pre-solve part, one PCG
// iteration (with and without reliable update) and the post-solve part can be
// captured as HIP graphs and relaunched every step.
struct BatchPos {
    const void* many = nullptr;
    dim3 grid, block;
    size_t arg_size = 0;
    size_t offset = 0;   // of the tank table inside the arena
};

struct BatchPhase {
    std::vector<BatchPos> pos;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> shape;   // layout the graph was captured for
};

struct ariel_batch {
    std::vector<ariel_sim*> tanks;
    int device = 0;
    hipStream_t stream = nullptr;
    BatchPhase pre, it_plain, it_refine, post;
    void* arena = nullptr;
    size_t arena_bytes = 0;
    unsigned char* host_arena = nullptr;   // pinned staging of the tables
    size_t host_bytes = 0;
    SolverScalars* sc_host = nullptr;      // pinned, one per tank
    bool graphs = false;
    int last_iters = 0;
    std::string err;
};

namespace {

int bfail(ariel_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    g_last_error = msg;
    return code;
}

#define BHIP(b, expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return bfail((b), ARIEL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// The launches of one tank's step, in four phases.
struct TankRec {
    Recorder pre, it_plain, it_refine, post;
};

void record_pre(ariel_sim* h, Recorder& r) {
    recorder() = &r;
    stage_sort(h);
    stage_density(h, 1);
    stage_mark(h);
    stage_p2g(h);
    stage_divergence(h);
    const Params p = h->params();
    hipStream_t s = h->stream;
    launch_precond_setup(s, p, h->A, h->L, h->diag, h->nup, h->pc);
    launch_fluid_segments(s, h->N, h->A, h->fl8);
    launch_ghost_levelset(s, h->N, h->A, h->L, h->Gt, h->sb);
    launch_solver_init(s, h->sc, h->d.max_iters, h->d.refine_every);
    launch_active_sets(s, h->N, h->fl8, h->seg_flag, h->seg_scan, h->seg_list, h->tile_any,
                       h->tile_lists, h->scan_sums, h->sc);
    set_active(h, true);
    launch_precond_build(s, p, h->A, h->diag, h->nup, h->pc, h->act, h->sc);
    enqueue_solve_setup(h, false);
    recorder() = nullptr;
}

void record_iteration(ariel_sim* h, Recorder& r, bool refine) {
    recorder() = &r;
    enqueue_iteration(h, 0, refine, false, false, false);
    recorder() = nullptr;
}

void record_post(ariel_sim* h, Recorder& r) {
    recorder() = &r;
    const Params p = h->params();
    const Reduce red = h->reduce();
    hipStream_t s = h->stream;
    enqueue_solve_finish(h);
    launch_true_residual_norm(s, p, h->A, h->L, h->P, h->D, red, h->act);
    stage_gradient(h);
    stage_extrapolate(h);
    stage_advect(h);
    stage_resample(h);
    recorder() = nullptr;
}

// Packs phase `which` of every tank into positions (tables appended to `host`).
int build_phase(ariel_batch* B, const std::vector<TankRec>& recs,
                Recorder TankRec::*which, BatchPhase& ph, std::vector<unsigned char>& host) {
    const size_t T = recs.size();
    const auto& first = (recs[0].*which).seq;
    ph.pos.clear();
    for (size_t i = 0; i < first.size(); ++i) {
        BatchPos bp;
        bp.many = first[i].many;
        bp.block = first[i].block;
        bp.arg_size = first[i].args.size();
        unsigned gx = 1;
        for (size_t t = 0; t < T; ++t) {
            const auto& seq = (recs[t].*which).seq;
            if (seq.size() != first.size() || seq[i].many != bp.many ||
                seq[i].block.x != bp.block.x || seq[i].args.size() != bp.arg_size)
                return bfail(B, ARIEL_ERR_INVALID, "batch: tanks record different launch sequences");
            gx = std::max(gx, seq[i].grid.x);
        }
        bp.grid = dim3(gx, (unsigned)T, 1);
        bp.offset = (host.size() + 255) / 256 * 256;
        host.resize(bp.offset + T * bp.arg_size);
        for (size_t t = 0; t < T; ++t)
            std::memcpy(host.data() + bp.offset + t * bp.arg_size, (recs[t].*which).seq[i].args.data(),
                        bp.arg_size);
        ph.pos.push_back(bp);
    }
    return ARIEL_OK;
}

std::vector<uint64_t> phase_shape(const BatchPhase& ph) {
    std::vector<uint64_t> v;
    for (const auto& p : ph.pos) {
        v.push_back((uint64_t)(uintptr_t)p.many);
        v.push_back(((uint64_t)p.grid.x << 32) | p.block.x);
        v.push_back(p.offset);
    }
    return v;
}

int issue_phase(ariel_batch* B, BatchPhase& ph) {
    for (const auto& p : ph.pos) {
        void* table = (unsigned char*)B->arena + p.offset;
        void* args[] = {&table};
        BHIP(B, hipLaunchKernel(p.many, p.grid, p.block, args, 0, B->stream));
    }
    return ARIEL_OK;
}

int run_phase(ariel_batch* B, BatchPhase& ph) {
    if (!B->graphs) return issue_phase(B, ph);
    const auto shape = phase_shape(ph);
    if (!ph.exec || shape != ph.shape) {
        if (ph.exec) hipGraphExecDestroy(ph.exec);
        if (ph.graph) hipGraphDestroy(ph.graph);
        ph.exec = nullptr;
        ph.graph = nullptr;
        BHIP(B, hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
        int rc = issue_phase(B, ph);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(B->stream, &g);
        if (rc != ARIEL_OK) return rc;
        if (e != hipSuccess) return bfail(B, ARIEL_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        ph.graph = g;
        BHIP(B, hipGraphInstantiate(&ph.exec, g, nullptr, nullptr, 0));
        ph.shape = shape;
    }
    BHIP(B, hipGraphLaunch(ph.exec, B->stream));
    return ARIEL_OK;
}

int read_batch_scalars(ariel_batch* B) {
    for (size_t t = 0; t < B->tanks.size(); ++t)
        BHIP(B, hipMemcpyAsync(&B->sc_host[t], B->tanks[t]->sc, sizeof(SolverScalars),
                               hipMemcpyDeviceToHost, B->stream));
    BHIP(B, hipStreamSynchronize(B->stream));
    return ARIEL_OK;
}

int batch_step_once(ariel_batch* B) {
    const size_t T = B->tanks.size();
    std::vector<TankRec> recs(T);
    for (size_t t = 0; t < T; ++t) {
        ariel_sim* h = B->tanks[t];
        h->frame++;
        record_pre(h, recs[t].pre);
        record_iteration(h, recs[t].it_plain, false);
        record_iteration(h, recs[t].it_refine, true);
        record_post(h, recs[t].post);
    }
    std::vector<unsigned char> host;
    TRY(build_phase(B, recs, &TankRec::pre, B->pre, host));
    TRY(build_phase(B, recs, &TankRec::it_plain, B->it_plain, host));
    TRY(build_phase(B, recs, &TankRec::it_refine, B->it_refine, host));
    TRY(build_phase(B, recs, &TankRec::post, B->post, host));
    if (host.size() > B->arena_bytes) {
        if (B->arena) hipFree(B->arena);
        if (B->host_arena) hipHostFree(B->host_arena);
        B->arena = nullptr;
        B->host_arena = nullptr;
        const size_t bytes = host.size() + host.size() / 4 + 4096;
        BHIP(B, hipMalloc(&B->arena, bytes));
        BHIP(B, hipHostMalloc((void**)&B->host_arena, bytes, hipHostMallocDefault));
        B->arena_bytes = B->host_bytes = bytes;
        // tables moved: every graph is captured again
        for (BatchPhase* ph : {&B->pre, &B->it_plain, &B->it_refine, &B->post}) ph->shape.clear();
    }
    // the previous step's graphs may still read the staging buffer's copy
    BHIP(B, hipStreamSynchronize(B->stream));
    std::memcpy(B->host_arena, host.data(), host.size());
    BHIP(B, hipMemcpyAsync(B->arena, B->host_arena, host.size(), hipMemcpyHostToDevice, B->stream));

    TRY(run_phase(B, B->pre));
    // PCG: iterations in chunks, one readback of every tank's scalars per chunk
    int max_iters = 0, k = 0;
    for (ariel_sim* h : B->tanks) {
        max_iters = std::max(max_iters, h->d.max_iters);
        k = h->d.refine_every;
    }
    int queued = 0;
    int chunk = std::max(4, B->last_iters - 3);
    for (;;) {
        const int n = std::min(chunk, max_iters - queued);
        for (int j = 0; j < n; ++j) {
            const int it = queued + j + 1;
            TRY(run_phase(B, (k > 0 && it % k == 0) ? B->it_refine : B->it_plain));
        }
        queued += n;
        TRY(read_batch_scalars(B));
        bool all_done = true;
        for (size_t t = 0; t < T; ++t) all_done = all_done && B->sc_host[t].done;
        if (all_done || queued >= max_iters) break;
        chunk = 4;
    }
    TRY(run_phase(B, B->post));
    TRY(read_batch_scalars(B));
    int most = 0;
    for (size_t t = 0; t < T; ++t) {
        ariel_sim* h = B->tanks[t];
        const SolverScalars& sc = B->sc_host[t];
        h->last_iters = sc.it;
        most = std::max(most, sc.it);
        h->stats.iterations = sc.it;
        h->stats.solve_status = sc.status;
        h->stats.rhs_norm = std::sqrt(sc.bnorm2);
        h->stats.residual_norm = std::sqrt(sc.rnorm2);
        h->stats.true_residual_norm = std::sqrt(sc.tmp);
        h->stats.true_residual_norm64 =
            h->d.refine_every > 0 ? std::sqrt(sc.tmp64) : h->stats.true_residual_norm;
        h->stats.fluid_cells = sc.fluid_cells;
        float md;
        uint32_t bits = sc.maxdiv_bits;
        std::memcpy(&md, &bits, 4);
        h->stats.max_divergence = md;
        h->stats.frame = h->frame;
    }
    B->last_iters = most;
    return ARIEL_OK;
}

void destroy_phase(BatchPhase& ph) {
    if (ph.exec) hipGraphExecDestroy(ph.exec);
    if (ph.graph) hipGraphDestroy(ph.graph);
    ph.exec = nullptr;
    ph.graph = nullptr;
}

// Autotuner: for each grid size, time P2G with each face-column tile edge on the
// first tank of that size and keep the fastest for all its tanks.
int autotune(ariel_batch* B) {
    std::vector<int> sizes;
    for (ariel_sim* h : B->tanks)
        if (std::find(sizes.begin(), sizes.end(), h->N) == sizes.end()) sizes.push_back(h->N);
    hipEvent_t e0, e1;
    BHIP(B, hipEventCreate(&e0));
    BHIP(B, hipEventCreate(&e1));
    for (int n : sizes) {
        ariel_sim* h = nullptr;
        for (ariel_sim* t : B->tanks)
            if (t->N == n) { h = t; break; }
        int best = h->p2g_tile;
        float best_ms = 1e30f;
        for (int tile : {4, 8, 16, 32}) {
            h->p2g_tile = tile;
            const Params p = h->params();
            for (int rep = 0; rep < 4; ++rep) {
                if (rep == 1) BHIP(B, hipEventRecord(e0, h->stream));
                launch_p2g(h->stream, p, h->pos[h->cur], h->vel[h->cur], h->offsets, h->A, h->u, h->du,
                           h->p2g_part, h->p2g_tflag);
            }
            BHIP(B, hipEventRecord(e1, h->stream));
            BHIP(B, hipEventSynchronize(e1));
            float ms = 0.f;
            BHIP(B, hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) {
                best_ms = ms;
                best = tile;
            }
        }
        for (ariel_sim* t : B->tanks)
            if (t->N == n) t->p2g_tile = best;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    // P2G left the tanks' grids as a fresh splat would; the next step redoes it
    return ARIEL_OK;
}

}  // namespace

extern "C" {

int ariel_batch_create(const ariel_scene_desc* descs, int32_t ntanks, ariel_batch** out) {
    if (!descs || !out || ntanks <= 0) return bfail(nullptr, ARIEL_ERR_INVALID, "batch_create: bad arguments");
    *out = nullptr;
    if (descs[0].precond == ARIEL_PRECOND_LEARNED || descs[0].precond == ARIEL_PRECOND_MG)
        return bfail(nullptr, ARIEL_ERR_UNSUPPORTED,
                     "batch_create: the learned and multigrid preconditioners run on single tanks only");
    for (int32_t t = 1; t < ntanks; ++t) {
        const ariel_scene_desc& a = descs[0];
        const ariel_scene_desc& b = descs[t];
        if (a.precond != b.precond || a.splat != b.splat || a.p2g != b.p2g || a.refine_every != b.refine_every ||
            a.resample != b.resample || a.max_iters != b.max_iters)
            return bfail(nullptr, ARIEL_ERR_INVALID,
                         "batch_create: tanks must share preconditioner, splat, P2G mode, refinement, "
                         "resampling and iteration limit");
    }
    ariel_batch* B = new ariel_batch();
    if (hipGetDevice(&B->device) != hipSuccess ||
        hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
        delete B;
        return bfail(nullptr, ARIEL_ERR_HIP, "batch_create: no device");
    }
    for (int32_t t = 0; t < ntanks; ++t) {
        ariel_sim* h = nullptr;
        const int rc = ariel_create(&descs[t], &h);
        if (rc != ARIEL_OK) {
            const std::string msg = g_last_error;
            ariel_batch_destroy(B);
            return bfail(nullptr, rc, "batch_create: tank " + std::to_string(t) + ": " + msg);
        }
        B->tanks.push_back(h);
    }
    if (hipHostMalloc((void**)&B->sc_host, sizeof(SolverScalars) * ntanks, hipHostMallocDefault) != hipSuccess) {
        ariel_batch_destroy(B);
        return bfail(nullptr, ARIEL_ERR_NOMEM, "batch_create: hipHostMalloc failed");
    }
    *out = B;
    return ARIEL_OK;
}

int ariel_batch_destroy(ariel_batch* B) {
    if (!B) return ARIEL_OK;
    if (B->stream) hipStreamSynchronize(B->stream);
    for (BatchPhase* ph : {&B->pre, &B->it_plain, &B->it_refine, &B->post}) destroy_phase(*ph);
    for (ariel_sim* h : B->tanks) ariel_destroy(h);
    if (B->arena) hipFree(B->arena);
    if (B->host_arena) hipHostFree(B->host_arena);
    if (B->sc_host) hipHostFree(B->sc_host);
    if (B->stream) hipStreamDestroy(B->stream);
    delete B;
    return ARIEL_OK;
}

const char* ariel_batch_last_error(const ariel_batch* B) {
    if (B) return B->err.c_str();
    return g_last_error.c_str();
}

int ariel_batch_size(ariel_batch* B, int32_t* n) {
    if (!B || !n) return bfail(B, ARIEL_ERR_INVALID, "null pointer");
    *n = (int32_t)B->tanks.size();
    return ARIEL_OK;
}

int ariel_batch_tank(ariel_batch* B, int32_t i, ariel_sim** out) {
    if (!B || !out) return bfail(B, ARIEL_ERR_INVALID, "null pointer");
    if (i < 0 || i >= (int32_t)B->tanks.size()) return bfail(B, ARIEL_ERR_INVALID, "tank index out of range");
    *out = B->tanks[i];
    return ARIEL_OK;
}

int ariel_batch_init(ariel_batch* B) {
    if (!B) return bfail(nullptr, ARIEL_ERR_INVALID, "null batch");
    for (size_t t = 0; t < B->tanks.size(); ++t) {
        const int rc = ariel_init(B->tanks[t]);
        if (rc != ARIEL_OK)
            return bfail(B, rc, "tank " + std::to_string(t) + ": " + B->tanks[t]->err);
    }
    return ARIEL_OK;
}

int ariel_batch_autotune(ariel_batch* B) {
    if (!B) return bfail(nullptr, ARIEL_ERR_INVALID, "null batch");
    for (ariel_sim* h : B->tanks)
        if (!h->initialized) return bfail(B, ARIEL_ERR_STATE, "autotune before ariel_batch_init");
    TRY(autotune(B));
    for (ariel_sim* h : B->tanks) BHIP(B, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

int ariel_batch_get_tuning(ariel_batch* B, int32_t i, int32_t* p2g_tile) {
    if (!B || !p2g_tile) return bfail(B, ARIEL_ERR_INVALID, "null pointer");
    if (i < 0 || i >= (int32_t)B->tanks.size()) return bfail(B, ARIEL_ERR_INVALID, "tank index out of range");
    *p2g_tile = B->tanks[i]->p2g_tile;
    return ARIEL_OK;
}

int ariel_batch_set_graphs(ariel_batch* B, int32_t enable) {
    if (!B) return bfail(nullptr, ARIEL_ERR_INVALID, "null batch");
    B->graphs = enable != 0;
    return ARIEL_OK;
}

int ariel_batch_step(ariel_batch* B, int32_t nsteps) {
    if (!B) return bfail(nullptr, ARIEL_ERR_INVALID, "null batch");
    if (nsteps < 0) return bfail(B, ARIEL_ERR_INVALID, "negative step count");
    for (ariel_sim* h : B->tanks) {
        if (!h->initialized) return bfail(B, ARIEL_ERR_STATE, "ariel_batch_step before ariel_batch_init");
        if (h->profile) return bfail(B, ARIEL_ERR_STATE, "profiling is per handle; disable it on batch tanks");
        // earlier single-tank work on the tanks' own streams comes first
        BHIP(B, hipStreamSynchronize(h->stream));
    }
    if (hipSetDevice(B->device) != hipSuccess) return bfail(B, ARIEL_ERR_HIP, "hipSetDevice failed");
    for (int32_t n = 0; n < nsteps; ++n) TRY(batch_step_once(B));
    BHIP(B, hipStreamSynchronize(B->stream));
    return ARIEL_OK;
}

int ariel_batch_step_sequential(ariel_batch* B, int32_t nsteps) {
    if (!B) return bfail(nullptr, ARIEL_ERR_INVALID, "null batch");
    BHIP(B, hipStreamSynchronize(B->stream));
    for (int32_t n = 0; n < nsteps; ++n)
        for (ariel_sim* h : B->tanks) {
            const int rc = ariel_step(h, 1);
            if (rc != ARIEL_OK) return bfail(B, rc, h->err);
        }
    for (ariel_sim* h : B->tanks) BHIP(B, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

}  // extern "C"
