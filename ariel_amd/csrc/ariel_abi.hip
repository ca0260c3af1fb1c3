// ariel_abi.hip — host side of the C ABI (include/ariel.h): device state of one
// simulation handle and the orchestration of FlipSim::Init / FlipSim::Step
// (src/sim/flip.cpp:42-128) as a sequence of HIP kernel launches on the
// handle's stream. Host work per step is launch-only apart from one small
// device->host read of the solver scalars per chunk of PCG iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ariel.h"
#include "ariel_device.h"
#include "ariel_internal.h"

using namespace ariel;

namespace {

thread_local std::string g_last_error = "";

struct Stage {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    bool recorded = false;
};

struct KTimer {
    int kernel;
    hipEvent_t e0, e1;
};

}  // namespace

void ariel::set_last_error(const std::string& msg) { g_last_error = msg; }

struct ariel_sim {
    ariel_scene_desc d{};
    int N = 0;
    int64_t nc = 0, nf = 0, np = 0;
    float maxd = 0.f;
    float max_density = 0.f;
    int frame = 0;
    bool initialized = false;
    int last_iters = 0;
    int p2g_tile = 8;             // P2G face-column tile edge (autotuned in batches)
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0;

    // particles (sorted by (cell, id)); b = current buffer
    float4* pos[2] = {nullptr, nullptr};
    float4* vel[2] = {nullptr, nullptr};
    int cur = 0;
    int32_t* cell_of = nullptr;
    int32_t* slots = nullptr;
    int32_t* sid = nullptr;       // particle ids next to slots (per-cell ordering)
    int32_t* counts = nullptr;
    int32_t* offsets = nullptr;
    int32_t* scan_sums = nullptr;
    uint8_t* accept = nullptr;
    int64_t np_cap = 0;

    // grids
    float* u[3] = {nullptr, nullptr, nullptr};
    float* du[3] = {nullptr, nullptr, nullptr};
    float *P = nullptr, *D = nullptr, *L = nullptr;
    uint8_t* A = nullptr;
    // pc: 2 N^3 floats, the factor in the natural layout and then the tile-major
    // copy the MIC(0) apply reads; Q is the apply's intermediate, tile-major
    float *R = nullptr, *Z = nullptr, *S = nullptr, *Q = nullptr, *pc = nullptr, *diag = nullptr;
    uint8_t* nup = nullptr;
    uint8_t* fl8 = nullptr;       // fluid bits per 8-cell z-segment
    double* P64 = nullptr;        // FP64 solution of the refined solve
    // active sets of the pressure solve (segments, tiles), rebuilt per solve
    int32_t *seg_flag = nullptr, *seg_scan = nullptr, *seg_list = nullptr;
    int32_t *tile_any = nullptr, *tile_lists = nullptr;
    Active act;
    // learned preconditioner: weights, activations, r_old (allocated on load)
    LearnedNet net;
    uint16_t* net_mem = nullptr;
    float* R_old = nullptr;
    bool net_ready = false;
    double* partials = nullptr;
    uint32_t* counter = nullptr;
    SolverScalars* sc = nullptr;
    SolverScalars* sc_host = nullptr;   // pinned

    Stage stages[ARIEL_NUM_STAGES];
    bool profile = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<KTimer> pending;
    ariel_kernel_stats kstats[ARIEL_NUM_KERNELS]{};
    hipEvent_t step_t0 = nullptr, step_t1 = nullptr, solve_t0 = nullptr, solve_t1 = nullptr;
    ariel_step_stats stats{};

    Params params() const {
        Params p;
        p.N = N;
        p.maxd = maxd;
        p.density = d.density;
        p.dt = d.dt;
        p.picflip = d.picflip;
        p.sharpen_re = d.sharpen_re;
        p.spring = d.spring;
        p.splat = d.splat;
        p.wall = d.wall;
        p.seed = d.seed;
        p.n_forces = d.n_forces;
        for (int f = 0; f < 4; ++f)
            for (int c = 0; c < 3; ++c) p.forces[f][c] = d.forces[f][c];
        p.max_density = max_density;
        p.frame = frame;
        p.mic_tau = d.mic_tau;
        p.mic_sigma = d.mic_sigma;
        // the learned preconditioner's fallback and set-up are the tiled MIC(0)
        p.precond = d.precond == ARIEL_PRECOND_LEARNED ? ARIEL_PRECOND_MIC0_TILED : d.precond;
        p.p2g_tile = p2g_tile;
        return p;
    }
    Reduce reduce() const { return Reduce{partials, counter, sc}; }
};

namespace ariel {
Recorder*& recorder() {
    static thread_local Recorder* r = nullptr;
    return r;
}
bool& batch_sizing() {
    static thread_local bool b = false;
    return b;
}
}  // namespace ariel

namespace {

int fail(ariel_sim* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    g_last_error = msg;
    return code;
}

#define HIPCHK(h, expr)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail((h), ARIEL_ERR_HIP,                                               \
                        std::string(#expr) + ": " + hipGetErrorString(e_));               \
    } while (0)

template <class T>
int dalloc(ariel_sim* h, T** p, int64_t n) {
    *p = nullptr;
    if (n <= 0) n = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(h, ARIEL_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return ARIEL_OK;
}

#define TRY(expr)                       \
    do {                                \
        int rc_ = (expr);               \
        if (rc_ != ARIEL_OK) return rc_; \
    } while (0)

int check_launch(ariel_sim* h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(h, ARIEL_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return ARIEL_OK;
}

hipEvent_t take_event(ariel_sim* h) {
    if (!h->ev_pool.empty()) {
        hipEvent_t e = h->ev_pool.back();
        h->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Algorithmic HBM bytes of one launch group: each input read once and each
// output written once. The solver groups count the cells they touch: the
// active segments (8 cells each) and, for the preconditioner, the active tiles
// (512 cells each; r and pc read, z written, and q written and read again on
// the red half).
double kernel_bytes(const ariel_sim* h, int k) {
    const double np = (double)h->np, nc = (double)h->nc;
    const double seg_cells = 8.0 * (double)h->act.nseg;
    const double tile_cells = 512.0 * (double)h->act.ntiles[2];
    switch (k) {
        case ARIEL_KERNEL_SORT: return 60.0 * np + 8.0 * nc;
        case ARIEL_KERNEL_DENSITY: return 40.0 * np;
        case ARIEL_KERNEL_MARK: return 4.0 * np + 9.0 * nc;
        case ARIEL_KERNEL_P2G: return 24.0 * np + 12.0 * nc;
        case ARIEL_KERNEL_SPMV: return 9.0 * seg_cells;
        case ARIEL_KERNEL_PRECOND: return 20.0 * tile_cells;
        case ARIEL_KERNEL_AXPY: return 24.0 * seg_cells;
        case ARIEL_KERNEL_GRADIENT: return 33.0 * nc;
        case ARIEL_KERNEL_EXTRAPOLATE: return 36.0 * nc;
        case ARIEL_KERNEL_ADVECT: return 48.0 * np + 24.0 * nc;
        case ARIEL_KERNEL_RESAMPLE: return 48.0 * np;
        default: return 0.0;
    }
}

// Each event record leaves the GPU idle for a few microseconds between the
// kernels it separates, so inside the PCG loop only one iteration in
// kProfEvery is timed (the launches do not depend on the iteration index).
constexpr int kProfEvery = 16;
inline bool prof_sampled(int it) { return (it - 1) % kProfEvery == 0; }

struct Prof {
    ariel_sim* h;
    int k;
    hipEvent_t e0 = nullptr;
    Prof(ariel_sim* hh, int kk, bool on = true) : h(hh), k(kk) {
        if (h->profile) h->kstats[k].calls += 1;
        if (h->profile && on) {
            e0 = take_event(h);
            if (e0) (void)hipEventRecord(e0, h->stream);
        }
    }
    ~Prof() {
        if (!e0) return;
        hipEvent_t e1 = take_event(h);
        if (!e1) return;
        (void)hipEventRecord(e1, h->stream);
        h->pending.push_back(KTimer{k, e0, e1});
    }
};

// Called after a stream synchronisation: fold finished timers into kstats.
void drain_timers(ariel_sim* h) {
    for (const KTimer& t : h->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.e0, t.e1) == hipSuccess) {
            h->kstats[t.kernel].launches += 1;
            h->kstats[t.kernel].total_ms += ms;
        }
        h->ev_pool.push_back(t.e0);
        h->ev_pool.push_back(t.e1);
    }
    h->pending.clear();
}

int validate(const ariel_scene_desc* d, std::string& msg) {
    if (!d) { msg = "null descriptor"; return ARIEL_ERR_INVALID; }
    const int n = d->dims[0];
    if (n != d->dims[1] || n != d->dims[2]) { msg = "grid must be cubic (dims equal)"; return ARIEL_ERR_INVALID; }
    if (n < 8 || n % 8 != 0) { msg = "grid edge must be a positive multiple of 8"; return ARIEL_ERR_UNSUPPORTED; }
    if (n > 1024) { msg = "grid edge above 1024"; return ARIEL_ERR_UNSUPPORTED; }
    if (!(d->density > 0.f)) { msg = "density must be positive"; return ARIEL_ERR_INVALID; }
    const double m = 1.0 / d->density;
    if (std::fabs(m - std::round(m)) > 1e-4 || std::round(m) < 1) {
        msg = "density must be 1/m for an integer m >= 1"; return ARIEL_ERR_INVALID;
    }
    if (!(d->dt > 0.f)) { msg = "dt must be positive"; return ARIEL_ERR_INVALID; }
    if (d->n_forces < 0 || d->n_forces > ARIEL_MAX_FORCES) { msg = "n_forces out of range"; return ARIEL_ERR_INVALID; }
    if (d->n_liquids < 0 || d->n_liquids > ARIEL_MAX_LIQUIDS) { msg = "n_liquids out of range"; return ARIEL_ERR_INVALID; }
    for (int l = 0; l < d->n_liquids; ++l)
        if (d->liquids[l].kind != ARIEL_LIQUID_BOX && d->liquids[l].kind != ARIEL_LIQUID_SPHERE) {
            msg = "unknown liquid kind"; return ARIEL_ERR_INVALID;
        }
    if (d->wall < 0 || 2 * d->wall >= n) { msg = "wall thickness out of range"; return ARIEL_ERR_INVALID; }
    if (d->precond == ARIEL_PRECOND_MIC0 || d->precond == ARIEL_PRECOND_REFERENCE) {
        msg = "preconditioner not available on the GPU path (use MIC0_TILED, MIC0_BLOCK, JACOBI "
              "or LEARNED)";
        return ARIEL_ERR_UNSUPPORTED;
    }
    if (d->precond != ARIEL_PRECOND_MIC0_TILED && d->precond != ARIEL_PRECOND_MIC0_BLOCK &&
        d->precond != ARIEL_PRECOND_JACOBI && d->precond != ARIEL_PRECOND_LEARNED) {
        msg = "unknown preconditioner"; return ARIEL_ERR_INVALID;
    }
    if ((d->precond == ARIEL_PRECOND_MIC0_TILED || d->precond == ARIEL_PRECOND_MIC0_BLOCK) &&
        d->tile != 8) {
        msg = "tiled MIC(0) uses 8^3 tiles (tile must be 8)"; return ARIEL_ERR_UNSUPPORTED;
    }
    if (!(d->tol > 0.f) && d->stop_rule == ARIEL_STOP_RELATIVE) { msg = "tol must be positive"; return ARIEL_ERR_INVALID; }
    if (d->max_iters < 1) { msg = "max_iters must be >= 1"; return ARIEL_ERR_INVALID; }
    if (d->refine_every < 0) { msg = "refine_every must be >= 0"; return ARIEL_ERR_INVALID; }
    if (d->stop_rule != ARIEL_STOP_RELATIVE && d->stop_rule != ARIEL_STOP_REFERENCE_ABS) { msg = "unknown stop rule"; return ARIEL_ERR_INVALID; }
    if (d->splat != ARIEL_SPLAT_SHARPEN && d->splat != ARIEL_SPLAT_TRILINEAR) { msg = "unknown splat kernel"; return ARIEL_ERR_INVALID; }
    return ARIEL_OK;
}

// ----------------------------------------------------------------- seeding
// Scene::GenerateParticles (scene.cpp:257-360) + AddLiquidParticle (436-455) with
// the deterministic order: liquids in descriptor order, interior cells in i,j,k
// order, lattice octants lexicographically; lattice spacing w = density*(1/maxd),
// point a at a*w + u*w, u = 0.5 (reference lattice) or a per-point uniform jitter
// u = (hash4(seed, 3l+axis, a*N*m+b, c) >> 8) * 2^-24.
struct Seeded {
    std::vector<float4> pos, vel;
};

void seed_range(const ariel_scene_desc& d, int N, int l, int ci0, int ci1, Seeded& out) {
    const float maxd = (float)N;
    const float thickness = 1.0f / maxd;
    const float w = d.density * thickness;
    const int m = (int)std::lround(1.0 / d.density);
    const uint32_t Nm = (uint32_t)(N * m);
    const int wall = d.wall;
    const ariel_liquid& lq = d.liquids[l];
    for (int ci = ci0; ci < ci1; ++ci)
        for (int cj = wall; cj < N - wall; ++cj)
            for (int ck = wall; ck < N - wall; ++ck)
                for (int oa = 0; oa < m; ++oa)
                    for (int ob = 0; ob < m; ++ob)
                        for (int oc = 0; oc < m; ++oc) {
                            const uint32_t a = (uint32_t)(ci * m + oa), b = (uint32_t)(cj * m + ob),
                                           c = (uint32_t)(ck * m + oc);
                            float jx = 0.5f, jy = 0.5f, jz = 0.5f;
                            if (d.jitter) {
                                const uint32_t key = a * Nm + b;
                                jx = (float)(hash4(d.seed, 3u * l + 0u, key, c) >> 8) * (1.0f / 16777216.0f);
                                jy = (float)(hash4(d.seed, 3u * l + 1u, key, c) >> 8) * (1.0f / 16777216.0f);
                                jz = (float)(hash4(d.seed, 3u * l + 2u, key, c) >> 8) * (1.0f / 16777216.0f);
                            }
                            const float x = (float)a * w + jx * w, y = (float)b * w + jy * w,
                                        z = (float)c * w + jz * w;
                            bool inside;
                            if (lq.kind == ARIEL_LIQUID_BOX) {
                                inside = x >= lq.lo[0] && x < lq.hi[0] && y >= lq.lo[1] && y < lq.hi[1] &&
                                         z >= lq.lo[2] && z < lq.hi[2];
                            } else {
                                const float dx = x - lq.center[0], dy = y - lq.center[1], dz = z - lq.center[2];
                                inside = dx * dx + dy * dy + dz * dz < lq.radius * lq.radius;
                            }
                            if (!inside) continue;
                            out.pos.push_back(make_float4(x, y, z, 0.0f));
                            out.vel.push_back(make_float4(lq.velocity[0], lq.velocity[1], lq.velocity[2], 10.0f));
                        }
}

void seed_particles(const ariel_scene_desc& d, int N, std::vector<float4>& pos, std::vector<float4>& vel) {
    pos.clear();
    vel.clear();
    const int wall = d.wall;
    unsigned hw = std::thread::hardware_concurrency();
    const int nthreads = (int)std::max(1u, std::min(hw, 16u));
    for (int l = 0; l < d.n_liquids; ++l) {
        const int span = N - 2 * wall;
        const int chunks = std::max(1, std::min(nthreads, span));
        std::vector<Seeded> parts(chunks);
        std::vector<std::thread> ts;
        for (int t = 0; t < chunks; ++t) {
            const int a0 = wall + span * t / chunks, a1 = wall + span * (t + 1) / chunks;
            ts.emplace_back([&, t, a0, a1] { seed_range(d, N, l, a0, a1, parts[t]); });
        }
        for (auto& th : ts) th.join();
        for (auto& p : parts) {
            pos.insert(pos.end(), p.pos.begin(), p.pos.end());
            vel.insert(vel.end(), p.vel.begin(), p.vel.end());
        }
    }
    for (size_t i = 0; i < pos.size(); ++i) {
        int32_t id = (int32_t)i;
        float f;
        std::memcpy(&f, &id, 4);
        pos[i].w = f;
    }
}

int ensure_particles(ariel_sim* h, int64_t np) {
    if (np <= h->np_cap && h->pos[0]) return ARIEL_OK;
    for (int b = 0; b < 2; ++b) {
        hipFree(h->pos[b]);
        hipFree(h->vel[b]);
        h->pos[b] = nullptr;
        h->vel[b] = nullptr;
    }
    hipFree(h->cell_of);
    hipFree(h->slots);
    hipFree(h->sid);
    hipFree(h->accept);
    h->cell_of = h->slots = h->sid = nullptr;
    h->accept = nullptr;
    for (int b = 0; b < 2; ++b) {
        TRY(dalloc(h, &h->pos[b], np));
        TRY(dalloc(h, &h->vel[b], np));
    }
    TRY(dalloc(h, &h->cell_of, np));
    TRY(dalloc(h, &h->slots, np));
    TRY(dalloc(h, &h->sid, np));
    TRY(dalloc(h, &h->accept, np));
    HIPCHK(h, hipMemsetAsync(h->accept, 1, (size_t)std::max<int64_t>(np, 1), h->stream));
    h->np_cap = np;
    return ARIEL_OK;
}

// ----------------------------------------------------------------- stages
int stage_sort(ariel_sim* h) {
    const int b = h->cur, o = 1 - b;
    Prof prof(h, ARIEL_KERNEL_SORT);
    launch_bin(h->stream, h->pos[b], h->np, h->N, h->maxd, h->cell_of, h->counts);
    launch_scan(h->stream, h->counts, h->offsets, h->nc, h->scan_sums);
    launch_scatter(h->stream, h->cell_of, h->np, h->offsets, h->counts, h->slots, h->pos[b], h->sid);
    launch_cellsort(h->stream, h->offsets, h->nc, h->sid, h->slots);
    launch_gather(h->stream, h->slots, h->np, h->pos[b], h->vel[b], h->pos[o], h->vel[o]);
    h->cur = o;
    return check_launch(h, "sort");
}

int stage_density(ariel_sim* h, int forces) {
    Prof prof(h, ARIEL_KERNEL_DENSITY);
    launch_density(h->stream, h->params(), h->pos[h->cur], h->vel[h->cur], h->np, h->offsets, forces);
    return check_launch(h, "density");
}

int stage_mark(ariel_sim* h) {
    Prof prof(h, ARIEL_KERNEL_MARK);
    launch_mark(h->stream, h->params(), h->vel[h->cur], h->offsets, h->A, h->L);
    return check_launch(h, "mark");
}

int stage_p2g(ariel_sim* h) {
    Prof prof(h, ARIEL_KERNEL_P2G);
    launch_p2g(h->stream, h->params(), h->pos[h->cur], h->vel[h->cur], h->offsets, h->A, h->u, h->du);
    return check_launch(h, "p2g");
}

int stage_divergence(ariel_sim* h) {
    launch_divergence(h->stream, h->params(), h->u, h->D);
    return check_launch(h, "divergence");
}

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
    if (refine) launch_refine_finish(s, h->A, h->P, h->P64, act, h->sc);
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
    if (refine) launch_refine_finish(s, h->A, h->P, h->P64, act, h->sc);
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
    return ARIEL_OK;
}

int stage_gradient(ariel_sim* h) {
    Prof prof(h, ARIEL_KERNEL_GRADIENT);
    launch_gradient(h->stream, h->params(), h->P, h->L, h->A, h->u);
    launch_grid_stats(h->stream, h->params(), h->u, h->A, h->sc);
    return check_launch(h, "gradient");
}

int stage_extrapolate(ariel_sim* h) {
    Prof prof(h, ARIEL_KERNEL_EXTRAPOLATE);
    launch_extrapolate(h->stream, h->params(), h->A, h->u, h->du);
    return check_launch(h, "extrapolate");
}

int stage_advect(ariel_sim* h) {
    Prof prof(h, ARIEL_KERNEL_ADVECT);
    launch_g2p_advect(h->stream, h->params(), h->pos[h->cur], h->vel[h->cur], h->np, h->u, h->du);
    return check_launch(h, "advect");
}

int stage_resample(ariel_sim* h) {
    if (!h->d.resample) return ARIEL_OK;
    TRY(stage_sort(h));
    const int b = h->cur, o = 1 - b;
    Prof prof(h, ARIEL_KERNEL_RESAMPLE);
    launch_resample(h->stream, h->params(), h->pos[b], h->vel[b], h->np, h->offsets, h->pos[o],
                    h->vel[o], h->accept);
    h->cur = o;
    return check_launch(h, "resample");
}

int run_stage(ariel_sim* h, int stage) {
    Stage& st = h->stages[stage];
    HIPCHK(h, hipEventRecord(st.t0, h->stream));
    int rc = ARIEL_OK;
    switch (stage) {
        case ARIEL_STAGE_SORT: rc = stage_sort(h); break;
        case ARIEL_STAGE_DENSITY: rc = stage_density(h, 1); break;
        case ARIEL_STAGE_MARK: rc = stage_mark(h); break;
        case ARIEL_STAGE_P2G: rc = stage_p2g(h); break;
        case ARIEL_STAGE_DIVERGENCE: rc = stage_divergence(h); break;
        case ARIEL_STAGE_SOLVE: rc = stage_solve(h); break;
        case ARIEL_STAGE_GRADIENT: rc = stage_gradient(h); break;
        case ARIEL_STAGE_EXTRAPOLATE: rc = stage_extrapolate(h); break;
        case ARIEL_STAGE_ADVECT: rc = stage_advect(h); break;
        case ARIEL_STAGE_RESAMPLE: rc = stage_resample(h); break;
        default: return fail(h, ARIEL_ERR_INVALID, "unknown stage");
    }
    if (rc != ARIEL_OK) return rc;
    HIPCHK(h, hipEventRecord(st.t1, h->stream));
    st.recorded = true;
    return ARIEL_OK;
}

int collect_stats(ariel_sim* h) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drain_timers(h);
    for (int i = 0; i < ARIEL_NUM_STAGES; ++i) {
        float ms = 0.f;
        if (h->stages[i].recorded && hipEventElapsedTime(&ms, h->stages[i].t0, h->stages[i].t1) == hipSuccess)
            h->stats.stage_ms[i] = ms;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->step_t0, h->step_t1) == hipSuccess) h->stats.step_ms = ms;
    if (hipEventElapsedTime(&ms, h->solve_t0, h->solve_t1) == hipSuccess) h->stats.solve_ms = ms;
    TRY(read_scalars(h));
    h->stats.fluid_cells = h->sc_host->fluid_cells;
    float md;
    uint32_t bits = h->sc_host->maxdiv_bits;
    std::memcpy(&md, &bits, 4);
    h->stats.max_divergence = md;
    (void)hipGetLastError();
    return ARIEL_OK;
}

int check_handle(ariel_sim* h) {
    if (!h) return fail(nullptr, ARIEL_ERR_INVALID, "null handle");
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return fail(h, ARIEL_ERR_HIP, hipGetErrorString(e));
    return ARIEL_OK;
}

int64_t field_size(const ariel_sim* h, int which) { return which <= ARIEL_FIELD_DUZ ? h->nf : h->nc; }

float* field_ptr(ariel_sim* h, int which) {
    switch (which) {
        case ARIEL_FIELD_UX: case ARIEL_FIELD_UY: case ARIEL_FIELD_UZ: return h->u[which];
        case ARIEL_FIELD_DUX: case ARIEL_FIELD_DUY: case ARIEL_FIELD_DUZ: return h->du[which - 3];
        case ARIEL_FIELD_P: return h->P;
        case ARIEL_FIELD_D: return h->D;
        case ARIEL_FIELD_L: return h->L;
        default: return nullptr;
    }
}

// Max-density calibration (flip.cpp:42-71): a 10^3 lattice at spacing
// density/maxd, sorted and run through ComputeDensity with m_max_density = 1;
// the largest density becomes m_max_density.
int calibrate_max_density(ariel_sim* h) {
    const float hsp = h->d.density / h->maxd;
    std::vector<float4> tp(1000), tv(1000, make_float4(0, 0, 0, 0));
    for (int i = 0, n = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j)
            for (int k = 0; k < 10; ++k, ++n) {
                int32_t id = n;
                float f;
                std::memcpy(&f, &id, 4);
                tp[n] = make_float4(((float)i + 0.5f) * hsp, ((float)j + 0.5f) * hsp, ((float)k + 0.5f) * hsp, f);
            }
    TRY(ensure_particles(h, 1000));
    h->cur = 0;
    h->np = 1000;
    HIPCHK(h, hipMemcpyAsync(h->pos[0], tp.data(), 1000 * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->vel[0], tv.data(), 1000 * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    TRY(stage_sort(h));
    h->max_density = 1.0f;
    TRY(stage_density(h, 0));
    HIPCHK(h, hipMemcpyAsync(tv.data(), h->vel[h->cur], 1000 * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    float md = 0.0f;
    for (const float4& v : tv) md = gmax(md, v.w);
    h->max_density = md;
    return check_launch(h, "max-density calibration");
}


}  // namespace

// =================================================================== C ABI
extern "C" {

int ariel_desc_defaults(ariel_scene_desc* d) {
    if (!d) return fail(nullptr, ARIEL_ERR_INVALID, "null descriptor");
    std::memset(d, 0, sizeof(*d));
    d->dims[0] = d->dims[1] = d->dims[2] = 32;
    d->density = 0.5f;
    d->dt = 0.005f;
    d->picflip = 0.95f;
    d->precond = ARIEL_PRECOND_MIC0_TILED;
    d->mic_tau = 0.97f;
    d->mic_sigma = 0.25f;
    d->tile = 8;
    d->tol = 1e-6f;
    d->max_iters = 500;
    d->refine_every = 0;
    d->stop_rule = ARIEL_STOP_RELATIVE;
    d->splat = ARIEL_SPLAT_SHARPEN;
    d->resample = 1;
    d->sharpen_re = 1.4f;
    d->spring = 50.0f;
    return ARIEL_OK;
}

int ariel_create(const ariel_scene_desc* d, ariel_sim** out) {
    if (!out) return fail(nullptr, ARIEL_ERR_INVALID, "null output pointer");
    *out = nullptr;
    std::string msg;
    int rc = validate(d, msg);
    if (rc != ARIEL_OK) return fail(nullptr, rc, msg);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(nullptr, ARIEL_ERR_HIP, "no HIP device visible");
    ariel_sim* h = new ariel_sim();
    h->d = *d;
    h->N = d->dims[0];
    h->nc = (int64_t)h->N * h->N * h->N;
    h->nf = (int64_t)(h->N + 1) * h->N * h->N;
    h->maxd = std::max(std::max((float)d->dims[0], (float)d->dims[2]), (float)d->dims[1]);
    hipGetDevice(&h->device);
    auto bail = [&](int code) {
        ariel_destroy(h);
        return code;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(nullptr, ARIEL_ERR_HIP, "hipStreamCreate failed"));
    for (int c = 0; c < 3; ++c) {
        if ((rc = dalloc(h, &h->u[c], h->nf)) || (rc = dalloc(h, &h->du[c], h->nf))) return bail(rc);
    }
    if ((rc = dalloc(h, &h->P, h->nc)) || (rc = dalloc(h, &h->D, h->nc)) || (rc = dalloc(h, &h->L, h->nc)) ||
        (rc = dalloc(h, &h->A, h->nc)) || (rc = dalloc(h, &h->R, h->nc)) || (rc = dalloc(h, &h->Z, h->nc)) ||
        (rc = dalloc(h, &h->S, h->nc)) || (rc = dalloc(h, &h->Q, h->nc)) || (rc = dalloc(h, &h->pc, 2 * h->nc)) ||
        (rc = dalloc(h, &h->diag, h->nc)) || (rc = dalloc(h, &h->nup, h->nc)) ||
        (rc = dalloc(h, &h->fl8, h->nc / 8)) ||
        (rc = dalloc(h, &h->counts, h->nc)) || (rc = dalloc(h, &h->offsets, h->nc + 1)) ||
        (rc = dalloc(h, &h->scan_sums, h->nc / 4096 + 4)) ||
        (rc = dalloc(h, &h->P64, h->d.refine_every > 0 ? h->nc : 1)) ||
        (rc = dalloc(h, &h->seg_flag, h->nc / 8)) || (rc = dalloc(h, &h->seg_scan, h->nc / 8 + 1)) ||
        (rc = dalloc(h, &h->seg_list, h->nc / 8)) || (rc = dalloc(h, &h->tile_any, h->nc / 512)) ||
        (rc = dalloc(h, &h->tile_lists, 3 * (h->nc / 512))) ||
        (rc = dalloc(h, &h->partials, 2 * (int64_t)kMaxPartials)) || (rc = dalloc(h, &h->counter, kCounterWords)) ||
        (rc = dalloc(h, &h->sc, 1)))
        return bail(rc);
    if (hipHostMalloc((void**)&h->sc_host, sizeof(SolverScalars), hipHostMallocDefault) != hipSuccess)
        return bail(fail(nullptr, ARIEL_ERR_NOMEM, "hipHostMalloc failed"));
    for (int i = 0; i < ARIEL_NUM_STAGES; ++i) {
        hipEventCreate(&h->stages[i].t0);
        hipEventCreate(&h->stages[i].t1);
    }
    hipEventCreate(&h->step_t0);
    hipEventCreate(&h->step_t1);
    hipEventCreate(&h->solve_t0);
    hipEventCreate(&h->solve_t1);
    hipStream_t s = h->stream;
    for (int c = 0; c < 3; ++c) {
        launch_fill_f32(s, h->u[c], h->nf, 0.0f);
        launch_fill_f32(s, h->du[c], h->nf, 0.0f);
    }
    launch_fill_f32(s, h->P, h->nc, 0.0f);
    launch_fill_f32(s, h->D, h->nc, 0.0f);
    launch_fill_f32(s, h->L, h->nc, 1.6f);   // macgrid.inl:82
    for (float* v : {h->R, h->Z, h->S, h->Q, h->diag}) launch_fill_f32(s, v, h->nc, 0.0f);
    launch_fill_f32(s, h->pc, 2 * h->nc, 0.0f);   // natural layout, then the MIC(0) apply's tile-major copy
    launch_fill_i32(s, h->counts, h->nc, 0);
    launch_fill_i32(s, h->offsets, h->nc + 1, 0);
    if (hipMemsetAsync(h->A, 0, (size_t)h->nc, s) != hipSuccess ||
        hipMemsetAsync(h->nup, 0, (size_t)h->nc, s) != hipSuccess ||
        hipMemsetAsync(h->counter, 0, sizeof(uint32_t) * kCounterWords, s) != hipSuccess ||
        hipMemsetAsync(h->sc, 0, sizeof(SolverScalars), s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return bail(fail(nullptr, ARIEL_ERR_HIP, "device initialisation failed"));
    *out = h;
    return ARIEL_OK;
}

int ariel_destroy(ariel_sim* h) {
    if (!h) return ARIEL_OK;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    for (int b = 0; b < 2; ++b) {
        hipFree(h->pos[b]);
        hipFree(h->vel[b]);
    }
    for (void* p : {(void*)h->cell_of, (void*)h->slots, (void*)h->sid, (void*)h->counts, (void*)h->offsets,
                    (void*)h->scan_sums, (void*)h->accept, (void*)h->P, (void*)h->D, (void*)h->L,
                    (void*)h->A, (void*)h->R, (void*)h->Z, (void*)h->S, (void*)h->Q, (void*)h->pc,
                    (void*)h->diag, (void*)h->nup, (void*)h->fl8, (void*)h->P64, (void*)h->partials, (void*)h->counter, (void*)h->sc,
                    (void*)h->seg_flag, (void*)h->seg_scan, (void*)h->seg_list, (void*)h->tile_any,
                    (void*)h->tile_lists, (void*)h->net_mem, (void*)h->R_old, (void*)h->net.wd})
        hipFree(p);
    for (int c = 0; c < 3; ++c) {
        hipFree(h->u[c]);
        hipFree(h->du[c]);
    }
    if (h->sc_host) hipHostFree(h->sc_host);
    for (int i = 0; i < ARIEL_NUM_STAGES; ++i) {
        if (h->stages[i].t0) hipEventDestroy(h->stages[i].t0);
        if (h->stages[i].t1) hipEventDestroy(h->stages[i].t1);
    }
    for (hipEvent_t e : {h->step_t0, h->step_t1, h->solve_t0, h->solve_t1})
        if (e) hipEventDestroy(e);
    drain_timers(h);
    for (hipEvent_t e : h->ev_pool) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return ARIEL_OK;
}

const char* ariel_last_error(const ariel_sim* h) {
    if (h) return h->err.c_str();
    return g_last_error.c_str();
}

int ariel_init(ariel_sim* h) {
    TRY(check_handle(h));
    TRY(calibrate_max_density(h));

    // Scene::GenerateParticles at frame 0, Sort, MarkCellTypes (flip.cpp:74-76).
    std::vector<float4> pos, vel;
    seed_particles(h->d, h->N, pos, vel);
    if (pos.size() >= (size_t)INT32_MAX) return fail(h, ARIEL_ERR_UNSUPPORTED, "too many particles");
    h->np = (int64_t)pos.size();
    h->np_cap = 0;
    TRY(ensure_particles(h, std::max<int64_t>(h->np, 1)));
    h->cur = 0;
    if (h->np) {
        HIPCHK(h, hipMemcpyAsync(h->pos[0], pos.data(), h->np * sizeof(float4), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->vel[0], vel.data(), h->np * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    }
    TRY(stage_sort(h));
    TRY(stage_mark(h));
    launch_fill_f32(h->stream, h->L, h->nc, 1.6f);   // Init marks A only; L keeps its background
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->frame = 0;
    h->initialized = true;
    h->stats.particles = h->np;
    h->stats.max_density = h->max_density;
    return check_launch(h, "init");
}

int ariel_step(ariel_sim* h, int32_t nsteps) {
    TRY(check_handle(h));
    if (!h->initialized) return fail(h, ARIEL_ERR_STATE, "ariel_step before ariel_init");
    if (!(h->max_density > 0.0f)) return fail(h, ARIEL_ERR_STATE, "max density not calibrated");
    if (nsteps < 0) return fail(h, ARIEL_ERR_INVALID, "negative step count");
    for (int n = 0; n < nsteps; ++n) {
        h->frame++;
        HIPCHK(h, hipEventRecord(h->step_t0, h->stream));
        for (int st = 0; st < ARIEL_NUM_STAGES; ++st) TRY(run_stage(h, st));
        HIPCHK(h, hipEventRecord(h->step_t1, h->stream));
    }
    if (nsteps > 0) TRY(collect_stats(h));
    h->stats.frame = h->frame;
    return ARIEL_OK;
}

int ariel_run_stage(ariel_sim* h, int32_t stage) {
    TRY(check_handle(h));
    if (!h->initialized && h->np == 0 && stage != ARIEL_STAGE_SOLVE)
        return fail(h, ARIEL_ERR_STATE, "no particles: call ariel_init or ariel_set_particles");
    if (stage < 0 || stage >= ARIEL_NUM_STAGES) return fail(h, ARIEL_ERR_INVALID, "unknown stage");
    if (stage == ARIEL_STAGE_DENSITY && !(h->max_density > 0.0f))
        return fail(h, ARIEL_ERR_STATE, "max density not calibrated");
    TRY(run_stage(h, stage));
    TRY(collect_stats(h));
    return ARIEL_OK;
}

int ariel_sync(ariel_sim* h) {
    TRY(check_handle(h));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

int ariel_get_frame(ariel_sim* h, int32_t* frame) {
    TRY(check_handle(h));
    if (!frame) return fail(h, ARIEL_ERR_INVALID, "null pointer");
    *frame = h->frame;
    return ARIEL_OK;
}

int ariel_get_dimensions(ariel_sim* h, int32_t dims[3]) {
    TRY(check_handle(h));
    if (!dims) return fail(h, ARIEL_ERR_INVALID, "null pointer");
    for (int i = 0; i < 3; ++i) dims[i] = h->d.dims[i];
    return ARIEL_OK;
}

int ariel_particle_count(ariel_sim* h, int64_t* n) {
    TRY(check_handle(h));
    if (!n) return fail(h, ARIEL_ERR_INVALID, "null pointer");
    *n = h->np;
    return ARIEL_OK;
}

int ariel_get_particles(ariel_sim* h, float* pos, float* vel, float* density, uint8_t* type,
                        int64_t cap, int64_t* n) {
    TRY(check_handle(h));
    if (n) *n = h->np;
    if (!pos && !vel && !density && !type) return ARIEL_OK;
    if (cap < h->np) return fail(h, ARIEL_ERR_CAPACITY, "particle buffer too small");
    if (type) std::memset(type, ARIEL_FLUID, (size_t)h->np);
    if (!pos && !vel && !density) return ARIEL_OK;
    if (h->np == 0) return ARIEL_OK;
    const int b = h->cur, o = 1 - b;
    // scatter to id order into the idle buffers
    launch_unsort(h->stream, h->pos[b], h->vel[b], nullptr, h->np, h->pos[o], h->vel[o], nullptr, nullptr);
    TRY(check_launch(h, "unsort"));
    std::vector<float4> hp(h->np), hv(h->np);
    HIPCHK(h, hipMemcpyAsync(hp.data(), h->pos[o], h->np * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(hv.data(), h->vel[o], h->np * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int64_t i = 0; i < h->np; ++i) {
        if (pos) { pos[3 * i] = hp[i].x; pos[3 * i + 1] = hp[i].y; pos[3 * i + 2] = hp[i].z; }
        if (vel) { vel[3 * i] = hv[i].x; vel[3 * i + 1] = hv[i].y; vel[3 * i + 2] = hv[i].z; }
        if (density) density[i] = hv[i].w;
    }
    return ARIEL_OK;
}

int ariel_set_particles(ariel_sim* h, const float* pos, const float* vel, const float* density, int64_t n) {
    TRY(check_handle(h));
    if (n < 0 || (n > 0 && !pos)) return fail(h, ARIEL_ERR_INVALID, "bad particle arrays");
    if (n >= INT32_MAX) return fail(h, ARIEL_ERR_UNSUPPORTED, "too many particles");
    if (!(h->max_density > 0.0f)) TRY(calibrate_max_density(h));
    TRY(ensure_particles(h, std::max<int64_t>(n, 1)));
    std::vector<float4> hp(n), hv(n);
    for (int64_t i = 0; i < n; ++i) {
        int32_t id = (int32_t)i;
        float f;
        std::memcpy(&f, &id, 4);
        hp[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], f);
        hv[i] = vel ? make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], density ? density[i] : 10.0f)
                    : make_float4(0.f, 0.f, 0.f, density ? density[i] : 10.0f);
    }
    h->cur = 0;
    h->np = n;
    if (n) {
        HIPCHK(h, hipMemcpyAsync(h->pos[0], hp.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->vel[0], hv.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->initialized = true;
    h->stats.particles = n;
    return ARIEL_OK;
}

int ariel_get_macgrid(ariel_sim* h, int32_t which, float* dst, int64_t cap) {
    TRY(check_handle(h));
    float* src = field_ptr(h, which);
    if (!src) return fail(h, ARIEL_ERR_INVALID, "unknown field");
    const int64_t n = field_size(h, which);
    if (!dst || cap < n) return fail(h, ARIEL_ERR_CAPACITY, "field buffer too small");
    HIPCHK(h, hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

int ariel_set_macgrid(ariel_sim* h, int32_t which, const float* src, int64_t n) {
    TRY(check_handle(h));
    float* dst = field_ptr(h, which);
    if (!dst) return fail(h, ARIEL_ERR_INVALID, "unknown field");
    if (!src || n != field_size(h, which)) return fail(h, ARIEL_ERR_INVALID, "field size mismatch");
    HIPCHK(h, hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

int ariel_get_flags(ariel_sim* h, uint8_t* dst, int64_t cap) {
    TRY(check_handle(h));
    if (!dst || cap < h->nc) return fail(h, ARIEL_ERR_CAPACITY, "flag buffer too small");
    HIPCHK(h, hipMemcpyAsync(dst, h->A, h->nc, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

int ariel_set_flags(ariel_sim* h, const uint8_t* src, int64_t n) {
    TRY(check_handle(h));
    if (!src || n != h->nc) return fail(h, ARIEL_ERR_INVALID, "flag size mismatch");
    for (int64_t i = 0; i < n; ++i)
        if (src[i] > 2) return fail(h, ARIEL_ERR_INVALID, "flag values must be 0, 1 or 2");
    HIPCHK(h, hipMemcpyAsync(h->A, src, n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

int ariel_get_sort(ariel_sim* h, int32_t* ids, int64_t ids_cap, int32_t* offsets, int64_t offsets_cap) {
    TRY(check_handle(h));
    if (ids) {
        if (ids_cap < h->np) return fail(h, ARIEL_ERR_CAPACITY, "ids buffer too small");
        if (h->np) {
            launch_unsort(h->stream, h->pos[h->cur], nullptr, nullptr, h->np, nullptr, nullptr, nullptr, h->slots);
            TRY(check_launch(h, "ids"));
            HIPCHK(h, hipMemcpyAsync(ids, h->slots, h->np * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        }
    }
    if (offsets) {
        if (offsets_cap < h->nc + 1) return fail(h, ARIEL_ERR_CAPACITY, "offsets buffer too small");
        HIPCHK(h, hipMemcpyAsync(offsets, h->offsets, (h->nc + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

int ariel_get_resample_accept(ariel_sim* h, uint8_t* dst, int64_t cap) {
    TRY(check_handle(h));
    if (!dst || cap < h->np) return fail(h, ARIEL_ERR_CAPACITY, "accept buffer too small");
    if (h->np == 0) return ARIEL_OK;
    // accept[] is indexed by the slot order of the resample pass, whose ids are
    // those of the current buffer (resample keeps the slot order).
    uint8_t* tmp = nullptr;
    TRY(dalloc(h, &tmp, h->np));
    launch_unsort(h->stream, h->pos[h->cur], nullptr, h->accept, h->np, nullptr, nullptr, tmp, nullptr);
    int rc = check_launch(h, "accept");
    if (rc == ARIEL_OK) {
        hipError_t e = hipMemcpyAsync(dst, tmp, h->np, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = fail(h, ARIEL_ERR_HIP, hipGetErrorString(e));
    }
    hipFree(tmp);
    return rc;
}

int ariel_get_max_density(ariel_sim* h, float* v) {
    TRY(check_handle(h));
    if (!v) return fail(h, ARIEL_ERR_INVALID, "null pointer");
    *v = h->max_density;
    return ARIEL_OK;
}

int ariel_set_max_density(ariel_sim* h, float v) {
    TRY(check_handle(h));
    if (!(v > 0.f)) return fail(h, ARIEL_ERR_INVALID, "max density must be positive");
    h->max_density = v;
    return ARIEL_OK;
}

int ariel_get_stats(ariel_sim* h, ariel_step_stats* s) {
    TRY(check_handle(h));
    if (!s) return fail(h, ARIEL_ERR_INVALID, "null pointer");
    h->stats.frame = h->frame;
    h->stats.particles = h->np;
    h->stats.max_density = h->max_density;
    *s = h->stats;
    return ARIEL_OK;
}

int ariel_apply_laplacian(ariel_sim* h, const float* x, float* y) {
    TRY(check_handle(h));
    if (!x || !y) return fail(h, ARIEL_ERR_INVALID, "null pointer");
    HIPCHK(h, hipMemcpyAsync(h->S, x, h->nc * sizeof(float), hipMemcpyHostToDevice, h->stream));
    launch_spmv_plain(h->stream, h->params(), h->A, h->L, h->S, h->Z);
    TRY(check_launch(h, "laplacian"));
    HIPCHK(h, hipMemcpyAsync(y, h->Z, h->nc * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

int ariel_apply_preconditioner(ariel_sim* h, const float* r, float* z) {
    TRY(check_handle(h));
    if (!r || !z) return fail(h, ARIEL_ERR_INVALID, "null pointer");
    const Params p = h->params();
    HIPCHK(h, hipMemcpyAsync(h->R, r, h->nc * sizeof(float), hipMemcpyHostToDevice, h->stream));
    TRY(prepare_solve(h));
    launch_fill_f32(h->stream, h->Z, h->nc, 0.0f);
    if (h->d.precond == ARIEL_PRECOND_LEARNED) {
        if (!h->net_ready) return fail(h, ARIEL_ERR_STATE, "learned preconditioner: no weights");
        // the network's input scale reads ||r||^2 over the FLUID cells
        std::vector<uint8_t> flags(h->nc);
        HIPCHK(h, hipMemcpyAsync(flags.data(), h->A, h->nc, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        double r2 = 0.0;
        for (int64_t c = 0; c < h->nc; ++c)
            if (flags[c] == ARIEL_FLUID) r2 += (double)r[c] * (double)r[c];
        HIPCHK(h, hipMemcpyAsync(&h->sc->rnorm2, &r2, sizeof(double), hipMemcpyHostToDevice, h->stream));
        launch_diag_weight(h->stream, h->N, h->A, h->diag, h->net.wd);
        launch_learned_apply(h->stream, h->net, h->N, h->R, h->A, h->diag, h->Z, h->sc, h->act.nseg);
    } else {
        launch_precond_apply(h->stream, p, h->fl8, h->pc, h->R, h->Q, h->Z, h->reduce(), h->act);
    }
    TRY(check_launch(h, "preconditioner"));
    HIPCHK(h, hipMemcpyAsync(z, h->Z, h->nc * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

int ariel_set_profiling(ariel_sim* h, int32_t enable) {
    TRY(check_handle(h));
    h->profile = enable != 0;
    return ARIEL_OK;
}

int ariel_reset_kernel_stats(ariel_sim* h) {
    TRY(check_handle(h));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drain_timers(h);
    for (auto& k : h->kstats) k = ariel_kernel_stats{};
    return ARIEL_OK;
}

int ariel_get_kernel_stats(ariel_sim* h, int32_t kernel, ariel_kernel_stats* out) {
    TRY(check_handle(h));
    if (kernel < 0 || kernel >= ARIEL_NUM_KERNELS || !out) return fail(h, ARIEL_ERR_INVALID, "bad kernel id");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drain_timers(h);
    *out = h->kstats[kernel];
    out->bytes_per_launch = kernel_bytes(h, kernel);
    return ARIEL_OK;
}

// ------------------------------------------------------------ state files
static const char kMagic[8] = {'A', 'R', 'I', 'E', 'L', 'S', 'T', '1'};

int ariel_save_state(ariel_sim* h, const char* path) {
    TRY(check_handle(h));
    if (!path) return fail(h, ARIEL_ERR_INVALID, "null path");
    std::vector<float> pos(3 * h->np), vel(3 * h->np), rho(h->np);
    int64_t n = 0;
    TRY(ariel_get_particles(h, pos.data(), vel.data(), rho.data(), nullptr, h->np, &n));
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(h, ARIEL_ERR_IO, std::string("cannot open ") + path);
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8;
    ok = ok && std::fwrite(&h->d, sizeof(h->d), 1, f) == 1;
    ok = ok && std::fwrite(&h->frame, sizeof(int), 1, f) == 1;
    ok = ok && std::fwrite(&h->max_density, sizeof(float), 1, f) == 1;
    ok = ok && std::fwrite(&h->np, sizeof(int64_t), 1, f) == 1;
    ok = ok && std::fwrite(pos.data(), sizeof(float), pos.size(), f) == pos.size();
    ok = ok && std::fwrite(vel.data(), sizeof(float), vel.size(), f) == vel.size();
    ok = ok && std::fwrite(rho.data(), sizeof(float), rho.size(), f) == rho.size();
    std::vector<float> buf;
    for (int which = 0; ok && which <= ARIEL_FIELD_L; ++which) {
        buf.resize(field_size(h, which));
        if (ariel_get_macgrid(h, which, buf.data(), (int64_t)buf.size()) != ARIEL_OK) { ok = false; break; }
        ok = std::fwrite(buf.data(), sizeof(float), buf.size(), f) == buf.size();
    }
    std::vector<uint8_t> flags(h->nc);
    if (ok && ariel_get_flags(h, flags.data(), h->nc) != ARIEL_OK) ok = false;
    ok = ok && std::fwrite(flags.data(), 1, flags.size(), f) == flags.size();
    std::fclose(f);
    if (!ok) return fail(h, ARIEL_ERR_IO, std::string("write failed: ") + path);
    return ARIEL_OK;
}

int ariel_load_state(ariel_sim* h, const char* path) {
    TRY(check_handle(h));
    if (!path) return fail(h, ARIEL_ERR_INVALID, "null path");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(h, ARIEL_ERR_IO, std::string("cannot open ") + path);
    char magic[8];
    ariel_scene_desc d;
    int frame = 0;
    float md = 0.f;
    int64_t np = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0;
    ok = ok && std::fread(&d, sizeof(d), 1, f) == 1;
    ok = ok && std::memcmp(&d, &h->d, sizeof(d)) == 0;
    ok = ok && std::fread(&frame, sizeof(int), 1, f) == 1;
    ok = ok && std::fread(&md, sizeof(float), 1, f) == 1;
    ok = ok && std::fread(&np, sizeof(int64_t), 1, f) == 1 && np >= 0 && np < INT32_MAX;
    if (!ok) {
        std::fclose(f);
        return fail(h, ARIEL_ERR_IO, "state file header mismatch");
    }
    std::vector<float> pos(3 * np), vel(3 * np), rho(np);
    ok = std::fread(pos.data(), sizeof(float), pos.size(), f) == pos.size();
    ok = ok && std::fread(vel.data(), sizeof(float), vel.size(), f) == vel.size();
    ok = ok && std::fread(rho.data(), sizeof(float), rho.size(), f) == rho.size();
    std::vector<std::vector<float>> fields(ARIEL_FIELD_L + 1);
    for (int which = 0; ok && which <= ARIEL_FIELD_L; ++which) {
        fields[which].resize(field_size(h, which));
        ok = std::fread(fields[which].data(), sizeof(float), fields[which].size(), f) == fields[which].size();
    }
    std::vector<uint8_t> flags(h->nc);
    ok = ok && std::fread(flags.data(), 1, flags.size(), f) == flags.size();
    std::fclose(f);
    if (!ok) return fail(h, ARIEL_ERR_IO, "state file truncated");
    TRY(ariel_set_particles(h, pos.data(), vel.data(), rho.data(), np));
    for (int which = 0; which <= ARIEL_FIELD_L; ++which)
        TRY(ariel_set_macgrid(h, which, fields[which].data(), (int64_t)fields[which].size()));
    TRY(ariel_set_flags(h, flags.data(), h->nc));
    h->frame = frame;
    h->max_density = md;
    // particles are unsorted after a load: sort so that every stage may follow
    TRY(stage_sort(h));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ARIEL_OK;
}

// Narrow-band level set of the FLUID particles: spheres of radius 0.5 in grid
// units (levelset.cpp:101-111, ParticlesToLevelSet with r = 0.5 at voxel size 1,
// positions scaled by maxd) sampled on the (N+1)^3 lattice of integer points.
// phi = min over particles of |x - p| - 0.5, clamped to the band of 3 voxels.
// The particle level set of the current particles on the host ((N+1)^3 nodes,
// x-major): rasterised on the GPU (export.hip), then copied out.
constexpr float kSdfBand = 3.0f, kSdfRadius = 0.5f;
int particle_sdf(ariel_sim* h, std::vector<float>& phi) {
    const int64_t M = h->N + 1, nm = M * M * M;
    SdfScratch w;
    float* dphi = nullptr;
    int rc = ARIEL_OK;
    if ((rc = dalloc(h, &w.counts, h->nc)) || (rc = dalloc(h, &w.offsets, h->nc + 1)) ||
        (rc = dalloc(h, &w.scan_sums, h->nc / 256 + 64)) || (rc = dalloc(h, &w.cell_of, h->np)) ||
        (rc = dalloc(h, &w.rank, h->np)) || (rc = dalloc(h, &w.sorted, h->np)) || (rc = dalloc(h, &dphi, nm))) {
    } else {
        launch_particle_sdf(h->stream, h->pos[h->cur], h->np, h->N, h->maxd, kSdfBand, kSdfRadius, w, dphi);
        rc = check_launch(h, "particle level set");
        phi.resize((size_t)nm);
        hipError_t e = hipSuccess;
        if (rc == ARIEL_OK) e = hipMemcpyAsync(phi.data(), dphi, nm * sizeof(float), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (rc == ARIEL_OK && e != hipSuccess) rc = fail(h, ARIEL_ERR_HIP, std::string("level set copy: ") + hipGetErrorString(e));
    }
    for (void* q : {(void*)w.counts, (void*)w.offsets, (void*)w.scan_sums, (void*)w.cell_of, (void*)w.rank,
                    (void*)w.sorted, (void*)dphi})
        if (q) (void)hipFree(q);
    return rc;
}

int ariel_export_levelset(ariel_sim* h, const char* path) {
    TRY(check_handle(h));
    if (!path) return fail(h, ARIEL_ERR_INVALID, "null path");
    std::vector<float> phi;
    TRY(particle_sdf(h, phi));
    const int M = h->N + 1;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(h, ARIEL_ERR_IO, std::string("cannot open ") + path);
    char header[256];
    int len = std::snprintf(header, sizeof(header),
                            "ARIEL_LEVELSET 1\ndims %d %d %d\nvoxel 1.0\nbackground %.6f\nfloat32 le\n", M, M, M,
                            kSdfBand);
    bool ok = std::fwrite(header, 1, (size_t)len, f) == (size_t)len;
    ok = ok && std::fwrite(phi.data(), sizeof(float), phi.size(), f) == phi.size();
    ok = std::fclose(f) == 0 && ok;
    if (!ok) return fail(h, ARIEL_ERR_IO, std::string("write failed: ") + path);
    return ARIEL_OK;
}

int ariel_export_levelset_noalloc(ariel_sim* h, float* phi, int64_t cap) {
    TRY(check_handle(h));
    if (!phi) return fail(h, ARIEL_ERR_INVALID, "null level-set buffer");
    const int64_t M = h->N + 1;
    if (cap < M * M * M) return fail(h, ARIEL_ERR_CAPACITY, "level-set buffer too small");
    // the grid and particle limits are checked before anything is queued
    std::string err;
    const hipError_t e = noalloc_levelset(h->stream, h->device, h->pos[h->cur], h->np, h->N, h->maxd,
                                          kSdfBand, kSdfRadius, phi, err);
    if (e == hipErrorInvalidValue && !err.empty()) return fail(h, ARIEL_ERR_CAPACITY, err);
    if (e != hipSuccess) return fail(h, ARIEL_ERR_HIP, err.empty() ? hipGetErrorString(e) : err);
    return ARIEL_OK;
}

int ariel_export_mesh(ariel_sim* h, const char* path) {
    TRY(check_handle(h));
    if (!path) return fail(h, ARIEL_ERR_INVALID, "null path");
    std::vector<float> phi, verts;
    std::vector<int32_t> quads;
    TRY(particle_sdf(h, phi));
    surface_nets(phi.data(), h->N + 1, verts, quads);
    if (!write_obj(path, verts, quads)) return fail(h, ARIEL_ERR_IO, std::string("cannot write ") + path);
    return ARIEL_OK;
}

int ariel_export_particles(ariel_sim* h, const char* path) {
    TRY(check_handle(h));
    if (!path) return fail(h, ARIEL_ERR_INVALID, "null path");
    const std::string p(path);
    if (p.size() < 4 || p.compare(p.size() - 4, 4, ".geo") != 0)
        return fail(h, ARIEL_ERR_UNSUPPORTED, "particle export: only .geo (Houdini classic ASCII) is built");
    std::vector<float> pos(3 * h->np), vel(3 * h->np);
    int64_t n = 0;
    TRY(ariel_get_particles(h, pos.data(), vel.data(), nullptr, nullptr, h->np, &n));
    for (float& x : pos) x = x * h->maxd;
    if (!write_geo(path, n, pos.data(), vel.data())) return fail(h, ARIEL_ERR_IO, std::string("cannot write ") + path);
    return ARIEL_OK;
}

int ariel_mesh_sdf(const float* phi, int32_t M, const char* path, int64_t* nverts, int64_t* nquads) {
    if (!phi || !path || M < 2 || M > 4097) {
        set_last_error("ariel_mesh_sdf: bad arguments");
        return ARIEL_ERR_INVALID;
    }
    std::vector<float> verts;
    std::vector<int32_t> quads;
    surface_nets(phi, M, verts, quads);
    if (nverts) *nverts = (int64_t)(verts.size() / 3);
    if (nquads) *nquads = (int64_t)(quads.size() / 4);
    if (!write_obj(path, verts, quads)) {
        set_last_error(std::string("cannot write ") + path);
        return ARIEL_ERR_IO;
    }
    return ARIEL_OK;
}

int ariel_solve_rhs(ariel_sim* h, const float* rhs, float* x, ariel_step_stats* stats) {
    TRY(check_handle(h));
    if (!rhs || !x) return fail(h, ARIEL_ERR_INVALID, "null pointer");
    HIPCHK(h, hipMemcpyAsync(h->D, rhs, h->nc * sizeof(float), hipMemcpyHostToDevice, h->stream));
    launch_fill_f32(h->stream, h->P, h->nc, 0.0f);
    TRY(stage_solve(h));
    HIPCHK(h, hipMemcpyAsync(x, h->P, h->nc * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->solve_t0, h->solve_t1) == hipSuccess) h->stats.solve_ms = ms;
    if (stats) *stats = h->stats;
    return ARIEL_OK;
}

int ariel_set_learned_weights(ariel_sim* h, const void* blob, int64_t bytes) {
    TRY(check_handle(h));
    if (!blob || bytes < 24) return fail(h, ARIEL_ERR_INVALID, "learned weights: short blob");
    const unsigned char* b = static_cast<const unsigned char*>(blob);
    if (std::memcmp(b, "ARIELNN1", 8) != 0) return fail(h, ARIEL_ERR_INVALID, "learned weights: bad magic");
    int32_t layers;
    int64_t count;
    std::memcpy(&layers, b + 8, 4);
    std::memcpy(&count, b + 16, 8);
    int64_t off[kLearnedLayers];
    const int64_t want = learned_layout(off);
    if (layers != kLearnedLayers || count != want || bytes != 24 + 2 * count)
        return fail(h, ARIEL_ERR_INVALID, "learned weights: layout does not match this build");
    if (h->N % 16 != 0)
        return fail(h, ARIEL_ERR_UNSUPPORTED, "learned preconditioner needs N divisible by 16");
    if (!h->net_mem) {
        const int64_t n3 = h->nc;
        // weights, then e0 16n3, e1 4n3, e2 0.75n3, e3 0.125n3, p 2n3, t 16n3, d2 0.75n3,
        // d1 4n3, d0 16n3 (elements, n3 = N^3; rounded up per buffer)
        const int64_t sizes[] = {want, 16 * n3, 4 * n3, (3 * n3) / 4 + 64, n3 / 8 + 64, 2 * n3,
                                 16 * n3, (3 * n3) / 4 + 64, 4 * n3, 16 * n3};
        int64_t total = 0, pos[10];
        for (int i = 0; i < 10; ++i) {
            pos[i] = total;
            total += (sizes[i] + 127) / 128 * 128;
        }
        // all three buffers or none: a failure leaves the handle without them
        uint16_t* mem = nullptr;
        float *r_old = nullptr, *wd = nullptr;
        int rc = dalloc(h, &mem, total);
        if (rc == ARIEL_OK) rc = dalloc(h, &r_old, h->nc);
        if (rc == ARIEL_OK) rc = dalloc(h, &wd, h->nc);
        if (rc != ARIEL_OK) {
            for (void* q : {(void*)mem, (void*)r_old, (void*)wd})
                if (q) (void)hipFree(q);
            return rc;
        }
        h->net_mem = mem;
        h->R_old = r_old;
        h->net.wd = wd;
        uint16_t* m = h->net_mem;
        h->net.weights = m + pos[0];
        h->net.e0 = m + pos[1];
        h->net.e1 = m + pos[2];
        h->net.e2 = m + pos[3];
        h->net.e3 = m + pos[4];
        h->net.p = m + pos[5];
        h->net.t = m + pos[6];
        h->net.d2 = m + pos[7];
        h->net.d1 = m + pos[8];
        h->net.d0 = m + pos[9];
        for (int i = 0; i < kLearnedLayers; ++i) h->net.off[i] = off[i];
    }
    HIPCHK(h, hipMemcpyAsync(h->net_mem, b + 24, (size_t)(2 * count), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->net_ready = true;
    return ARIEL_OK;
}

int ariel_selftest_bf16(int64_t* mismatches) {
    if (!mismatches) return fail(nullptr, ARIEL_ERR_INVALID, "selftest_bf16: null pointer");
    *mismatches = -1;
    hipError_t e = run_bf16_selftest(mismatches);
    if (e != hipSuccess)
        return fail(nullptr, ARIEL_ERR_HIP, std::string("selftest_bf16: ") + hipGetErrorString(e));
    return ARIEL_OK;
}

int ariel_selftest_division(int32_t which, float a, float lo, float hi, int64_t* mismatches) {
    if (!mismatches || (which != 0 && which != 1) || !(lo >= 0.0f) || !(hi >= lo) ||
        !std::isfinite(hi) || !(a > 0.0f) || !std::isfinite(a))
        return fail(nullptr, ARIEL_ERR_INVALID, "selftest_division: bad arguments");
    *mismatches = -1;
    hipError_t e = run_division_selftest(which, a, lo, hi, mismatches);
    if (e != hipSuccess)
        return fail(nullptr, ARIEL_ERR_HIP, std::string("selftest_division: ") + hipGetErrorString(e));
    return ARIEL_OK;
}

}  // extern "C"

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
    const Reduce red = h->reduce();
    hipStream_t s = h->stream;
    launch_precond_setup(s, p, h->A, h->L, h->diag, h->nup, h->pc);
    launch_fluid_segments(s, h->N, h->A, h->fl8);
    launch_solver_init(s, h->sc, h->d.max_iters, h->d.refine_every);
    launch_active_sets(s, h->N, h->fl8, h->seg_flag, h->seg_scan, h->seg_list, h->tile_any,
                       h->tile_lists, h->scan_sums, h->sc);
    set_active(h, true);
    launch_precond_build(s, p, h->A, h->diag, h->nup, h->pc, h->act, h->sc);
    const int refine = h->d.refine_every > 0;
    launch_residual0(s, p, h->A, h->L, h->P, h->D, h->R, red, (double)h->d.tol, h->d.stop_rule,
                     h->P64, refine);
    if (refine) launch_fill_f32(s, h->P, h->nc, 0.0f);
    launch_fill_f32(s, h->Z, h->nc, 0.0f);
    launch_precond_apply(s, p, h->fl8, h->pc, h->R, h->Q, h->Z, red, h->act);
    launch_update_s(s, h->Z, h->S, red, 1, h->act);
    recorder() = nullptr;
}

void record_iteration(ariel_sim* h, Recorder& r, bool refine) {
    recorder() = &r;
    const Params p = h->params();
    const Reduce red = h->reduce();
    hipStream_t s = h->stream;
    launch_spmv_dot(s, p, h->A, h->L, h->S, h->Z, red, h->act);
    launch_axpy2(s, h->P, h->S, h->R, h->Z, red, h->act);
    if (refine) launch_reliable_update(s, p, h->A, h->L, h->P, h->P64, h->D, h->R, red, h->act);
    launch_precond_apply(s, p, h->fl8, h->pc, h->R, h->Q, h->Z, red, h->act);
    launch_update_s(s, h->Z, h->S, red, 0, h->act);
    recorder() = nullptr;
}

void record_post(ariel_sim* h, Recorder& r) {
    recorder() = &r;
    const Params p = h->params();
    const Reduce red = h->reduce();
    hipStream_t s = h->stream;
    if (h->d.refine_every > 0) launch_refine_finish(s, h->A, h->P, h->P64, h->act, h->sc);
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
                launch_p2g(h->stream, p, h->pos[h->cur], h->vel[h->cur], h->offsets, h->A, h->u, h->du);
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
    if (descs[0].precond == ARIEL_PRECOND_LEARNED)
        return bfail(nullptr, ARIEL_ERR_UNSUPPORTED,
                     "batch_create: the learned preconditioner runs on single tanks only");
    for (int32_t t = 1; t < ntanks; ++t) {
        const ariel_scene_desc& a = descs[0];
        const ariel_scene_desc& b = descs[t];
        if (a.precond != b.precond || a.splat != b.splat || a.refine_every != b.refine_every ||
            a.resample != b.resample || a.max_iters != b.max_iters)
            return bfail(nullptr, ARIEL_ERR_INVALID,
                         "batch_create: tanks must share preconditioner, splat, refinement, "
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
