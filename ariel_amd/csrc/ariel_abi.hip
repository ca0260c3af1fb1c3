// ariel_abi.hip — host side of the C ABI (include/ariel.h): device state of one
// simulation handle and the orchestration of FlipSim::Init / FlipSim::Step
// (src/sim/flip.cpp:42-128) as a sequence of HIP kernel launches on the
// handle's stream. Host work per step is launch-only apart from one small
// device->host read of the solver scalars per chunk of PCG iterations.
#include "ariel_handle.h"

using namespace ariel;
using namespace ariel::host;

thread_local std::string ariel::host::g_last_error = "";

void ariel::set_last_error(const std::string& msg) { g_last_error = msg; }

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

namespace ariel {
namespace host {

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
// output written once. The solver groups count the cells of the active tiles
// (512 each) that they sweep: the operator reads z and s_prev and writes s and
// y (16 B) and reads the FLUID bits (1 B per 8 cells, counted as 1 B per cell
// with the class reads at the free surface); the preconditioner reads r and
// pc, writes z, and writes and reads again q on the red half (20 B); the AXPY
// reads p, s, r, y and writes p, r (24 B). With the tiled red-black MIC(0) the
// AXPY rides on the preconditioner's first pass over each colour, where r is
// read once for both: that group then moves 20 + 24 - 4 = 40 B per cell.
double kernel_bytes(const ariel_sim* h, int k) {
    const double np = (double)h->np, nc = (double)h->nc;
    const double tile_cells = 512.0 * (double)h->act.ntiles[2];
    switch (k) {
        case ARIEL_KERNEL_SORT: return 60.0 * np + 8.0 * nc;
        case ARIEL_KERNEL_DENSITY: return 40.0 * np;
        case ARIEL_KERNEL_MARK: return 4.0 * np + 9.0 * nc;
        case ARIEL_KERNEL_P2G: return 24.0 * np + 12.0 * nc;
        case ARIEL_KERNEL_SPMV: return 17.0 * tile_cells;
        case ARIEL_KERNEL_PRECOND:
            // multigrid: the two level-0 passes read r, the diagonal and the
            // classes (9 B per cell) and the second writes z
            if (h->d.precond == ARIEL_PRECOND_MG) return 22.0 * tile_cells;
            return (h->d.precond == ARIEL_PRECOND_MIC0_TILED ? 40.0 : 20.0) * tile_cells;
        case ARIEL_KERNEL_AXPY: return 24.0 * tile_cells;
        case ARIEL_KERNEL_GRADIENT: return 33.0 * nc;
        case ARIEL_KERNEL_EXTRAPOLATE: return 36.0 * nc;
        case ARIEL_KERNEL_ADVECT: return 48.0 * np + 24.0 * nc;
        case ARIEL_KERNEL_RESAMPLE: return 48.0 * np;
        default: return 0.0;
    }
}

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
        d->precond != ARIEL_PRECOND_JACOBI && d->precond != ARIEL_PRECOND_LEARNED &&
        d->precond != ARIEL_PRECOND_MG) {
        msg = "unknown preconditioner"; return ARIEL_ERR_INVALID;
    }
    if ((d->precond == ARIEL_PRECOND_MIC0_TILED || d->precond == ARIEL_PRECOND_MIC0_BLOCK ||
         d->precond == ARIEL_PRECOND_MG) &&
        d->tile != 8) {
        msg = "tiled MIC(0) uses 8^3 tiles (tile must be 8)"; return ARIEL_ERR_UNSUPPORTED;
    }
    if (!(d->tol > 0.f) && d->stop_rule == ARIEL_STOP_RELATIVE) { msg = "tol must be positive"; return ARIEL_ERR_INVALID; }
    if (d->max_iters < 1) { msg = "max_iters must be >= 1"; return ARIEL_ERR_INVALID; }
    if (d->refine_every < 0) { msg = "refine_every must be >= 0"; return ARIEL_ERR_INVALID; }
    if (d->stop_rule != ARIEL_STOP_RELATIVE && d->stop_rule != ARIEL_STOP_REFERENCE_ABS) { msg = "unknown stop rule"; return ARIEL_ERR_INVALID; }
    if (d->splat != ARIEL_SPLAT_SHARPEN && d->splat != ARIEL_SPLAT_TRILINEAR) { msg = "unknown splat kernel"; return ARIEL_ERR_INVALID; }
    if (d->p2g != ARIEL_P2G_ORDERED && d->p2g != ARIEL_P2G_TILED) { msg = "unknown P2G mode"; return ARIEL_ERR_INVALID; }
    if (d->p2g == ARIEL_P2G_TILED && d->splat == ARIEL_SPLAT_SHARPEN && !(d->sharpen_re <= 1.5f)) {
        // the tiled scatter visits the window without its first cell along each
        // tangential axis, which holds only zero weights while RE <= 1.5
        msg = "tiled P2G needs a Sharpen radius <= 1.5 cells"; return ARIEL_ERR_UNSUPPORTED;
    }
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

// Scene::GenerateParticles at a step's frame (flip.cpp:98): the particles the
// scene's liquid geoms emit on it join the simulation, appended in id order
// after the existing ones (the particle arrays grow, keeping their contents).
int emit_particles(ariel_sim* h) {
    std::vector<float4> pos, vel;
    emit_geom_liquids(*h->geom, h->d, h->N, h->frame, pos, vel);
    if (pos.empty()) return ARIEL_OK;
    const int64_t add = (int64_t)pos.size(), n = h->np + add;
    if (n >= INT32_MAX) return fail(h, ARIEL_ERR_UNSUPPORTED, "too many particles");
    for (int64_t i = 0; i < add; ++i) {
        const int32_t id = (int32_t)(h->np + i);
        std::memcpy(&pos[i].w, &id, 4);
    }
    if (n > h->np_cap) {
        // keep the current particles across the reallocation
        std::vector<float4> hp(h->np), hv(h->np);
        if (h->np) {
            HIPCHK(h, hipMemcpyAsync(hp.data(), h->pos[h->cur], h->np * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipMemcpyAsync(hv.data(), h->vel[h->cur], h->np * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
        }
        TRY(ensure_particles(h, std::max<int64_t>(n, h->np_cap + h->np_cap / 2)));
        h->cur = 0;
        if (h->np) {
            HIPCHK(h, hipMemcpyAsync(h->pos[0], hp.data(), h->np * sizeof(float4), hipMemcpyHostToDevice, h->stream));
            HIPCHK(h, hipMemcpyAsync(h->vel[0], hv.data(), h->np * sizeof(float4), hipMemcpyHostToDevice, h->stream));
        }
    }
    HIPCHK(h, hipMemcpyAsync(h->pos[h->cur] + h->np, pos.data(), add * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->vel[h->cur] + h->np, vel.data(), add * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));   // the host vectors go out of scope
    h->np = n;
    h->stats.particles = n;
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
    launch_p2g(h->stream, h->params(), h->pos[h->cur], h->vel[h->cur], h->offsets, h->A, h->u, h->du,
               h->p2g_part, h->p2g_tflag);
    return check_launch(h, "p2g");
}

int stage_divergence(ariel_sim* h) {
    launch_divergence(h->stream, h->params(), h->u, h->D);
    return check_launch(h, "divergence");
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

// Coarse levels of the multigrid preconditioner, allocated on first use.
int ensure_mg(ariel_sim* h) {
    if (h->mg.type[1] || h->N <= 8) {
        if (h->mg.L == 0) mg_plan(h->N, h->mg);
        return ARIEL_OK;
    }
    mg_plan(h->N, h->mg);
    for (int l = 1; l <= h->mg.L; ++l) {
        const int64_t n = (int64_t)h->mg.M[l] * h->mg.M[l] * h->mg.M[l];
        TRY(dalloc(h, &h->mg.type[l], n));
        TRY(dalloc(h, &h->mg.d[l], n));
        TRY(dalloc(h, &h->mg.r[l], n));
        TRY(dalloc(h, &h->mg.z[l], n));
        TRY(dalloc(h, &h->mg.zo[l], n));
    }
    return ARIEL_OK;
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


}  // namespace host
}  // namespace ariel

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
    h->part_cap = partial_capacity(h->N);
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
        (rc = dalloc(h, &h->Rt, h->nc)) || (rc = dalloc(h, &h->Zt, h->nc)) || (rc = dalloc(h, &h->St, 2 * h->nc)) ||
        (rc = dalloc(h, &h->Yt, h->nc)) || (rc = dalloc(h, &h->Pt, h->nc)) || (rc = dalloc(h, &h->Gt, h->nc)) ||
        (rc = dalloc(h, &h->sb, h->nc / 8)) ||
        (rc = dalloc(h, &h->diag, h->nc)) || (rc = dalloc(h, &h->nup, h->nc)) ||
        (rc = dalloc(h, &h->fl8, h->nc / 8)) ||
        (rc = dalloc(h, &h->counts, h->nc)) || (rc = dalloc(h, &h->offsets, h->nc + 1)) ||
        (rc = dalloc(h, &h->scan_sums, h->nc / 4096 + 4)) ||
        (rc = dalloc(h, &h->P64, h->d.refine_every > 0 ? h->nc : 1)) ||
        (rc = dalloc(h, &h->seg_flag, h->nc / 8)) || (rc = dalloc(h, &h->seg_scan, h->nc / 8 + 1)) ||
        (rc = dalloc(h, &h->seg_list, h->nc / 8)) || (rc = dalloc(h, &h->tile_any, h->nc / 512)) ||
        (rc = dalloc(h, &h->tile_lists, 3 * (h->nc / 512))) ||
        (rc = dalloc(h, &h->partials, 3 * h->part_cap)) || (rc = dalloc(h, &h->counter, kCounterWords)) ||
        (rc = dalloc(h, &h->sc, 1)))
        return bail(rc);
    if (h->d.precond == ARIEL_PRECOND_MG && (rc = ensure_mg(h)) != ARIEL_OK) return bail(rc);
    if (h->d.p2g == ARIEL_P2G_TILED) {
        const int64_t ntiles = p2g_tile_count(h->N);
        if ((rc = dalloc(h, &h->p2g_part, ntiles * kP2gTilePart)) || (rc = dalloc(h, &h->p2g_tflag, ntiles)))
            return bail(rc);
    }
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
    for (float* v : {h->R, h->Z, h->S, h->Q, h->diag, h->Rt, h->Zt, h->Yt, h->Pt}) launch_fill_f32(s, v, h->nc, 0.0f);
    launch_fill_f32(s, h->St, 2 * h->nc, 0.0f);
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
                    (void*)h->Rt, (void*)h->Zt, (void*)h->St, (void*)h->Yt, (void*)h->Pt, (void*)h->Gt, (void*)h->sb,
                    (void*)h->diag, (void*)h->nup, (void*)h->fl8, (void*)h->P64, (void*)h->partials, (void*)h->counter, (void*)h->sc,
                    (void*)h->seg_flag, (void*)h->seg_scan, (void*)h->seg_list, (void*)h->tile_any,
                    (void*)h->tile_lists, (void*)h->p2g_part, (void*)h->p2g_tflag, (void*)h->net_mem, (void*)h->R_old, (void*)h->net.wd})
        hipFree(p);
    for (int c = 0; c < 3; ++c) {
        hipFree(h->u[c]);
        hipFree(h->du[c]);
    }
    for (int l = 1; l < kMgMaxLevels; ++l)
        for (void* p : {(void*)h->mg.type[l], (void*)h->mg.d[l], (void*)h->mg.r[l], (void*)h->mg.z[l],
                        (void*)h->mg.zo[l]})
            hipFree(p);
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

// The solver fields of `solver` over the scene's settings and forces.
void take_solver(ariel_scene_desc& d, const ariel_scene_desc* solver) {
    if (!solver) return;
    d.picflip = solver->picflip;
    d.wall = solver->wall;
    d.seed = solver->seed;
    d.jitter = solver->jitter;
    d.precond = solver->precond;
    d.mic_tau = solver->mic_tau;
    d.mic_sigma = solver->mic_sigma;
    d.tile = solver->tile;
    d.tol = solver->tol;
    d.max_iters = solver->max_iters;
    d.refine_every = solver->refine_every;
    d.stop_rule = solver->stop_rule;
    d.splat = solver->splat;
    d.resample = solver->resample;
    d.sharpen_re = solver->sharpen_re;
    d.spring = solver->spring;
    d.p2g = solver->p2g;
}

int ariel_create_from_json(const char* json, const char* base_dir, const ariel_scene_desc* solver,
                           ariel_sim** out) {
    if (!out || !json) return fail(nullptr, ARIEL_ERR_INVALID, "null argument");
    *out = nullptr;
    ariel_scene_desc d;
    auto geom = std::make_unique<SceneGeom>();
    std::string err;
    const int rc = scene_from_json(json, base_dir, &d, geom.get(), err);
    if (rc != ARIEL_OK) return fail(nullptr, rc, "create_from_json: " + err);
    take_solver(d, solver);
    TRY(ariel_create(&d, out));
    (*out)->geom = std::move(geom);
    return ARIEL_OK;
}

int ariel_get_desc(ariel_sim* h, ariel_scene_desc* out) {
    TRY(check_handle(h));
    if (!out) return fail(h, ARIEL_ERR_INVALID, "null pointer");
    *out = h->d;
    return ARIEL_OK;
}

int ariel_scene_emit(const char* json, const char* base_dir, const ariel_scene_desc* solver,
                     int32_t frame, float* pos, float* vel, int64_t cap, int64_t* n) {
    if (!json || !n) return fail(nullptr, ARIEL_ERR_INVALID, "null argument");
    ariel_scene_desc d;
    SceneGeom geom;
    std::string err;
    const int rc = scene_from_json(json, base_dir, &d, &geom, err);
    if (rc != ARIEL_OK) return fail(nullptr, rc, "scene_emit: " + err);
    take_solver(d, solver);
    std::string msg;
    const int vr = validate(&d, msg);
    if (vr != ARIEL_OK) return fail(nullptr, vr, msg);
    std::vector<float4> p, v;
    emit_geom_liquids(geom, d, d.dims[0], frame, p, v);
    *n = (int64_t)p.size();
    if (!pos && !vel) return ARIEL_OK;
    if (cap < *n) return fail(nullptr, ARIEL_ERR_CAPACITY, "particle buffer too small");
    for (size_t i = 0; i < p.size(); ++i) {
        if (pos) { pos[3 * i] = p[i].x; pos[3 * i + 1] = p[i].y; pos[3 * i + 2] = p[i].z; }
        if (vel) { vel[3 * i] = v[i].x; vel[3 * i + 1] = v[i].y; vel[3 * i + 2] = v[i].z; }
    }
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
    if (h->geom) {
        const size_t n0 = pos.size();
        emit_geom_liquids(*h->geom, h->d, h->N, 0, pos, vel);
        for (size_t i = n0; i < pos.size(); ++i) {
            const int32_t id = (int32_t)i;
            std::memcpy(&pos[i].w, &id, 4);
        }
    }
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
        if (h->geom) TRY(emit_particles(h));
        HIPCHK(h, hipEventRecord(h->step_t0, h->stream));
        for (int st = 0; st < ARIEL_NUM_STAGES; ++st) TRY(run_stage(h, st));
        HIPCHK(h, hipEventRecord(h->step_t1, h->stream));
    }
    if (nsteps > 0) TRY(collect_stats(h));
    h->stats.frame = h->frame;
    return ARIEL_OK;
}

int ariel_step_autotune(ariel_sim* h, const int32_t* candidates, int32_t n, ariel_solver_trial* trials,
                        int32_t* chosen) {
    TRY(check_handle(h));
    if (!h->initialized) return fail(h, ARIEL_ERR_STATE, "ariel_step_autotune before ariel_init");
    if (!(h->max_density > 0.0f)) return fail(h, ARIEL_ERR_STATE, "max density not calibrated");
    if (!candidates || n < 1 || n > 8) return fail(h, ARIEL_ERR_INVALID, "autotune: 1 to 8 candidates");
    for (int i = 0; i < n; ++i) {
        const int c = candidates[i];
        if (c != ARIEL_PRECOND_MIC0_TILED && c != ARIEL_PRECOND_MIC0_BLOCK && c != ARIEL_PRECOND_JACOBI &&
            c != ARIEL_PRECOND_MG)
            return fail(h, ARIEL_ERR_INVALID, "autotune: candidates are MIC0_TILED, MIC0_BLOCK, JACOBI or MG");
        if (c != ARIEL_PRECOND_JACOBI && h->d.tile != 8)
            return fail(h, ARIEL_ERR_UNSUPPORTED, "autotune: tiled preconditioners need tile 8");
        if (c == ARIEL_PRECOND_MG) TRY(ensure_mg(h));
    }
    h->frame++;
    if (h->geom) TRY(emit_particles(h));
    HIPCHK(h, hipEventRecord(h->step_t0, h->stream));
    for (int st = 0; st < ARIEL_STAGE_SOLVE; ++st) TRY(run_stage(h, st));
    // every trial starts from the step's warm start, kept in the staging buffer S
    // (no stage of a step uses it)
    const size_t bytes = (size_t)h->nc * sizeof(float);
    HIPCHK(h, hipMemcpyAsync(h->S, h->P, bytes, hipMemcpyDeviceToDevice, h->stream));
    const int saved_iters = h->last_iters;
    int best = -1;
    float best_ms = 0.0f;
    for (int i = 0; i < n; ++i) {
        h->d.precond = candidates[i];
        ariel_solver_trial tr{candidates[i], 0, 0, 0.0f};
        for (int rep = 0; rep < 2; ++rep) {
            HIPCHK(h, hipMemcpyAsync(h->P, h->S, bytes, hipMemcpyDeviceToDevice, h->stream));
            h->last_iters = saved_iters;
            TRY(stage_solve(h));
            TRY(collect_stats(h));
            tr.iterations = h->stats.iterations;
            tr.status = h->stats.solve_status;
            tr.solve_ms = (float)h->stats.solve_ms;
        }
        if (trials) trials[i] = tr;
        if (tr.status == ARIEL_SOLVE_CONVERGED && (best < 0 || tr.solve_ms < best_ms)) {
            best = i;
            best_ms = tr.solve_ms;
        }
    }
    if (best < 0) best = 0;
    if (chosen) *chosen = best;
    h->d.precond = candidates[best];
    HIPCHK(h, hipMemcpyAsync(h->P, h->S, bytes, hipMemcpyDeviceToDevice, h->stream));
    h->last_iters = saved_iters;
    for (int st = ARIEL_STAGE_SOLVE; st < ARIEL_NUM_STAGES; ++st) TRY(run_stage(h, st));
    HIPCHK(h, hipEventRecord(h->step_t1, h->stream));
    TRY(collect_stats(h));
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
    TRY(prepare_solve(h));
    HIPCHK(h, hipMemcpyAsync(h->R, r, h->nc * sizeof(float), hipMemcpyHostToDevice, h->stream));
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
        // the apply works on the tile-major vectors over the active tiles; cells
        // outside them are not FLUID and keep z = 0
        launch_tile_convert(h->stream, h->N, h->R, h->Rt, h->act, h->sc, true, false);
        precond_apply(h, h->reduce());
        launch_tile_convert(h->stream, h->N, h->Zt, h->Z, h->act, h->sc, false, false);
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
constexpr float kTrailRmin = 0.01f;   // raster.setRmin(.01f), levelset.cpp:104
int particle_sdf(ariel_sim* h, std::vector<float>& phi) {
    const int64_t M = h->N + 1, nm = M * M * M;
    SdfScratch w;
    float* dphi = nullptr;
    int rc = ARIEL_OK;
    if ((rc = dalloc(h, &w.counts, h->nc)) || (rc = dalloc(h, &w.offsets, h->nc + 1)) ||
        (rc = dalloc(h, &w.scan_sums, h->nc / 256 + 64)) || (rc = dalloc(h, &w.cell_of, h->np)) ||
        (rc = dalloc(h, &w.rank, h->np)) || (rc = dalloc(h, &w.sorted, h->np)) || (rc = dalloc(h, &dphi, nm))) {
    } else {
        if (h->levelset_mode == ARIEL_LEVELSET_TRAILS)
            launch_trail_sdf(h->stream, h->pos[h->cur], h->vel[h->cur], h->np, h->N, h->maxd, kSdfBand,
                             kSdfRadius, kTrailRmin, dphi);
        else
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

int ariel_set_levelset_mode(ariel_sim* h, int32_t mode) {
    TRY(check_handle(h));
    if (mode != ARIEL_LEVELSET_SPHERES && mode != ARIEL_LEVELSET_TRAILS)
        return fail(h, ARIEL_ERR_INVALID, "unknown level-set mode");
    h->levelset_mode = mode;
    return ARIEL_OK;
}

int ariel_export_levelset(ariel_sim* h, const char* path) {
    TRY(check_handle(h));
    if (!path) return fail(h, ARIEL_ERR_INVALID, "null path");
    std::vector<float> phi;
    TRY(particle_sdf(h, phi));
    const int M = h->N + 1;
    const std::string ps(path);
    if (ps.size() >= 4 && ps.compare(ps.size() - 4, 4, ".vdb") == 0) {
        std::string err;
        if (!write_vdb(path, phi.data(), M, kSdfBand, "surface", err)) return fail(h, ARIEL_ERR_IO, err);
        return ARIEL_OK;
    }
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
    if (partio_format(path) == 0)
        return fail(h, ARIEL_ERR_UNSUPPORTED, "particle export: the extension must be .geo or .bgeo (optionally .gz)");
    std::string err;
    std::vector<float> pos(3 * h->np), vel(3 * h->np);
    int64_t n = 0;
    TRY(ariel_get_particles(h, pos.data(), vel.data(), nullptr, nullptr, h->np, &n));
    for (float& x : pos) x = x * h->maxd;
    const int rc = write_partio(path, n, pos.data(), vel.data(), err);
    if (rc != ARIEL_OK) return fail(h, rc, err);
    return ARIEL_OK;
}

int ariel_write_vdb(const float* phi, int32_t M, float bg, const char* path) {
    if (!phi || !path || M < 1 || M > 4096 || !(bg > 0.0f)) {
        set_last_error("ariel_write_vdb: bad arguments");
        return ARIEL_ERR_INVALID;
    }
    std::string err;
    if (!write_vdb(path, phi, M, bg, "surface", err)) {
        set_last_error(err);
        return ARIEL_ERR_IO;
    }
    return ARIEL_OK;
}

int ariel_write_particles(const char* path, int64_t n, const float* pos, const float* vel) {
    if (!path || n < 0 || (n > 0 && (!pos || !vel))) {
        set_last_error("ariel_write_particles: bad arguments");
        return ARIEL_ERR_INVALID;
    }
    std::string err;
    const int rc = write_partio(path, n, pos, vel, err);
    if (rc != ARIEL_OK) set_last_error(err);
    return rc;
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
    if (!mismatches || (which < 0 || which > 2) || !(lo >= 0.0f) || !(hi >= lo) ||
        !std::isfinite(hi) || !(a > 0.0f) || !std::isfinite(a))
        return fail(nullptr, ARIEL_ERR_INVALID, "selftest_division: bad arguments");
    *mismatches = -1;
    hipError_t e = run_division_selftest(which, a, lo, hi, mismatches);
    if (e != hipSuccess)
        return fail(nullptr, ARIEL_ERR_HIP, std::string("selftest_division: ") + hipGetErrorString(e));
    return ARIEL_OK;
}

}  // extern "C"
