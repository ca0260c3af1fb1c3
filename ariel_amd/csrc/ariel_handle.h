// ariel_handle.h — host-side state of one simulation handle (the opaque
// ariel_sim of include/ariel.h) and the helpers shared by the host files:
// ariel_abi.hip (handle lifetime, stages of FlipSim::Step, accessors, state and
// export entry points), solve.hip (the pressure solve's orchestration) and
// batch.hip (independent tanks stepped together).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ariel.h"
#include "ariel_device.h"
#include "ariel_internal.h"

namespace ariel {
namespace host {

struct Stage {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    bool recorded = false;
};

struct KTimer {
    int kernel;
    hipEvent_t e0, e1;
};

// Message of the last handle-less failure on the calling thread.
extern thread_local std::string g_last_error;

}  // namespace host
}  // namespace ariel

// Internal header of the library's host files: the device-side types of
// ariel_internal.h are used unqualified below.
using namespace ariel;

struct ariel_sim {
    ariel_scene_desc d{};
    // the reference scene's geometry (ariel_create_from_json): liquid geoms
    // that emit particles on the frames they are in
    std::unique_ptr<ariel::SceneGeom> geom;
    int N = 0;
    int64_t nc = 0, nf = 0, np = 0;
    float maxd = 0.f;
    float max_density = 0.f;
    int frame = 0;
    bool initialized = false;
    int last_iters = 0;
    int p2g_tile = 8;             // P2G face-column tile edge (autotuned in batches)
    int levelset_mode = ARIEL_LEVELSET_SPHERES;   // export level set: spheres or trails
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
    // ARIEL_P2G_TILED: per-tile face partials and the tiles' particle flags
    float* p2g_part = nullptr;
    uint8_t* p2g_tflag = nullptr;
    // pc: 2 N^3 floats, the factor in the natural layout and then the tile-major
    // copy the preconditioner applies read. R, Z, S: natural-layout staging of
    // the accessors and of the learned network's input and output. The solve's
    // iteration vectors are tile-major: Rt, Zt, St (two buffers), Yt = A s, the
    // iterate Pt, and Q, the MIC(0) apply's intermediate.
    float *R = nullptr, *Z = nullptr, *S = nullptr, *Q = nullptr, *pc = nullptr, *diag = nullptr;
    float *Rt = nullptr, *Zt = nullptr, *St = nullptr, *Yt = nullptr, *Pt = nullptr;
    float* Gt = nullptr;   // ghost level set of the solve (L, NaN at SOLID), tile-major
    uint8_t* sb = nullptr;  // SOLID bits per tile-major 8-cell line
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
    float* R_old = nullptr;   // tile-major
    bool net_ready = false;
    MgLevels mg;                  // coarse levels of ARIEL_PRECOND_MG
    double* partials = nullptr;   // 3 * part_cap: two-value partials, then the SpMV's
    int64_t part_cap = 0;
    uint32_t* counter = nullptr;
    SolverScalars* sc = nullptr;
    SolverScalars* sc_host = nullptr;   // pinned

    ariel::host::Stage stages[ARIEL_NUM_STAGES];
    bool profile = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<ariel::host::KTimer> pending;
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
        p.p2g = d.p2g;
        return p;
    }
    Reduce reduce() const { return Reduce{partials, counter, sc, partials + 2 * part_cap}; }
};

namespace ariel {
namespace host {

inline int fail(ariel_sim* h, int code, const std::string& msg) {
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

#define TRY(expr)                       \
    do {                                \
        int rc_ = (expr);               \
        if (rc_ != ARIEL_OK) return rc_; \
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

int check_launch(ariel_sim* h, const char* what);
int check_handle(ariel_sim* h);
hipEvent_t take_event(ariel_sim* h);
double kernel_bytes(const ariel_sim* h, int k);
void drain_timers(ariel_sim* h);
int read_scalars(ariel_sim* h);

// Each event record leaves the GPU idle for a few microseconds between the
// kernels it separates, so inside the PCG loop only one iteration in
// kProfEvery is timed (the launches do not depend on the iteration index).
constexpr int kProfEvery = 16;
inline bool prof_sampled(int it) { return (it - 1) % kProfEvery == 0; }

// Times the launches queued during its lifetime as one group of kernel k
// (events on the handle's stream), when profiling is on and `on` is set.
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

// stages of FlipSim::Step (ariel_abi.hip, solve.hip)
int stage_sort(ariel_sim* h);
int stage_density(ariel_sim* h, int forces);
int stage_mark(ariel_sim* h);
int stage_p2g(ariel_sim* h);
int stage_divergence(ariel_sim* h);
int stage_solve(ariel_sim* h);
int stage_gradient(ariel_sim* h);
int stage_extrapolate(ariel_sim* h);
int stage_advect(ariel_sim* h);
int stage_resample(ariel_sim* h);

// pressure solve helpers (solve.hip)
void set_active(ariel_sim* h, bool device_counts);
int prepare_solve(ariel_sim* h);
// The solve's launches after prepare_solve (solve.hip): set-up, one iteration
// (refine: with the reliable update; learned: flexible PCG, mic_only once the
// host has seen the fallback; prof: sampled kernel timing), and the pressure
// from the iterate.
void enqueue_solve_setup(ariel_sim* h, bool learned);
void enqueue_iteration(ariel_sim* h, int it, bool refine, bool learned, bool mic_only, bool prof);
void enqueue_solve_finish(ariel_sim* h);
// z = M^-1 r on the tile-major vectors with z.r and beta (every preconditioner
// but the learned one)
void precond_apply(ariel_sim* h, const Reduce& red);

}  // namespace host
}  // namespace ariel
