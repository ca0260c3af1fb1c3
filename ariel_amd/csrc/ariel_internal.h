// ariel_internal.h — device state and kernel launchers of the FLIP step.
//
// Layout in HBM (one simulation handle, N = grid edge, Np particles):
//   particles, sorted by (cell, id), ping-pong pair:
//     pos[b] : float4 (x, y, z, id as int bits)          16 B/particle
//     vel[b] : float4 (ux, uy, uz, density)              16 B/particle
//   sort: cell_of[Np], slots[Np] (int32), counts[N^3] (int32, zero between
//         sorts), offsets[N^3+1] (int32)
//   MAC grid: u[3], du[3] (faces, float), P, D, L (cells, float), A (cells, u8)
//   solver: diag, pc (natural, then its tile-major copy), nup (u8), fl8 (FLUID
//         bits per 8-cell z-segment); the iteration vectors r, z, s (two
//         buffers), y = A s, the iterate p and q, tile-major (tile t's 512 cells
//         at [512 t, 512 t + 512), cell (il, jl, kl) at 64 il + 8 jl + kl);
//         x64 (double, natural) for reliable updates; reduction partials (double)
#pragma once

#include <array>
#include <type_traits>
#include <algorithm>

#include <hip/hip_runtime.h>

#include "../../include/ariel.h"
#include <stdint.h>

#include <cstring>
#include <string>
#include <vector>

namespace ariel {

constexpr int kAir = 0, kFluid = 1, kSolid = 2;
constexpr int kBlock = 256;      // default threads per block of streaming kernels
constexpr int kMaxPartials = 65536;
constexpr int kCounterWords = 33 * 32;   // arrival counters of the reductions

// Reduction/solver scalars living in device memory. Written by the one-block
// finalize kernel of each reduction, read by later kernels; no host round trip
// inside a PCG iteration.
struct SolverScalars {
    double bnorm2;      // ||b||^2 over fluid cells
    double rnorm2;      // ||r||^2
    double thr;         // stop threshold on rnorm2
    double rho;         // z.r
    double sigma;       // s.As
    double alpha;       // rho / sigma
    double beta;        // rho_new / rho
    double tmp;         // scratch
    int32_t done;       // 1 once converged / failed / out of iterations
    int32_t it;         // completed iterations
    int32_t status;     // ARIEL_SOLVE_*
    int32_t max_iters;
    int32_t refine_every;
    int32_t fluid_cells;
    uint32_t maxdiv_bits;  // max |div| as float bits (non-negative floats order as ints)
    int32_t pad;
    int64_t nseg;          // active z-segments (counted by the active-set build)
    int32_t ntile[3];      // active tiles: red, black, all
    int32_t flex;          // flexible PCG (learned preconditioner)
    int32_t use_mic;       // flexible PCG has fallen back to MIC(0)
    int32_t restart;       // next search direction restarts (beta = 0)
    double rnorm2_prev;    // ||r||^2 at the start of the fallback test's window
    int32_t fallback_it;   // iteration at which flexible PCG switched to MIC(0)
    int32_t fallback_reason;  // 1: slow progress, 2: not a descent direction
    double rnorm2_0;       // ||r||^2 of the solve's initial residual
    double tmp64;          // ||b - A x64||^2 of the FP64 solution (report)
};

// Cells the pressure solve touches. A z-segment is 8 cells (cells 8s..8s+7) and
// is active when one of them is FLUID; an 8^3 tile is active when one of its 64
// segments is. Lists are ascending, so block partials keep a fixed order.
// Everything outside them stays as the solve's full-grid set-up left it
// (zero off-FLUID), which the operator, the preconditioner halos and the dot
// products never read as FLUID values.
// A count of -1 means "read it from the device scalars" (batched launches);
// the caps are the largest possible counts, used to size such launches.
struct Active {
    const int32_t* segs = nullptr;
    int64_t nseg = 0;
    const int32_t* tiles[3] = {nullptr, nullptr, nullptr};   // red, black, all
    int64_t ntiles[3] = {0, 0, 0};
    int64_t seg_cap = 0;
    int64_t tile_cap[3] = {0, 0, 0};
};

struct Params {
    int N;
    float maxd;
    float density;     // lattice spacing in cells
    float dt;
    float picflip;
    float sharpen_re;
    float spring;
    int splat;
    int wall;
    uint32_t seed;
    int n_forces;
    float forces[4][3];
    float max_density;
    int frame;
    float mic_tau, mic_sigma;
    int precond;
    int p2g_tile;      // edge of the face-column tiles that order P2G blocks
    int p2g;           // ARIEL_P2G_* (ordered gather or tiled scatter)
};

// Reduction workspace: per-block double partials + an arrival counter.
struct Reduce {
    double* partials;      // 2 * partial_capacity(N): one or two values per block
    uint32_t* counter;     // 33 arrival counters, 128 bytes apart (zero between reductions)
    SolverScalars* sc;     // device scalars
    double* sig;           // partial_capacity(N): the SpMV's s.As per block, apart from
                           // `partials` so the first MIC(0) pass can sum them while it
                           // writes its own
};
// Blocks a reducing solver kernel may launch on an N^3 grid (one partial each).
int64_t partial_capacity(int N);

// Coarse levels of the multigrid preconditioner (ARIEL_PRECOND_MG): level l >= 1
// is a dense grid of edge M[l] = ceil(M[l-1] / 2), M[0] = N, down to the
// coarsest, L, of edge <= 8 (L = 0: no coarse level).
constexpr int kMgMaxLevels = 12;
struct MgLevels {
    int L = 0;
    int M[kMgMaxLevels] = {};
    uint8_t* type[kMgMaxLevels] = {};
    float* d[kMgMaxLevels] = {};
    float* r[kMgMaxLevels] = {};
    float* z[kMgMaxLevels] = {};
    float* zo[kMgMaxLevels] = {};
};
// edges of the levels for grid edge N (fills L and M)
void mg_plan(int N, MgLevels& mg);
// coarse cell classes and diagonals from the solve's classes A (once per solve)
void launch_mg_build(hipStream_t s, int N, const uint8_t* A, const MgLevels& mg);
// z = B r (one V-cycle) on the active tiles, tile-major r and z; z.r into the
// partials and FIN_RHO (beta)
void launch_mg_apply(hipStream_t s, int N, const uint8_t* A, const float* diag, const MgLevels& mg,
                     const float* Rt, float* Zt, const Reduce& red, const Active& act);

// Exhaustive division self-check (selftest.hip); lo, hi non-negative.
hipError_t run_bf16_selftest(int64_t* mismatches);
hipError_t run_division_selftest(int which, float a, float lo, float hi, int64_t* mismatches);

// Learned preconditioner (learned.hip): packed BF16 weights and activation
// buffers of the 4-level network.
constexpr int kLearnedLayers = 11;
struct LearnedNet {
    const uint16_t* weights = nullptr;
    int64_t off[kLearnedLayers] = {};
    uint16_t *e0 = nullptr, *e1 = nullptr, *e2 = nullptr, *e3 = nullptr;
    uint16_t *p = nullptr, *t = nullptr, *d2 = nullptr, *d1 = nullptr, *d0 = nullptr;
    float* wd = nullptr;   // sqrt(6/diag) on FLUID cells, 0 elsewhere (per solve)
};
// Offsets of the layers in the packed weights (BF16 elements); returns the total.
int64_t learned_layout(int64_t* off);
// z = net(r) on the FLUID cells (0 elsewhere); skipped when sc says done or
// fallen back to MIC(0). nseg = active segments (-1: from sc) for the scale.
void launch_learned_apply(hipStream_t s, const LearnedNet& net, int N, const float* r,
                          const uint8_t* A, const float* diag, float* z, const SolverScalars* sc,
                          int64_t nseg);
// The calling thread's last error message (ariel_last_error(NULL)).
void set_last_error(const std::string& msg);

// net.wd from the cell classes and the diagonal of the current matrix.
void launch_diag_weight(hipStream_t s, int N, const uint8_t* A, const float* diag, float* wd);

// ---------------------------------------------------------------- launching
// Every kernel is a functor K { static constexpr int kThreads; struct Args;
// static __device__ void run(const Args&); }. For one simulation, launch<K>
// passes the Args by value (k_one<K>). A batch of independent tanks runs the
// same launch sequence once for all of them: k_many<K> is launched with
// gridDim.y = number of tanks and reads tank blockIdx.y's Args from a device
// table. While a Recorder is installed on the calling thread, launch<K> records
// the launch (kernel, geometry, Args bytes) instead of issuing it.
struct LaunchRecord {
    const void* many = nullptr;   // host stub of k_many<K>
    dim3 grid, block;
    std::vector<unsigned char> args;
};
struct Recorder {
    std::vector<LaunchRecord> seq;
};
Recorder*& recorder();        // thread-local (ariel_abi.hip)
// Batch sizing (thread-local): launchers give reductions and active-set loops
// grids that do not depend on the tank, and the kernels take the active-set
// counts from the tank's device scalars, so one recorded sequence serves every
// tank and every step.
bool& batch_sizing();

// A kernel may ask for a minimum number of resident waves per SIMD
// (K::kMinWavesPerEU), which caps the registers the compiler gives each lane.
template <class K, class = void>
struct min_waves_per_eu {
    static constexpr int value = 1;
};
template <class K>
struct min_waves_per_eu<K, std::void_t<decltype(K::kMinWavesPerEU)>> {
    static constexpr int value = K::kMinWavesPerEU;
};

template <class K>
__global__ __launch_bounds__(K::kThreads, min_waves_per_eu<K>::value) void k_one(typename K::Args a) {
    K::run(a);
}
template <class K>
__global__ __launch_bounds__(K::kThreads, min_waves_per_eu<K>::value) void k_many(
    const typename K::Args* __restrict__ t) {
    K::run(t[blockIdx.y]);
}

template <class K>
inline void launch(hipStream_t s, dim3 grid, dim3 block, const typename K::Args& a) {
    if (Recorder* r = recorder()) {
        LaunchRecord rec;
        rec.many = reinterpret_cast<const void*>(&k_many<K>);
        rec.grid = grid;
        rec.block = block;
        rec.args.resize(sizeof(a));
        std::memcpy(rec.args.data(), &a, sizeof(a));
        r->seq.push_back(std::move(rec));
        return;
    }
    hipLaunchKernelGGL(k_one<K>, grid, block, 0, s, a);
}

// ---------------------------------------------------------------- launchers
// particles (kernels_particles.hip)
void launch_bin(hipStream_t s, const float4* pos, int64_t np, int N, float maxd, int32_t* cell_of,
                int32_t* counts);
void launch_scan(hipStream_t s, const int32_t* in, int32_t* out, int64_t n, int32_t* block_sums);
void launch_scatter(hipStream_t s, const int32_t* cell_of, int64_t np, const int32_t* offsets,
                    int32_t* counts, int32_t* slots, const float4* pos, int32_t* sid);
void launch_cellsort(hipStream_t s, const int32_t* offsets, int64_t nc, int32_t* sid,
                     int32_t* slots);
void launch_gather(hipStream_t s, const int32_t* slots, int64_t np, const float4* pos_in,
                   const float4* vel_in, float4* pos_out, float4* vel_out);
void launch_density(hipStream_t s, const Params& p, const float4* pos, float4* vel, int64_t np,
                    const int32_t* offsets, int apply_forces);
void launch_mark(hipStream_t s, const Params& p, const float4* vel, const int32_t* offsets,
                 uint8_t* A, float* L);
// ARIEL_P2G_TILED keeps, per 8^3 cell tile, the (weight, weighted velocity) sums of
// every face its cells reach: 3 components x 900 faces x 2 floats, and one flag
// byte per tile (0: the tile holds no particle and wrote no sums).
constexpr int kP2gTileEdge = 8;
constexpr int64_t kP2gTileFaces = 900;
constexpr int64_t kP2gTilePart = 3 * 2 * kP2gTileFaces;
inline int64_t p2g_tile_count(int N) {
    const int64_t t = (N + kP2gTileEdge - 1) / kP2gTileEdge;
    return t * t * t;
}
void launch_p2g(hipStream_t s, const Params& p, const float4* pos, const float4* vel,
                const int32_t* offsets, const uint8_t* A, float* const u[3], float* const du[3],
                float* part, uint8_t* tflag);
void launch_g2p_advect(hipStream_t s, const Params& p, float4* pos, float4* vel, int64_t np,
                       float* const u[3], float* const du[3]);
void launch_resample(hipStream_t s, const Params& p, const float4* pos, const float4* vel,
                     int64_t np, const int32_t* offsets, float4* pos_out, float4* vel_out,
                     uint8_t* accept);
void launch_unsort(hipStream_t s, const float4* pos, const float4* vel, const uint8_t* accept,
                   int64_t np, float4* pos_out, float4* vel_out, uint8_t* accept_out,
                   int32_t* ids_out);
void launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v);

// frame export (export.hip)
struct SdfScratch {
    int32_t* counts = nullptr;     // N^3
    int32_t* offsets = nullptr;    // N^3 + 1
    int32_t* scan_sums = nullptr;  // block sums of launch_scan over N^3
    int32_t* cell_of = nullptr;    // np
    int32_t* rank = nullptr;       // np
    float4* sorted = nullptr;      // np
};
// phi of the particle level set on the (N+1)^3 nodes (grid units).
void launch_particle_sdf(hipStream_t s, const float4* pos, int64_t np, int N, float maxd, float band,
                         float rad, SdfScratch& w, float* phi);
// phi of the particles' velocity trails on the (N+1)^3 nodes (grid units).
void launch_trail_sdf(hipStream_t s, const float4* pos, const float4* vel, int64_t np, int N, float maxd,
                      float band, float rad, float rmin, float* phi);
// Dual mesh of phi's zero surface on an M^3 node grid: vertices (x, y, z in node
// units) and quads (4 vertex indices each, normals out of the negative side).
void surface_nets(const float* phi, int M, std::vector<float>& verts, std::vector<int32_t>& quads);
bool write_obj(const char* path, const std::vector<float>& verts, const std::vector<int32_t>& quads);
// The same level set with its workspace in a fixed device array (noalloc.hip):
// no device allocation. Fails (err set, hipErrorInvalidValue) beyond the
// array's limits; device = the current device.
hipError_t noalloc_levelset(hipStream_t s, int device, const float4* pos, int64_t np, int N,
                            float maxd, float band, float rad, float* phi, std::string& err);
// scene_geom.hip: the reference's scene geometry (meshes, per-frame transforms,
// emitters) on the host
struct GeomMesh {
    std::vector<std::array<float, 3>> v;
    std::vector<std::array<int, 4>> f;   // 1-based vertex indices; f[3] == 0: a triangle
    std::array<float, 3> lo{}, hi{};
};
struct GeomXform {
    float t[3] = {0, 0, 0}, r[3] = {0, 0, 0}, s[3] = {1, 1, 1};   // rotation in degrees
};
struct GeomNode {   // MeshContainer (mesh.hpp): per-frame meshes and transforms
    std::vector<int> mesh, xform;
    int offset = 0, interval = 1;
    bool pre_persist = false, post_persist = false;
    int frames() const { return (int)std::min(mesh.size(), xform.size()); }
};
struct GeomLiquid {
    int geom;
    float velocity[3];
};
struct SceneGeom {
    std::vector<GeomMesh> meshes;
    std::vector<GeomXform> xforms;
    std::vector<GeomNode> geoms;
    std::vector<GeomLiquid> liquids;
};
bool read_obj_mesh(const std::string& path, GeomMesh& g, std::string& err);
void cube_mesh(GeomMesh& g, const float p0[3], const float p1[3]);
void sphere_mesh(GeomMesh& g, const float center[3], float radius, int subdivs);
bool geom_in_frame(const GeomNode& n, float frame);
bool geom_contains(const SceneGeom& s, const GeomNode& n, float frame, const float p[3]);
bool geom_bounds(const SceneGeom& s, const GeomNode& n, float frame, float lo[3], float hi[3]);
void emit_geom_liquids(const SceneGeom& s, const ariel_scene_desc& d, int N, int frame,
                       std::vector<float4>& pos, std::vector<float4>& vel);
// scene_json.hip: the reference's JSON scene (settings, forces and, with geom,
// every geometry node and liquid; base_dir resolves mesh files)
int scene_from_json(const char* text, const char* base_dir, ariel_scene_desc* d, SceneGeom* geom,
                    std::string& err);

// files.hip: the OpenVDB level-set file and the Partio particle files
bool write_vdb(const char* path, const float* phi, int M, float bg, const char* grid_name, std::string& err);
int write_partio(const char* path, int64_t n, const float* pos, const float* vel, std::string& err);
int partio_format(const char* path);   // 1 ".geo", 2 ".bgeo" (either + ".gz"), 0 other
void launch_fill_f32(hipStream_t s, float* p, int64_t n, float v);

// grid (kernels_grid.hip)
void launch_divergence(hipStream_t s, const Params& p, float* const u[3], float* D);
void launch_gradient(hipStream_t s, const Params& p, const float* P, const float* L,
                     const uint8_t* A, float* const u[3]);
void launch_extrapolate(hipStream_t s, const Params& p, const uint8_t* A, float* const u[3],
                        float* const du[3]);
void launch_grid_stats(hipStream_t s, const Params& p, float* const u[3], const uint8_t* A,
                       SolverScalars* sc);

// pressure solve (kernels_pcg.hip)
void launch_precond_setup(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                          float* diag, uint8_t* nup, float* pc);
void launch_precond_build(hipStream_t s, const Params& p, const uint8_t* A, const float* diag,
                          const uint8_t* nup, float* pc, const Active& act, SolverScalars* sc);
// Fluid bits of each 8-cell z-segment (input of the tiled preconditioner apply).
void launch_fluid_segments(hipStream_t s, int N, const uint8_t* A, uint8_t* fl8);
// Active segment and tile lists from fl8; counts into sc->nseg / sc->ntile.
// seg_flag, seg_scan: N^3/8 (+1) ints; tile_any: (N/8)^3 ints; scan_sums: the
// block sums of launch_scan.
void launch_active_sets(hipStream_t s, int N, const uint8_t* fl8, int32_t* seg_flag,
                        int32_t* seg_scan, int32_t* seg_list, int32_t* tile_any,
                        int32_t* tile_lists, int32_t* scan_sums, SolverScalars* sc);
// The iteration's p += alpha s, r -= alpha y, carried by the preconditioner
// apply (tiled red-black MIC(0) only): the tile-major iterate, both
// search-direction buffers, y, and N^3.
struct IterAxpy {
    float* P;
    const float* S;
    const float* Y;
    int64_t nc;
};
// z = M^-1 r (r, q, z tile-major); the last pass finalises rho_new = z.r and
// beta in the scalars. With ax (tiled red-black MIC(0)), the first pass over
// each colour first applies the iteration's AXPY to its tiles and the finalize
// also takes ||r||^2 and the stop test.
void launch_precond_apply(hipStream_t s, const Params& p, const uint8_t* fl8, const float* pc,
                          float* r, float* q, float* z, const Reduce& red,
                          const Active& act, const IterAxpy* ax = nullptr);
// r = b - A p from the warm start P (natural; cleared off FLUID) into the
// tile-major Rt and Pt (refine: x64 = P, Pt untouched).
void launch_residual0(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                      float* P, const float* D, float* Rt, const Reduce& red, double tol,
                      int stop_rule, double* P64, int refine, float* Pt);
// The ghost level set G of a solve (L with NaN at SOLID cells) and the SOLID
// bits sb of every 8-cell line, both tile-major over every tile (the
// operator's ghost terms in the tile-major solve).
void launch_ghost_levelset(hipStream_t s, int N, const uint8_t* A, const float* L, float* G,
                           uint8_t* sb);
// s = z (first iteration) or z + beta s_prev, y = A s, alpha = rho / s.y. St
// holds two tile-major buffers selected by the iteration's parity.
// fold_sigma: no finalize; the first MIC(0) pass with the AXPY sums the
// partials itself (alpha = rho / s.As)
void launch_spmv_dot(hipStream_t s, const Params& p, const float* G, const uint8_t* sb,
                     const uint8_t* fl8, const float* Zt, float* St, float* Yt, const Reduce& red,
                     const Active& act, bool fold_sigma = false);
// p += alpha s, r -= alpha y, ||r||^2 and the stop test (tile-major).
void launch_axpy2(hipStream_t s, float* Pt, const float* St, float* Rt, const float* Yt, int64_t nc,
                  const Reduce& red, const Active& act);
// Reliable update every refine_every iterations: x64 += p, p = 0, r = b - A x64 (FP64).
void launch_reliable_update(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                            float* Pt, double* P64, const float* D, float* Rt, const Reduce& red,
                            const Active& act);
// The pressure P (natural) from the iterate Pt: (float)(x64 + p) when P64 is
// given, else p; 0 off FLUID.
void launch_refine_finish(hipStream_t s, int N, const uint8_t* A, float* P, const float* Pt,
                          const double* P64, const Active& act, const SolverScalars* sc);
// Natural <-> tile-major copy over the active tiles (gate: only while flexible
// PCG is on the network).
void launch_tile_convert(hipStream_t s, int N, const float* src, float* dst, const Active& act,
                         const SolverScalars* sc, bool to_tiles, bool gate);
void launch_spmv_plain(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                       const float* x, float* y);
void launch_solver_init(hipStream_t s, SolverScalars* sc, int max_iters, int refine_every,
                        int flex = 0);
// Flexible PCG helpers: dst = src on the active tiles; z.r and z.r_old with the
// Polak-Ribiere finalize (check: only the descent test of the network's z).
void launch_copy_list(hipStream_t s, const float* src, float* dst, const Active& act,
                      const SolverScalars* sc);
void launch_pr_dot(hipStream_t s, const float* z, const float* r, const float* r_old,
                   const Reduce& red, const Active& act, bool check);
// ||b - A p||^2 in FP64 into sc->tmp (R untouched).
void launch_true_residual_norm(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                               const float* P, const float* D, const Reduce& red,
                               const Active& act);
// End of a solve with reliable updates: x64 += p, p = 0, then ||b - A x64||^2 of
// the FP64 solution in FP64 into sc->tmp64 (R untouched). launch_refine_finish
// follows and rounds x64 to the pressure.
void launch_true_residual_norm64(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                                 float* Pt, double* P64, const float* D, const Reduce& red,
                                 const Active& act);

// reduction helpers
int reduce_blocks(int64_t n, int per_thread);

}  // namespace ariel
