/*
 * ariel.h — C ABI of the MI355X-native FLIP step.
 *
 * This is the drop-in boundary for Ariel's `fluidCore::FlipSim`
 * (reference: src/sim/flip.hpp:20-34). Every entry point below names the
 * reference interface it replaces. The library owns all device memory of a
 * simulation handle; host buffers are caller-owned and copied. Calls on one
 * handle are serialised on that handle's HIP stream; distinct handles may be
 * driven from distinct threads.
 *
 * Every function returns an int status: ARIEL_OK (0) or a negative
 * ARIEL_ERR_* code; ariel_last_error() returns the message of the last failure
 * on that handle (or the last handle-less failure when h is NULL).
 *
 * Units follow the reference: particle positions live in the unit box [0,1]^3
 * (grid coordinates are maxd * p), velocities in unit-box lengths per second.
 * Grids are cubic, N = dims[0] = dims[1] = dims[2]; storage is x-slowest,
 * z-fastest (Grid<T>::m_rawgrid[x][y][z], src/grid/grid.inl:11-62):
 *   cell (i,j,k)   -> (i*N + j)*N + k                       (N^3 values)
 *   x-face (i,j,k) -> (i*N + j)*N + k,     i in [0,N]       ((N+1)N^2 values)
 *   y-face (i,j,k) -> (i*(N+1) + j)*N + k, j in [0,N]       ((N+1)N^2 values)
 *   z-face (i,j,k) -> (i*N + j)*(N+1) + k, k in [0,N]       ((N+1)N^2 values)
 */
#ifndef ARIEL_H
#define ARIEL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define ARIEL_OK               0
#define ARIEL_ERR_INVALID     -1   /* bad argument or descriptor            */
#define ARIEL_ERR_HIP         -2   /* HIP runtime error                     */
#define ARIEL_ERR_NOMEM       -3   /* device allocation failed              */
#define ARIEL_ERR_STATE       -4   /* call out of order (e.g. step before init) */
#define ARIEL_ERR_CAPACITY    -5   /* caller buffer too small               */
#define ARIEL_ERR_UNSUPPORTED -6   /* feature outside this build            */
#define ARIEL_ERR_IO          -7   /* file open/read/write failed           */

/* ---- cell classes (src/grid/macgrid.inl:13) ---------------------------- */
#define ARIEL_AIR   0
#define ARIEL_FLUID 1
#define ARIEL_SOLID 2

/* ---- liquid primitives (replace JSON liquid geoms, sceneloader.cpp:153-171) */
#define ARIEL_LIQUID_BOX    0      /* lo <= p < hi on every axis            */
#define ARIEL_LIQUID_SPHERE 1      /* |p - center|^2 < radius^2             */

/* ---- pressure preconditioners ------------------------------------------ */
#define ARIEL_PRECOND_MIC0_TILED   0  /* MIC(0), tile red-black ordering       */
#define ARIEL_PRECOND_MIC0         1  /* MIC(0), natural (lexicographic) order */
#define ARIEL_PRECOND_JACOBI       2  /* diagonal                              */
#define ARIEL_PRECOND_REFERENCE    3  /* solver.inl:104-135,232-286 verbatim   */
#define ARIEL_PRECOND_MIC0_BLOCK   4  /* MIC(0) inside tiles, no coupling      */
#define ARIEL_PRECOND_LEARNED      5  /* learned conv net inside flexible PCG  */
#define ARIEL_PRECOND_MG           6  /* multigrid V(1,1) cycle (no reference
                                         counterpart; single tanks)            */

/* ---- PCG stop rules ----------------------------------------------------- */
#define ARIEL_STOP_RELATIVE      0   /* ||r||_2 <= tol * ||b||_2 (FP64 norms) */
#define ARIEL_STOP_REFERENCE_ABS 1   /* r.r <= 1e-2 * N^3 (solver.inl:322,339) */

/* ---- P2G kernels --------------------------------------------------------- */
#define ARIEL_SPLAT_SHARPEN   0      /* reference (particlegridoperations.inl:146-249) */
#define ARIEL_SPLAT_TRILINEAR 1      /* hat weights over the same window       */

/* How P2G sums each face (particlegridoperations.inl:146-249 defines the sums). */
#define ARIEL_P2G_ORDERED 0   /* per-face gather in the reference's order: bit-exact      */
#define ARIEL_P2G_TILED   1   /* scatter from 8^3 cell tiles: each lane sums its cell's
                                 particles into its 54 candidate faces in registers, the
                                 waves add those sums into an LDS face tile in a fixed
                                 order, and a second pass combines the <= 8 tiles that
                                 cover a face in a fixed order. Deterministic, within
                                 1e-6 relative of the ordered sums (not bit-exact).
                                 Sharpen radius <= 1.5 cells or the hat kernel.         */

/* ---- grid fields for ariel_get_macgrid / ariel_set_macgrid -------------- */
#define ARIEL_FIELD_UX   0   /* m_u_x                                   */
#define ARIEL_FIELD_UY   1   /* m_u_y                                   */
#define ARIEL_FIELD_UZ   2   /* m_u_z                                   */
#define ARIEL_FIELD_DUX  3   /* m_mgrid_previous.m_u_x (prev, then du)  */
#define ARIEL_FIELD_DUY  4
#define ARIEL_FIELD_DUZ  5
#define ARIEL_FIELD_P    6   /* pressure, warm start across steps       */
#define ARIEL_FIELD_D    7   /* divergence, negated (solver RHS)         */
#define ARIEL_FIELD_L    8   /* CellSDF level set                        */

/* ---- stages of FlipSim::Step, for stage-wise parity -------------------- */
#define ARIEL_STAGE_SORT        0  /* ParticleGrid::Sort                    */
#define ARIEL_STAGE_DENSITY     1  /* ComputeDensity + ApplyExternalForces  */
#define ARIEL_STAGE_MARK        2  /* MarkCellTypes + BuildSDF (A and L)    */
#define ARIEL_STAGE_P2G         3  /* Splat + StorePreviousGrid + boundary  */
#define ARIEL_STAGE_DIVERGENCE  4  /* Project divergence + FlipGrid         */
#define ARIEL_STAGE_SOLVE       5  /* BuildPreconditioner + PCG             */
#define ARIEL_STAGE_GRADIENT    6  /* SubtractPressureGradient + boundary   */
#define ARIEL_STAGE_EXTRAPOLATE 7  /* ExtrapolateVelocity + SubtractPrevious*/
#define ARIEL_STAGE_ADVECT      8  /* SolvePicFlip + AdvectParticles        */
#define ARIEL_STAGE_RESAMPLE    9  /* ResampleParticles                     */
#define ARIEL_NUM_STAGES       10

/* ---- solver status ------------------------------------------------------- */
#define ARIEL_SOLVE_CONVERGED  0
#define ARIEL_SOLVE_MAX_ITERS  1
#define ARIEL_SOLVE_BREAKDOWN  2   /* s.As <= 0                              */
#define ARIEL_SOLVE_NONFINITE  3   /* NaN/Inf in a reduction                 */

/* ---- kernel groups for per-kernel timing (ariel_get_kernel_stats) -------- */
#define ARIEL_KERNEL_SORT        0   /* bin + scan + scatter + cell order + gather */
#define ARIEL_KERNEL_DENSITY     1
#define ARIEL_KERNEL_MARK        2
#define ARIEL_KERNEL_P2G         3
#define ARIEL_KERNEL_SPMV        4   /* z = A s with the s.z reduction (PCG)    */
#define ARIEL_KERNEL_PRECOND     5   /* one preconditioner apply (all passes)   */
#define ARIEL_KERNEL_AXPY        6   /* p += a s, r -= a z with ||r||^2         */
#define ARIEL_KERNEL_GRADIENT    7
#define ARIEL_KERNEL_EXTRAPOLATE 8
#define ARIEL_KERNEL_ADVECT      9   /* G2P FLIP/PIC blend + advection          */
#define ARIEL_KERNEL_RESAMPLE   10   /* spring + resample pass (sort excluded)  */
#define ARIEL_NUM_KERNELS       11

#define ARIEL_MAX_FORCES  4
#define ARIEL_MAX_LIQUIDS 8

typedef struct ariel_liquid {
    int32_t kind;           /* ARIEL_LIQUID_*                               */
    float   lo[3];          /* box corner (unit coordinates)                */
    float   hi[3];          /* box corner                                   */
    float   center[3];      /* sphere centre                                */
    float   radius;         /* sphere radius                                */
    float   velocity[3];    /* initial velocity (sceneloader.cpp:158-167)   */
} ariel_liquid;

/* Scene + solver description. ariel_desc_defaults() fills the reference's
 * defaults (sceneloader.cpp:411-443, flip.cpp:15-30, solver.inl:107). */
typedef struct ariel_scene_desc {
    int32_t  dims[3];       /* grid edge per axis; must be equal            */
    float    density;       /* lattice spacing in cells, 1/m for integer m  */
    float    dt;            /* step size                                    */
    float    picflip;       /* FLIP fraction of the PIC/FLIP blend          */
    int32_t  n_forces;
    float    forces[ARIEL_MAX_FORCES][3];
    int32_t  n_liquids;
    ariel_liquid liquids[ARIEL_MAX_LIQUIDS];
    int32_t  wall;          /* static SOLID wall thickness in cells         */
    uint32_t seed;          /* seed of the octant jitter and resampler RNG  */
    int32_t  jitter;        /* 1: uniform jitter inside each lattice octant */
    int32_t  precond;       /* ARIEL_PRECOND_*                              */
    float    mic_tau;       /* MIC(0) modification (0.97)                   */
    float    mic_sigma;     /* MIC(0) safety (0.25)                         */
    int32_t  tile;          /* tile edge of the tiled preconditioners       */
    float    tol;           /* relative residual target                     */
    int32_t  max_iters;
    int32_t  refine_every;  /* true-residual recompute period (0 = never)   */
    int32_t  stop_rule;     /* ARIEL_STOP_*                                 */
    int32_t  splat;         /* ARIEL_SPLAT_*                                */
    int32_t  resample;      /* 1: run ResampleParticles each step           */
    float    sharpen_re;    /* P2G Sharpen radius in cells (1.4)            */
    float    spring;        /* resampler spring force (50)                  */
    int32_t  p2g;           /* ARIEL_P2G_* (0: ordered, the reference's sums) */
    int32_t  reserved[7];
} ariel_scene_desc;

typedef struct ariel_step_stats {
    int32_t frame;              /* FlipSim::m_frame                          */
    int32_t iterations;         /* PCG iterations of the last solve          */
    int32_t solve_status;       /* ARIEL_SOLVE_*                             */
    int32_t fluid_cells;        /* cells classified FLUID                    */
    int64_t particles;
    double  rhs_norm;           /* ||b||_2 over fluid cells                  */
    double  residual_norm;      /* recursive ||r||_2 at exit                 */
    double  true_residual_norm; /* ||b - A p||_2 recomputed at exit          */
    double  max_divergence;     /* max |div u| over fluid cells after projection */
    double  max_density;        /* FlipSim::m_max_density                    */
    double  stage_ms[ARIEL_NUM_STAGES]; /* device time of each stage, last step */
    double  step_ms;            /* device time of the last step              */
    double  solve_ms;           /* device time of the last pressure solve    */
    int32_t precond_fallback;   /* learned preconditioner, last solve: 0 network used throughout,
                                   1 switched to MIC(0) for slow progress, 2 for a network
                                   output that was no descent direction */
    int32_t fallback_iter;      /* iteration of that switch (0 if none) */
    double  true_residual_norm64; /* ||b - A x||_2 of the FP64 solution x of a solve with
                                     reliable updates, before its rounding to the FP32
                                     pressure (equal to true_residual_norm otherwise) */
} ariel_step_stats;

typedef struct ariel_kernel_stats {
    int64_t launches;          /* timed launches (groups) since the last reset */
    double  total_ms;          /* summed device time of the timed launches     */
    double  bytes_per_launch;  /* algorithmic HBM bytes of one launch          */
    int64_t calls;             /* all launches (groups) while profiling, timed
                                  or not: total time ~ total_ms / launches * calls */
} ariel_kernel_stats;

typedef struct ariel_sim ariel_sim;

/* Fill d with the reference defaults: 32^3, density 0.5, dt 0.005
 * (sceneloader.cpp:411-443), FLIP 0.95 (flip.cpp:27), Sharpen RE 1.4
 * (particlegridoperations.inl:149), spring 50 (particleresampler.inl:42),
 * MIC(0) tau 0.97 sigma 0.25, tol 1e-6, 500 iterations, no liquids. */
int ariel_desc_defaults(ariel_scene_desc* d);

/* FlipSim::FlipSim(maxres, density, stepsize, scene, verbose)
 * (src/sim/flip.cpp:15-30): validates d and allocates device state. */
int ariel_create(const ariel_scene_desc* d, ariel_sim** out);

/* FlipSim::~FlipSim (src/sim/flip.cpp:32-40). NULL is accepted. */
int ariel_destroy(ariel_sim* h);

/* Message of the last failure on h (h may be NULL). Never NULL. */
const char* ariel_last_error(const ariel_sim* h);

/* FlipSim::Init (src/sim/flip.cpp:42-77): max-density calibration, seeding
 * (Scene::GenerateParticles, src/scene/scene.cpp:257-360), sort, cell marks. */
int ariel_init(ariel_sim* h);

/* nsteps x FlipSim::Step (src/sim/flip.cpp:92-128), without export. */
int ariel_step(ariel_sim* h, int32_t nsteps);

/* One solver trial of ariel_step_autotune. */
typedef struct ariel_solver_trial {
    int32_t precond;        /* ARIEL_PRECOND_* tried                       */
    int32_t iterations;     /* PCG iterations of its solve                 */
    int32_t status;         /* ARIEL_SOLVE_*                               */
    float   solve_ms;       /* device time of its second solve             */
} ariel_solver_trial;

/* Solver autotuner (no reference counterpart): one FlipSim::Step whose pressure
 * solve is first run with each of the n candidate preconditioners (ARIEL_PRECOND_
 * MIC0_TILED, _MIC0_BLOCK, _JACOBI, _MG; 1 <= n <= 8), twice each from the same
 * warm start on the same right-hand side, and timed; the fastest candidate that
 * converged becomes the handle's preconditioner, and the step finishes with its
 * solve, so the step equals an ariel_step with that preconditioner. trials (n
 * entries, may be NULL) receives the measurements, *chosen (may be NULL) the
 * index of the pick. */
int ariel_step_autotune(ariel_sim* h, const int32_t* candidates, int32_t n, ariel_solver_trial* trials,
                        int32_t* chosen);

/* One stage of Step on the current device state (ARIEL_STAGE_*). Stages after
 * ARIEL_STAGE_SORT assume the particles are sorted. Used by parity tests. */
int ariel_run_stage(ariel_sim* h, int32_t stage);

/* Block until all work queued on h has finished. */
int ariel_sync(ariel_sim* h);

/* FlipSim::m_frame (src/sim/flip.hpp:33). */
int ariel_get_frame(ariel_sim* h, int32_t* frame);

/* FlipSim::GetDimensions (src/sim/flip.cpp:747-749). */
int ariel_get_dimensions(ariel_sim* h, int32_t dims[3]);

/* Number of particles (size of FlipSim::GetParticles()). */
int ariel_particle_count(ariel_sim* h, int64_t* n);

/* FlipSim::GetParticles (src/sim/flip.cpp:743-745): copies particles out in
 * particle-vector order (particle id order). pos/vel are n*3 floats, density
 * n floats, type n bytes (Particle::m_type, src/grid/macgrid.inl:36-42: every
 * particle of this build is ARIEL_FLUID, the reference's solid particles belong
 * to the solid geometry it does not run); any pointer may be NULL. cap is the
 * capacity in particles. */
int ariel_get_particles(ariel_sim* h, float* pos, float* vel, float* density,
                        uint8_t* type, int64_t cap, int64_t* n);

/* Replace all particles (id = index). Used to load a state; the particles are
 * unsorted afterwards, so the next stage to run must be ARIEL_STAGE_SORT. On a
 * handle that was never initialised, the max-density calibration of
 * FlipSim::Init (src/sim/flip.cpp:42-71) runs first, so that ComputeDensity
 * divides by the reference's m_max_density. */
int ariel_set_particles(ariel_sim* h, const float* pos, const float* vel,
                        const float* density, int64_t n);

/* MAC-grid accessor (no reference counterpart: m_mgrid is private,
 * src/sim/flip.hpp:52). which = ARIEL_FIELD_*; cap in floats. */
int ariel_get_macgrid(ariel_sim* h, int32_t which, float* dst, int64_t cap);
int ariel_set_macgrid(ariel_sim* h, int32_t which, const float* src, int64_t n);

/* Cell classes m_A (src/grid/macgrid.inl:21-34) as one byte per cell. */
int ariel_get_flags(ariel_sim* h, uint8_t* dst, int64_t cap);
int ariel_set_flags(ariel_sim* h, const uint8_t* src, int64_t n);

/* Result of the last particle sort (ParticleGrid::Sort, particlegrid.cpp:140-172):
 * ids[np] lists particle ids cell by cell, each cell in id order; offsets[N^3+1]
 * is the exclusive prefix sum of per-cell counts. Either pointer may be NULL. */
int ariel_get_sort(ariel_sim* h, int32_t* ids, int64_t ids_cap,
                   int32_t* offsets, int64_t offsets_cap);

/* Accept mask of the last ResampleParticles (particleresampler.inl:103-117),
 * one byte per particle in id order. */
int ariel_get_resample_accept(ariel_sim* h, uint8_t* dst, int64_t cap);

/* FlipSim::m_max_density (src/sim/flip.cpp:42-71). */
int ariel_get_max_density(ariel_sim* h, float* v);
int ariel_set_max_density(ariel_sim* h, float v);

int ariel_get_stats(ariel_sim* h, ariel_step_stats* s);

/* Reference matrix-free operator y = A x (ComputeAx, solver.inl:196-230) on
 * the current cell classes and level set; x and y are N^3 host floats. */
int ariel_apply_laplacian(ariel_sim* h, const float* x, float* y);

/* Preconditioner of the configured kind built on the current A and L
 * (BuildPreconditioner, solver.inl:104-135) and applied: z = M^-1 r. */
int ariel_apply_preconditioner(ariel_sim* h, const float* r, float* z);

/* Per-kernel device timing with HIP events recorded on the handle's stream
 * around each launch group (no reference counterpart: measurement only).
 * Enabling adds two event records per launch group, each of which idles the
 * GPU for a few microseconds; inside the PCG loop only one iteration in 16 is
 * timed for that reason. Stats accumulate until reset. */
int ariel_set_profiling(ariel_sim* h, int32_t enable);
int ariel_reset_kernel_stats(ariel_sim* h);
int ariel_get_kernel_stats(ariel_sim* h, int32_t kernel, ariel_kernel_stats* out);

/* Binary state dump/restore: particles, all grid fields, flags, max density,
 * frame counter. The descriptor must match the file's. */
int ariel_save_state(ariel_sim* h, const char* path);
int ariel_load_state(ariel_sim* h, const char* path);

/* FlipSim(maxres, density, stepsize, Scene*, ...) with the Scene of
 * SceneLoader (src/scene/sceneloader.cpp:11-474): the reference's JSON scene,
 * text in json, mesh files read relative to base_dir. Settings (dim, density,
 * step_size) and global forces come from the scene, the solver fields of
 * *solver (every other field; defaults if solver is NULL). Geometry: OBJ mesh
 * files (faces of 3 or 4 vertices) and the box and sphere generators,
 * tessellated as the reference does; transforms with translation, rotation
 * (degrees, R = Rz Ry Rx) and scale, interpolated between frames; animated
 * mesh sequences; geoms with frame offset, interval and pre/post persistence.
 * Every liquid geom emits particles (lattice points it contains, by ray
 * parity along +z, scene.cpp:362-378, 436-455) at Init and on every step whose
 * frame it is in (Scene::GenerateParticles, scene.cpp:257-300), so a
 * persistent emitter keeps adding particles. Solid geometry returns
 * ARIEL_ERR_UNSUPPORTED (no solid constraints in this build). */
int ariel_create_from_json(const char* json, const char* base_dir, const ariel_scene_desc* solver,
                           ariel_sim** out);

/* FlipSim::GetScene (src/sim/flip.hpp:31): the descriptor the handle runs with
 * (for ariel_create_from_json, the scene's settings and forces with the solver
 * fields). */
int ariel_get_desc(ariel_sim* h, ariel_scene_desc* out);

/* The particles the scene's liquid geoms emit on `frame` (host only, no GPU):
 * positions and velocities in the order they join a simulation; *n is set to
 * their count (pos and vel may be NULL to query it). No reference counterpart. */
int ariel_scene_emit(const char* json, const char* base_dir, const ariel_scene_desc* solver,
                     int32_t frame, float* pos, float* vel, int64_t cap, int64_t* n);

/* Scene::ExportParticles VDB path (src/scene/scene.cpp:84-101, LevelSet from
 * particles, src/grid/levelset.cpp:101-111): the fluid particles' level set,
 * rasterised on the GPU on the (N+1)^3 nodes of a grid of voxel size 1 in grid
 * units (a particle sits at p * maxd), limited to the narrow band value 3:
 *  - ARIEL_LEVELSET_SPHERES (the default; rasterizeSpheres, levelset.cpp:108):
 *    min over particles of |x - p| - 0.5;
 *  - ARIEL_LEVELSET_TRAILS (what the reference runs: rasterizeTrails,
 *    levelset.cpp:109; the C++ facade's choice): each particle a trail of
 *    spheres from radius 0.5 at p * maxd backwards along its velocity, the
 *    radius falling to 0.01 over the trail's length |v| (OpenVDB
 *    ParticlesToLevelSet's rule, restated).
 * ariel_set_levelset_mode picks one for the handle's later exports (the level
 * set, and the mesh of ariel_export_mesh).
 * A path ending in ".vdb" is written as an OpenVDB file (LevelSet::
 * WriteVDBGridToFile, levelset.cpp:195-201; file format 222, one level-set
 * FloatGrid "surface", voxel size 1, active nodes |phi| < 3); any other path as
 * a raw little-endian float32 volume (x-major) after a text header. */
#define ARIEL_LEVELSET_SPHERES 0
#define ARIEL_LEVELSET_TRAILS  1
int ariel_set_levelset_mode(ariel_sim* h, int32_t mode);
int ariel_export_levelset(ariel_sim* h, const char* path);

/* Allocation-free variant of ariel_export_levelset (no counterpart in the
 * reference; it is for callers that cannot allocate device memory: code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). The same
 * level set (ARIEL_LEVELSET_SPHERES), bit for bit, written into the caller's DEVICE buffer phi of cap
 * floats (at least (N+1)^3, x-major) instead of a file, queued on the handle's
 * stream without synchronising (ariel_sync(h) waits for it). Its workspace is a
 * fixed device array of about 26 MiB that the library holds on every device it
 * runs on until the process ends; it calls no hipMalloc/hipFree. Limits: grid
 * edge <= ARIEL_NOALLOC_MAX_N and particles <= ARIEL_NOALLOC_MAX_PARTICLES,
 * beyond which it returns ARIEL_ERR_CAPACITY. Every call on one device shares
 * that workspace: calls that overlap on one device (other streams, other host
 * threads) race and must be serialised by the caller. */
#define ARIEL_NOALLOC_MAX_N         64
#define ARIEL_NOALLOC_MAX_PARTICLES 1048576
int ariel_export_levelset_noalloc(ariel_sim* h, float* phi, int64_t cap);

/* Scene::ExportParticles OBJ path (scene.cpp:93-99, LevelSet::WriteObjToFile,
 * levelset.cpp:113-173): the zero surface of the same level set as a dual
 * quad mesh (one vertex per voxel the surface crosses, one quad per crossed
 * voxel edge, normals out of the liquid), in grid units, written like
 * Obj::WriteObj ("v x y z" lines, then "f a b c d" with 1-based indices). */
int ariel_export_mesh(ariel_sim* h, const char* path);

/* Scene::ExportParticles Partio path (scene.cpp:50-82): every fluid particle
 * with its position in grid units, velocity "v" and integer "id" (its index in
 * the file), in GetParticles order. The format follows the file extension as
 * Partio::write's does: ".geo" (Houdini classic ASCII geometry) or ".bgeo"
 * (its binary form), each optionally gzipped with ".gz" appended; other
 * extensions give ARIEL_ERR_UNSUPPORTED. */
int ariel_export_particles(ariel_sim* h, const char* path);

/* The writers of ariel_export_levelset and ariel_export_particles on a caller's
 * data (host only, no GPU): the M^3 node volume phi (x-major) with background
 * bg as an OpenVDB file, and n particles (positions already in grid units,
 * velocities) as the Partio file of path's extension. No reference counterpart
 * (the reference writes through OpenVDB and Partio). */
int ariel_write_vdb(const float* phi, int32_t M, float bg, const char* path);
int ariel_write_particles(const char* path, int64_t n, const float* pos, const float* vel);

/* The mesher of ariel_export_mesh on a caller's level set (host only, no GPU):
 * phi holds M^3 node values, x-major (the volume ariel_export_levelset writes);
 * writes the OBJ to path and returns the vertex and quad counts. No reference
 * counterpart (the reference meshes inside OpenVDB). */
int ariel_mesh_sdf(const float* phi, int32_t M, const char* path, int64_t* nverts, int64_t* nquads);

/* Weights of the learned preconditioner (ARIEL_PRECOND_LEARNED; no reference
 * counterpart): the packed BF16 blob written by ariel_amd/learned/train.py --
 * 8-byte magic "ARIELNN1", int32 layer count (11), int32 reserved, int64 value
 * count, then that many little-endian BF16 values in the kernels' layout.
 * Required before the first solve of a handle with that preconditioner;
 * allocates the network's activation buffers. */
int ariel_set_learned_weights(ariel_sim* h, const void* blob, int64_t bytes);

/* Solve A x = rhs for the current cell classes and level set with the
 * configured preconditioner, from x = 0 (no reference counterpart: it builds
 * the learned preconditioner's training pairs and serves the tests). rhs and x
 * are N^3 host floats; overwrites the handle's divergence and pressure. */
int ariel_solve_rhs(ariel_sim* h, const float* rhs, float* x, ariel_step_stats* stats);

/* The reference's JSON scene (src/scene/sceneloader.cpp) as a descriptor:
 * defaults, then settings (density, step_size, dim), globalforces, box and
 * sphere mesh generators with translation/scale transforms, one-frame mesh
 * geoms and liquid sim entries with velocities (geometry in grid cells).
 * Meshes from files, rotations, animation, persistent emitters and solids give
 * ARIEL_ERR_UNSUPPORTED; malformed JSON gives ARIEL_ERR_INVALID. */
int ariel_desc_from_json(const char* json_text, ariel_scene_desc* out);

/* ---------------------------------------------------------------- batches
 * Independent tanks stepped together (no counterpart in the reference, which
 * runs one FlipSim per process). Every stage of a step is one kernel launch
 * over all tanks; the launch sequence is replayed from per-tank argument tables
 * and can be captured as HIP graphs. Each tank's results equal those of
 * stepping it alone with ariel_step. Tanks must share preconditioner, splat,
 * refinement period, resampling switch and iteration limit; grid sizes and
 * scenes may differ. */
typedef struct ariel_batch ariel_batch;

int ariel_batch_create(const ariel_scene_desc* descs, int32_t ntanks, ariel_batch** out);
int ariel_batch_destroy(ariel_batch* b);
const char* ariel_batch_last_error(const ariel_batch* b);
int ariel_batch_size(ariel_batch* b, int32_t* ntanks);
/* The tank's handle for the per-handle accessors (owned by the batch). */
int ariel_batch_tank(ariel_batch* b, int32_t i, ariel_sim** tank);
/* ariel_init on every tank. */
int ariel_batch_init(ariel_batch* b);
/* Autotuner: for each grid size, times P2G with face-column tile edges
 * {4, 8, 16, 32} on one tank of that size and keeps the fastest for all. */
int ariel_batch_autotune(ariel_batch* b);
int ariel_batch_get_tuning(ariel_batch* b, int32_t i, int32_t* p2g_tile);
/* Capture the replayed launches as HIP graphs (off by default). */
int ariel_batch_set_graphs(ariel_batch* b, int32_t enable);
/* nsteps batched steps (one launch per stage for all tanks). */
int ariel_batch_step(ariel_batch* b, int32_t nsteps);
/* nsteps steps as a loop of ariel_step over the tanks: the baseline. */
int ariel_batch_step_sequential(ariel_batch* b, int32_t nsteps);

/* ---------------------------------------------------------------- training
 * Training of the learned preconditioner's network on the GPU with this
 * library's own forward (BF16 MFMA) and backward (FP32) kernels and Adam on
 * FP32 master weights (no counterpart in the reference). Weights are the 11
 * layers in the packed layout of the ARIELNN1 blob, as FP32. Data pointers are
 * DEVICE pointers to `batch` consecutive grids of N^3 floats each (N a
 * multiple of 16): residuals r, exact errors e with A e = r, the matrix
 * diagonal diag (ADiag, unscaled) and the FLUID mask fluid (1/0). The loss of a
 * pair is the signed A-angle 1 - (r.z)|r.z|/((r.e)(z.Az)) of the network
 * output z. */
typedef struct ariel_trainer ariel_trainer;

int ariel_trainer_create(ariel_trainer** out);
int ariel_trainer_destroy(ariel_trainer* t);
const char* ariel_trainer_last_error(const ariel_trainer* t);
/* n = the packed weight count (248064); set resets the Adam state. */
int ariel_trainer_set_weights(ariel_trainer* t, const float* w, int64_t n);
int ariel_trainer_get_weights(ariel_trainer* t, float* w, int64_t n);
/* Mean loss over the batch into *loss; with update != 0 one Adam step (learning
 * rate lr, gradient clipped to norm clip if clip > 0); grad (host, n floats,
 * may be NULL) receives the batch-mean gradient. */
int ariel_trainer_step(ariel_trainer* t, int32_t N, int32_t batch, const float* r, const float* e,
                       const float* diag, const float* fluid, float lr, float clip, int32_t update,
                       double* loss, float* grad);
/* z = network(r) for each grid of the batch (device pointers). */
int ariel_trainer_forward(ariel_trainer* t, int32_t N, int32_t batch, const float* r,
                          const float* diag, const float* fluid, float* z);

/* Self-check of the device division routines against IEEE division on the
 * current device (no counterpart in the reference: it guards the bit-exactness
 * of Smooth and Sharpen inside the kernels). For every float x with
 * lo <= x <= hi (lo >= 0):
 *   which = 0: x / a by the fixed-divisor routine (ComputeDensity's Smooth);
 *              for x < 2^-100 the compared value is max(1 - x/a, 0);
 *   which = 1: a / x by the reciprocal-based routine (Sharpen);
 *   which = 2: 1.96f / x (the Sharpen numerator RE^2 = 1.4^2 of P2G) by the
 *              one-correction routine P2G uses for that radius; a is unused.
 * *mismatches receives the number of x whose result differs in any bit. */
int ariel_selftest_division(int32_t which, float a, float lo, float hi, int64_t* mismatches);

/* Self-check of the network kernels' BF16 rounding (no counterpart in the
 * reference): the hardware conversion of two floats to packed BF16 against
 * round-to-nearest-even on the bit pattern, for every finite float. */
int ariel_selftest_bf16(int64_t* mismatches);

#ifdef __cplusplus
}
#endif

#endif /* ARIEL_H */
