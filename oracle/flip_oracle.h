/*
 * flip_oracle.h — CPU restatement of Ariel's FLIP step (TEST INFRASTRUCTURE).
 *
 * This library is the correctness oracle and the timed CPU baseline. It is
 * only ever loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, and only as the checker: the product path never links
 * or calls it.
 *
 * The descriptor layout is identical, field for field, to ariel_scene_desc so
 * that one host-side struct drives both; the oracle keeps its own copy of the
 * declaration so that it does not depend on the product's headers.
 */
#ifndef FLIP_ORACLE_H
#define FLIP_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_liquid {
    int32_t kind;
    float   lo[3];
    float   hi[3];
    float   center[3];
    float   radius;
    float   velocity[3];
} oracle_liquid;

typedef struct oracle_desc {
    int32_t  dims[3];
    float    density;
    float    dt;
    float    picflip;
    int32_t  n_forces;
    float    forces[4][3];
    int32_t  n_liquids;
    oracle_liquid liquids[8];
    int32_t  wall;
    uint32_t seed;
    int32_t  jitter;
    int32_t  precond;
    float    mic_tau;
    float    mic_sigma;
    int32_t  tile;
    float    tol;
    int32_t  max_iters;
    int32_t  refine_every;
    int32_t  stop_rule;
    int32_t  splat;
    int32_t  resample;
    float    sharpen_re;
    float    spring;
    int32_t  reserved[8];
} oracle_desc;

typedef struct oracle_stats {
    int32_t frame;
    int32_t iterations;
    int32_t solve_status;
    int32_t fluid_cells;
    int64_t particles;
    double  rhs_norm;
    double  residual_norm;
    double  true_residual_norm;
    double  max_divergence;
    double  max_density;
    double  stage_ms[10];
    double  step_ms;
    double  solve_ms;
    int32_t precond_fallback;   /* always 0: the oracle has no learned preconditioner */
    int32_t pad_;
    double  true_residual_norm64; /* FP64 solution of a refined solve, before rounding */
} oracle_stats;

typedef struct oracle_sim oracle_sim;

/* mode of oracle_step / oracle_stage(SOLVE) */
#define ORACLE_MODE_EXTENDED 0  /* the configured precond/stop rule (E3-E5)   */
#define ORACLE_MODE_FAITHFUL 1  /* solver.inl verbatim: PRef quirk, FP32 dots,
                                   absolute stop 1e-2*N^3, N^3 iterations    */

oracle_sim* oracle_create(const oracle_desc* d, int threads);
void   oracle_destroy(oracle_sim* s);
int    oracle_init(oracle_sim* s);
int    oracle_step(oracle_sim* s, int mode);
int    oracle_stage(oracle_sim* s, int stage, int mode);
int64_t oracle_num_particles(const oracle_sim* s);
void   oracle_get_particles(const oracle_sim* s, float* pos, float* vel, float* density);
void   oracle_set_particles(oracle_sim* s, const float* pos, const float* vel,
                            const float* density, int64_t n);
int64_t oracle_field_size(const oracle_sim* s, int field);
void   oracle_get_field(const oracle_sim* s, int field, float* dst);
void   oracle_set_field(oracle_sim* s, int field, const float* src);
void   oracle_get_flags(const oracle_sim* s, uint8_t* dst);
void   oracle_set_flags(oracle_sim* s, const uint8_t* src);
void   oracle_get_sort(const oracle_sim* s, int32_t* ids, int32_t* offsets);
void   oracle_get_accept(const oracle_sim* s, uint8_t* dst);
float  oracle_get_max_density(const oracle_sim* s);
void   oracle_set_max_density(oracle_sim* s, float v);
void   oracle_get_stats(const oracle_sim* s, oracle_stats* st);
int    oracle_get_frame(const oracle_sim* s);
/* FlipSim::m_frame for stage-wise runs (Step increments it before its stages). */
void   oracle_set_frame(oracle_sim* s, int frame);

/* Reference operator y = A x on the current A/L (solver.inl:196-230). */
void   oracle_apply_A(const oracle_sim* s, const float* x, float* y);
/* The same operator evaluated in double precision (true residual, E5). */
void   oracle_apply_A_f64(const oracle_sim* s, const float* x, double* y);
/* Build the configured preconditioner on the current A/L and apply it. */
void   oracle_apply_precond(oracle_sim* s, const float* r, float* z);
/* Preconditioner values (1/sqrt(pivot)) of the last build, N^3 floats. */
void   oracle_get_precond(const oracle_sim* s, float* dst);
/* FP64 PCG on the current A, L and RHS D (negated divergence), natural-order
 * MIC(0) in double, to relative residual tol; p_out N^3 doubles (may be NULL).
 * Returns the iteration count, or -1 if it did not converge. */
int    oracle_solve_fp64(const oracle_sim* s, double tol, int max_iters, double* p_out,
                         double* rel_residual);

/* Closed-form kernels (src/math/kernels.inl:25-38). */
float  oracle_smooth(float r2, float h);
float  oracle_sharpen(float r2, float h);
float  oracle_sqrlength(const float* p0, const float* p1);
/* Shared counter-based hash (E2/E6). */
uint32_t oracle_hash4(uint32_t seed, uint32_t a, uint32_t b, uint32_t c);
/* Trilinear staggered interpolation of one face component (InterpolateVelocity). */
void   oracle_interpolate_velocity(const oracle_sim* s, int which_grid, const float* p, float* u);

#ifdef __cplusplus
}
#endif

#endif
