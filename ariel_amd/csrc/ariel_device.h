// ariel_device.h — device-side arithmetic shared by the FLIP kernels.
//
// Every helper reproduces the reference's float expression and evaluation order
// so that the kernels agree bit for bit with a serial evaluation (sources are
// compiled with -ffp-contract=off and IEEE division/sqrt):
//   glm::max(x,y) = (x<y)?y:x, glm::min(x,y) = (y<x)?y:x.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ariel {

__host__ __device__ __forceinline__ float gmax(float x, float y) { return (x < y) ? y : x; }
__host__ __device__ __forceinline__ float gmin(float x, float y) { return (y < x) ? y : x; }
// gmax(x, c) as one v_max_f32, for x that is never NaN and never -0.0 (where
// gmax would keep -0.0 and the hardware max may give +0.0): differences
// 1 - q and q - 1 of finite q (an exact zero difference is +0.0 under
// round-to-nearest) and sums of squares.
__device__ __forceinline__ float gmax_num(float x, float c) { return fmaxf(x, c); }

// Sqrlength (src/math/kernels.inl:25-30)
__device__ __forceinline__ float sqrlength(float ax, float ay, float az, float bx, float by,
                                           float bz) {
    float a = ax - bx, b = ay - by, c = az - bz;
    return a * a + b * b + c * c;
}
// Smooth (kernels.inl:32-34)
__device__ __forceinline__ float smooth(float r2, float h) {
    return gmax(1.0f - r2 / (h * h), 0.0f);
}
// Sharpen (kernels.inl:36-38)
__device__ __forceinline__ float sharpen(float r2, float h) {
    return gmax(h * h / gmax(r2, 1.0e-5f) - 1.0f, 0.0f);
}

// Correctly rounded x / d for a divisor that is fixed over many calls, given
// yd = 1.0f / d (itself correctly rounded). q = x*yd lies within two ulps of
// x/d; one residual correction brings it within one ulp, and then, yd being the
// correctly rounded reciprocal, a second correction yields the correctly rounded
// quotient (Markstein's theorem, IBM J. Res. Dev. 34(1), 1990; Muller et al.,
// Handbook of Floating-Point Arithmetic, 2nd ed., sec. 4.7). The residuals
// r = x - q*d are exact fmas unless they underflow, which needs x below about
// 2^-100 for the divisors used here; such quotients are below 2^-80 and
// smooth_fixed() returns exactly 1 for them either way.
// ariel_selftest_division compares it with IEEE division for every float x in
// a range.
__device__ __forceinline__ float div_fixed(float x, float d, float yd) {
    float q = x * yd;
    float r = fmaf(-q, d, x);
    q = fmaf(r, yd, q);
    r = fmaf(-q, d, x);
    return fmaf(r, yd, q);
}
// smooth(r2, h) with hh = h*h and yhh = 1.0f / hh: same value, bit for bit.
__device__ __forceinline__ float smooth_fixed(float r2, float hh, float yhh) {
    return gmax_num(1.0f - div_fixed(r2, hh, yhh), 0.0f);
}
// smooth(r2, h) when hh = h*h is a power of two and yhh = 1.0f / hh (then exact):
// r2 / hh and r2 * yhh are both the exact value r2 * 2^-e correctly rounded, so
// the product is the quotient, bit for bit, for every r2 (subnormal results
// included). Power-of-two grids at the reference's lattice densities give such
// radii (4 * 0.5 / 256 = 2^-7).
// One fma, 1 - r2*yhh rounded once, gives the bits of the multiply followed by
// the rounded subtraction: the product is exact, or below 2^-126 where both
// forms round to 1.
__device__ __forceinline__ float smooth_pow2(float r2, float yhh) {
    return gmax_num(fmaf(-r2, yhh, 1.0f), 0.0f);
}
__host__ __device__ inline bool is_pow2_float(float x) {
    uint32_t b;
    __builtin_memcpy(&b, &x, 4);
    const uint32_t e = (b >> 23) & 0xFFu;
    return (b >> 31) == 0 && (b & 0x7FFFFFu) == 0 && e > 0 && e < 0xFFu;
}

// Correctly rounded n / d: Newton-Raphson on the hardware reciprocal (one
// reciprocal refinement, a quotient, two residual corrections), the arithmetic
// of the compiler's IEEE expansion without its range scaling and special-value
// fix-up, so it holds while n, d and n/d are finite, nonzero and within
// [2^-60, 2^60], where that scaling is the identity.
__device__ __forceinline__ float div_nr(float n, float d) {
    float y = __builtin_amdgcn_rcpf(d);
    const float e = fmaf(-d, y, 1.0f);
    y = fmaf(e, y, y);
    float q = n * y;
    float r = fmaf(-d, q, n);
    q = fmaf(r, y, q);
    r = fmaf(-d, q, n);
    return fmaf(r, y, q);
}
// sharpen(r2, h) with hh = h*h and the division by div_nr: same value for
// 2^-60 <= hh <= 2^60 (max(r2, 1e-5) keeps the divisor in range).
__device__ __forceinline__ float sharpen_nr(float r2, float hh) {
    return gmax_num(div_nr(hh, gmax_num(r2, 1.0e-5f)) - 1.0f, 0.0f);
}

// The reference's P2G Sharpen numerator, RE^2 with RE = 1.4 cells
// (particlegridoperations.inl:149), as a float.
constexpr float kSharpenHH = 1.4f * 1.4f;
// n / d for n = kSharpenHH by the hardware reciprocal and one residual
// correction. Compared with IEEE division for every float d in [1e-5, 4]
// (ariel_selftest_division, which = 2): equal bit for bit. max(r2, 1e-5) puts the
// divisor at or above 1e-5, and beyond 4 the quotient is below 1/2, where the
// Sharpen weight is 0 whatever the last bit.
__device__ __forceinline__ float div_sharpen_hh(float d) {
    const float y = __builtin_amdgcn_rcpf(d);
    const float q = kSharpenHH * y;
    const float r = fmaf(-d, q, kSharpenHH);
    return fmaf(r, y, q);
}
// sharpen(r2, 1.4) with the division by div_sharpen_hh: same value for every r2.
__device__ __forceinline__ float sharpen_re14(float r2) {
    return gmax_num(div_sharpen_hh(gmax_num(r2, 1.0e-5f)) - 1.0f, 0.0f);
}

// Workgroups are dispatched to the 8 XCDs in turn (workgroup b to XCD b % 8).
// xcd_block maps them so that each XCD takes a contiguous eighth of the logical
// blocks: blocks that run together then share neighbour cells in their XCD's
// L2 instead of spreading every neighbourhood over all eight. The grid must be
// xcd_grid(nb) blocks; logical blocks >= nb return early.
constexpr int kXcds = 8;
__host__ __device__ __forceinline__ int64_t xcd_grid(int64_t nb) {
    return (nb + kXcds - 1) / kXcds * kXcds;
}
__device__ __forceinline__ int64_t xcd_block(int64_t nb) {
    const int64_t per = (nb + kXcds - 1) / kXcds;
    const int64_t b = blockIdx.x;
    if (b >= per * kXcds) return nb;   // beyond this grid (a launch sized for a larger one)
    return (b % kXcds) * per + b / kXcds;
}

// 32-bit integer finaliser (lowbias32) and the 4-key hash used for the octant
// jitter and the resampler's coincident-neighbour jitter.
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
__host__ __device__ __forceinline__ uint32_t hash4(uint32_t seed, uint32_t a, uint32_t b,
                                                   uint32_t c) {
    return mix32(mix32(mix32(mix32(seed) ^ a) ^ b) ^ c);
}

// Particle cell (ParticleGrid::Sort, particlegrid.cpp:156-158; ComputeDensity,
// flip.cpp:711-713): (int)max(0, min((int)maxd-1, (int)maxd * p)).
__device__ __forceinline__ int cell_coord(float p, float maxd) {
    float fm = (float)(int)maxd;
    return (int)gmax(0.0f, gmin(fm - 1.0f, fm * p));
}

__host__ __device__ __forceinline__ int64_t cidx(int N, int i, int j, int k) {
    return ((int64_t)i * N + j) * N + k;
}
// face index of component c at (i,j,k), see include/ariel.h
__host__ __device__ __forceinline__ int64_t fidx(int N, int c, int i, int j, int k) {
    if (c == 0) return ((int64_t)i * N + j) * N + k;
    if (c == 1) return ((int64_t)i * (N + 1) + j) * N + k;
    return ((int64_t)i * N + j) * (N + 1) + k;
}

// Interpolate (particlegridoperations.inl:106-118) on one staggered component with
// extents (nx, ny, nz) = n (floats).
__device__ __forceinline__ float interpolate(const float* __restrict__ q, int N, int comp, float px,
                                             float py, float pz, float nx, float ny, float nz) {
    float x = gmax(0.0f, gmin(nx, px));
    float y = gmax(0.0f, gmin(ny, py));
    float z = gmax(0.0f, gmin(nz, pz));
    int i = (int)gmin(x, nx - 2);
    int j = (int)gmin(y, ny - 2);
    int k = (int)gmin(z, nz - 2);
    float fi1 = (float)(i + 1) - x, fi0 = x - (float)i;
    float fj1 = (float)(j + 1) - y, fj0 = y - (float)j;
    float q000 = q[fidx(N, comp, i, j, k)], q100 = q[fidx(N, comp, i + 1, j, k)];
    float q010 = q[fidx(N, comp, i, j + 1, k)], q110 = q[fidx(N, comp, i + 1, j + 1, k)];
    float q001 = q[fidx(N, comp, i, j, k + 1)], q101 = q[fidx(N, comp, i + 1, j, k + 1)];
    float q011 = q[fidx(N, comp, i, j + 1, k + 1)], q111 = q[fidx(N, comp, i + 1, j + 1, k + 1)];
    float term1 = (fi1 * q000 + fi0 * q100) * fj1;
    float term2 = (fi1 * q010 + fi0 * q110) * fj0;
    float term3 = (fi1 * q001 + fi0 * q101) * fj1;
    float term4 = (fi1 * q011 + fi0 * q111) * fj0;
    return ((float)(k + 1) - z) * (term1 + term2) + (z - (float)k) * (term3 + term4);
}

// InterpolateVelocity (particlegridoperations.inl:120-133): maxd on every axis.
__device__ __forceinline__ void interpolate_velocity(float* const* g, int N, float px, float py,
                                                     float pz, float& ux, float& uy, float& uz) {
    const float x = (float)N, y = (float)N, z = (float)N;
    ux = interpolate(g[0], N, 0, x * px, y * py - 0.5f, z * pz - 0.5f, x + 1, y, z);
    uy = interpolate(g[1], N, 1, x * px - 0.5f, y * py, z * pz - 0.5f, x, y + 1, z);
    uz = interpolate(g[2], N, 2, x * px - 0.5f, y * py - 0.5f, z * pz, x, y, z + 1);
}

// Wave-level FP64 sum (64-lane wavefront) via DPP-capable shuffles.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

}  // namespace ariel
