// kernels_pcg.hip — pressure solve for gfx950: matrix-free 7-point operator,
// FP32 vectors with FP64 reductions, and MIC(0) under a tiled red-black ordering.
//
// PCG state never leaves the device inside an iteration: every reduction writes
// per-block FP64 partials, and a one-block finalize kernel sums them in a fixed
// order and updates the scalars (alpha, beta, residual, stop flag) in
// SolverScalars. Later kernels read those scalars and return early once `done`
// is set, so the host can queue iterations in chunks without synchronising.
// The iterations run over ascending lists of the active z-segments and 8^3 tiles
// (those holding a FLUID cell), built once per solve.
//
// MIC(0) (Bridson, "Fluid Simulation for Computer Graphics", ch. 5; the
// reference's BuildPreconditioner/ApplyPreconditioner, solver.inl:104-286) is
// factored under an ordering where 8^3 tiles are coloured red/black by
// (ti+tj+tk)&1, red tiles precede black ones and cells inside a tile follow the
// natural order. A black tile's cross-tile neighbours are all red (lower), a red
// tile's all black (upper). One 64-lane wave owns a tile, lane = (il,jl) owns the
// z-line of 8 cells, and the in-tile triangular solves run as a 22-step
// wavefront over i+j+k through LDS.
#include "ariel_device.h"
#include "ariel_internal.h"

namespace ariel {

namespace {

constexpr int kRedBlocks = 2048;   // cap of grid-stride reduction grids
constexpr int kTile = 8;
constexpr int kPad = 9;            // LDS stride per z-line (bank-conflict padding)

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// Grids of the solver kernels as functions of the work counts, shared by the
// host launchers and the kernels. A kernel launched on more blocks than its own
// grid (a batched launch sized for the largest tank) lets the extra blocks
// return, strides by its own grid and writes partials only for it, so each
// tank's reduction order is the one a launch of its own would have.
__host__ __device__ inline int64_t list_grid(int64_t n) {
    const int64_t b = (n + kBlock - 1) / kBlock;
    return b < 1 ? 1 : (b < kRedBlocks ? b : kRedBlocks);
}
__host__ __device__ inline int spmv_per_thread(int64_t nseg) {
    const int64_t cap = (int64_t)kBlock * kMaxPartials;
    const int64_t per = (nseg + cap - 1) / cap;
    return per < 1 ? 1 : (int)per;
}
__host__ __device__ inline int64_t spmv_grid(int64_t nseg) {
    const int64_t per_block = (int64_t)kBlock * spmv_per_thread(nseg);
    const int64_t nb = (nseg + per_block - 1) / per_block;
    return nb < 1 ? 1 : nb;
}
constexpr int kTilesPerBlock = 4;
__host__ __device__ inline int64_t mic_grid(int64_t ntiles) {
    const int64_t nb = (ntiles + kTilesPerBlock - 1) / kTilesPerBlock;
    return nb < 1 ? 1 : nb;
}
// Counts passed as -1 are read from the tank's device scalars (batched launches,
// which are recorded before the counts exist).
__device__ __forceinline__ int64_t seg_count(int64_t n, const SolverScalars* sc) {
    return n >= 0 ? n : sc->nseg;
}
__device__ __forceinline__ int64_t tile_count(int64_t n, const SolverScalars* sc, int which) {
    return n >= 0 ? n : (int64_t)sc->ntile[which];
}

// ------------------------------------------------------------ reductions
// Each reducing kernel stores one FP64 partial per block and value
// (part[slot*NV + v], plain stores); a one-block finalize kernel launched after
// it sums the partials in a fixed order and updates the solver scalars. The
// kernel boundary makes the partials visible. (A "last block finalises" scheme
// needs a device-scope release per block, which on gfx950 writes back the L2 of
// the block's XCD each time; it made every reducing kernel several times slower.)
template <int NV>
__device__ void block_partials(double (&v)[NV], double* part, int slot) {
    __shared__ double sm[NV][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double t = wave_sum(v[i]);
        if (lane == 0) sm[i][w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double t = 0.0;
            for (int ww = 0; ww < nw; ++ww) t += sm[i][ww];
            part[slot * NV + i] = t;
        }
    }
    // sm is shared by every call: a later call in the same kernel may not
    // overwrite it before thread 0 has read it
    __syncthreads();
}

__device__ __forceinline__ bool finite(double x) { return isfinite(x); }

// ------------------------------------------------------------ operator
// ComputeAx + XRef (solver.inl:138-154, 196-230): y_c = (6 x_c - sum XRef_q) / h,
// h = 1/(n*n); XRef = x_q (FLUID), x_c (SOLID or clamped out of range),
// L_q / min(1e-6, L_c) * x_c (AIR); terms in the order +x,-x,+y,-y,+z,-z.
__device__ __forceinline__ float xref(const uint8_t* __restrict__ A, const float* __restrict__ L,
                                      const float* __restrict__ x, int N, int pi, int pj, int pk,
                                      float xc, int64_t c) {
    const int a = min(max(0, pi), N - 1), b = min(max(0, pj), N - 1), cc = min(max(0, pk), N - 1);
    const int64_t q = cidx(N, a, b, cc);
    const uint8_t t = A[q];
    if (t == kFluid) return x[q];
    if (t == kSolid) return xc;
    return L[q] / gmin(1.0e-6f, L[c]) * xc;
}

__device__ __forceinline__ float spmv_cell(const uint8_t* __restrict__ A, const float* __restrict__ L,
                                           const float* __restrict__ x, int N, int i, int j, int k,
                                           int64_t c, float h) {
    const float xc = x[c];
    return (6.0f * xc - xref(A, L, x, N, i + 1, j, k, xc, c) - xref(A, L, x, N, i - 1, j, k, xc, c) -
            xref(A, L, x, N, i, j + 1, k, xc, c) - xref(A, L, x, N, i, j - 1, k, xc, c) -
            xref(A, L, x, N, i, j, k + 1, xc, c) - xref(A, L, x, N, i, j, k - 1, xc, c)) / h;
}

// The same operator in FP64 from FP32 inputs (true residual).
__device__ __forceinline__ double spmv_cell_f64(const uint8_t* __restrict__ A,
                                                const float* __restrict__ L,
                                                const float* __restrict__ x, int N, int i, int j,
                                                int k, int64_t c) {
    const double xc = x[c];
    double acc = 6.0 * xc;
    const int q[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k},
                         {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const int a = min(max(0, q[m][0]), N - 1), b = min(max(0, q[m][1]), N - 1),
                  cc = min(max(0, q[m][2]), N - 1);
        const int64_t n = cidx(N, a, b, cc);
        const uint8_t t = A[n];
        if (t == kFluid) acc -= (double)x[n];
        else if (t == kSolid) acc -= xc;
        else acc -= (double)L[n] / (double)gmin(1.0e-6f, L[c]) * xc;
    }
    return (double)N * (double)N * acc;
}

// The same operator on an FP64 vector (solution accumulated by reliable updates).
__device__ __forceinline__ double spmv_cell_f64d(const uint8_t* __restrict__ A,
                                                 const float* __restrict__ L,
                                                 const double* __restrict__ x, int N, int i,
                                                 int j, int k, int64_t c) {
    const double xc = x[c];
    double acc = 6.0 * xc;
    const int q[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k},
                         {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const int a = min(max(0, q[m][0]), N - 1), b = min(max(0, q[m][1]), N - 1),
                  cc = min(max(0, q[m][2]), N - 1);
        const int64_t n = cidx(N, a, b, cc);
        const uint8_t t = A[n];
        if (t == kFluid) acc -= x[n];
        else if (t == kSolid) acc -= xc;
        else acc -= (double)L[n] / (double)gmin(1.0e-6f, L[c]) * xc;
    }
    return (double)N * (double)N * acc;
}

__device__ __forceinline__ void decode_cell(int N, int64_t c64, int& i, int& j, int& k) {
    const uint32_t c = (uint32_t)c64, n = (uint32_t)N;   // N^3 < 2^31 for N <= 1024
    k = (int)(c % n);
    const uint32_t q = c / n;
    j = (int)(q % n);
    i = (int)(q / n);
}

// ------------------------------------------------------------ solver kernels
struct KSolverInit {
    static constexpr int kThreads = kBlock;
    struct Args {
        SolverScalars* sc;
        int max_iters;
        int refine_every;
        int flex;
    };
    static __device__ void run(const Args& kargs) {
        SolverScalars* sc = kargs.sc;
        int max_iters = kargs.max_iters;
        int refine_every = kargs.refine_every;

        if (threadIdx.x != 0) return;
        sc->flex = kargs.flex;
        sc->use_mic = 0;
        sc->restart = 0;
        sc->rnorm2_prev = 0.0;
        sc->fallback_it = 0;
        sc->fallback_reason = 0;
        sc->bnorm2 = sc->rnorm2 = sc->thr = sc->rho = sc->sigma = 0.0;
        sc->alpha = sc->beta = sc->tmp = 0.0;
        sc->done = 0;
        sc->it = 0;
        sc->status = 1;
        sc->max_iters = max_iters;
        sc->refine_every = refine_every;
        sc->fluid_cells = 0;
        sc->maxdiv_bits = 0u;
    }
};

// Warm start: P := 0 off FLUID (what the reference's first Op does), r = b - A p,
// ||b||^2 and ||r||^2 over FLUID cells; stop threshold; early exit if converged.
// With reliable updates (refine): x64 = P, r = b - A x64 in FP64; P is cleared
// by a separate fill once every thread has read it.
struct KResidual0 {
    static constexpr int kThreads = kBlock;
    struct Args {
        Params p;
        const uint8_t* A;
        const float* L;
        float* P;
        const float* D;
        float* R;
        Reduce red;
        double tol;
        int stop_rule;
        double* P64;
        int refine;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const uint8_t* __restrict__ A = kargs.A;
        const float* __restrict__ L = kargs.L;
        float* __restrict__ P = kargs.P;
        const float* __restrict__ D = kargs.D;
        float* __restrict__ R = kargs.R;
        const Reduce& red = kargs.red;
        double tol = kargs.tol;
        int stop_rule = kargs.stop_rule;
        double* __restrict__ P64 = kargs.P64;
        int refine = kargs.refine;

        const int N = p.N;
        const int64_t nc = (int64_t)N * N * N;
        const float n = (float)N;
        const float h = 1.0f / (n * n);
        const int64_t nb = list_grid(nc);
        if ((int64_t)blockIdx.x >= nb) return;
        double v[2] = {0.0, 0.0};
        for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
             c += nb * blockDim.x) {
            if (A[c] != kFluid) {
                if (!refine) P[c] = 0.0f;
                else P64[c] = 0.0;
                R[c] = 0.0f;
                continue;
            }
            int i, j, k;
            decode_cell(N, c, i, j, k);
            const float d = D[c];
            v[0] += (double)d * (double)d;
            if (refine) {
                const double r = (double)d - spmv_cell_f64(A, L, P, N, i, j, k, c);
                P64[c] = (double)P[c];
                R[c] = (float)r;
                v[1] += r * r;
            } else {
                const float ax = spmv_cell(A, L, P, N, i, j, k, c, h);
                const float r = d + (-1.0f) * ax;
                R[c] = r;
                v[1] += (double)r * (double)r;
            }
        }
        block_partials<2>(v, red.partials, blockIdx.x);
    }
};

// One XRef term for a neighbour with class t, value xq, level set lq; cell value
// xc with level set lc (solver.inl:138-154).
__device__ __forceinline__ float xref_term(uint8_t t, float xq, const float* __restrict__ L,
                                           int64_t q, float lc, float xc) {
    if (t == kFluid) return xq;
    if (t == kSolid) return xc;
    return L[q] / gmin(1.0e-6f, lc) * xc;
}

// y = A x on the quad of 4 consecutive cells t (cells 4t..4t+3) with vector loads
// (x and the classes of the quad and of its +-y and +-x neighbour quads) and the
// reference expression per cell; adds y.x over its FLUID cells to acc.
__device__ __forceinline__ void spmv_quad(const Params& p, const uint8_t* __restrict__ A,
                                          const float* __restrict__ L,
                                          const float* __restrict__ x, float* __restrict__ y,
                                          int64_t t, int i, int j, int k0, double& acc) {
    const int N = p.N;
    const int64_t NN = (int64_t)N * N;
    const float n = (float)N;
    const float h = 1.0f / (n * n);
    const float4* __restrict__ X4 = reinterpret_cast<const float4*>(x);
    const uchar4* __restrict__ A4 = reinterpret_cast<const uchar4*>(A);
    float4* __restrict__ Y4 = reinterpret_cast<float4*>(y);
    const int64_t c = 4 * t;
    const uchar4 ac = A4[t];
    if (ac.x != kFluid && ac.y != kFluid && ac.z != kFluid && ac.w != kFluid) {
        Y4[t] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float4 xc4 = X4[t];
    const float xs[4] = {xc4.x, xc4.y, xc4.z, xc4.w};
    const uint8_t as[4] = {ac.x, ac.y, ac.z, ac.w};
    // neighbour quads (clamped to self at the domain boundary)
    const bool hxp = i + 1 < N, hxm = i > 0, hyp = j + 1 < N, hym = j > 0;
    float4 xxp = xc4, xxm = xc4, xyp = xc4, xym = xc4;
    uchar4 axp = ac, axm = ac, ayp = ac, aym = ac;
    if (hxp) { xxp = X4[t + NN / 4]; axp = A4[t + NN / 4]; }
    if (hxm) { xxm = X4[t - NN / 4]; axm = A4[t - NN / 4]; }
    if (hyp) { xyp = X4[t + N / 4]; ayp = A4[t + N / 4]; }
    if (hym) { xym = X4[t - N / 4]; aym = A4[t - N / 4]; }
    float xzm = 0.f, xzp = 0.f;
    uint8_t azm = kSolid, azp = kSolid;
    if (k0 > 0) { xzm = x[c - 1]; azm = A[c - 1]; }
    if (k0 + 4 < N) { xzp = x[c + 4]; azp = A[c + 4]; }
    const float vxp[4] = {xxp.x, xxp.y, xxp.z, xxp.w}, vxm[4] = {xxm.x, xxm.y, xxm.z, xxm.w};
    const float vyp[4] = {xyp.x, xyp.y, xyp.z, xyp.w}, vym[4] = {xym.x, xym.y, xym.z, xym.w};
    const uint8_t txp[4] = {axp.x, axp.y, axp.z, axp.w}, txm[4] = {axm.x, axm.y, axm.z, axm.w};
    const uint8_t typ[4] = {ayp.x, ayp.y, ayp.z, ayp.w}, tym[4] = {aym.x, aym.y, aym.z, aym.w};
    float out[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        if (as[m] != kFluid) {
            out[m] = 0.0f;
            continue;
        }
        const int64_t cm = c + m;
        const float xcm = xs[m];
        const bool need_l = txp[m] == kAir || txm[m] == kAir || typ[m] == kAir || tym[m] == kAir ||
                            (m < 3 ? as[m + 1] == kAir : azp == kAir) ||
                            (m > 0 ? as[m - 1] == kAir : azm == kAir);
        const float lc = need_l ? L[cm] : 0.0f;
        // +x, -x, +y, -y: a missing neighbour clamps to the cell itself (FLUID)
        const float t1 = hxp ? xref_term(txp[m], vxp[m], L, cm + NN, lc, xcm) : xcm;
        const float t2 = hxm ? xref_term(txm[m], vxm[m], L, cm - NN, lc, xcm) : xcm;
        const float t3 = hyp ? xref_term(typ[m], vyp[m], L, cm + N, lc, xcm) : xcm;
        const float t4 = hym ? xref_term(tym[m], vym[m], L, cm - N, lc, xcm) : xcm;
        float t5, t6;
        if (m < 3) t5 = xref_term(as[m + 1], xs[m + 1], L, cm + 1, lc, xcm);
        else t5 = (k0 + 4 < N) ? xref_term(azp, xzp, L, cm + 1, lc, xcm) : xcm;
        if (m > 0) t6 = xref_term(as[m - 1], xs[m - 1], L, cm - 1, lc, xcm);
        else t6 = (k0 > 0) ? xref_term(azm, xzm, L, cm - 1, lc, xcm) : xcm;
        const float yc = (6.0f * xcm - t1 - t2 - t3 - t4 - t5 - t6) / h;
        out[m] = yc;
        acc += (double)yc * (double)xcm;
    }
    Y4[t] = make_float4(out[0], out[1], out[2], out[3]);
}

// z = A s, sigma = z.s, alpha = rho / sigma over the active segments (two quads
// each); y elsewhere keeps the zeros of the solve's set-up.
// One quad per thread (neighbouring lanes take the two quads of a segment, so a
// wave's loads of the list's segments are contiguous); blocks take contiguous
// runs of the list and each XCD a contiguous eighth of the blocks, so the +-x
// neighbour planes of a segment are read by the same XCD that reads them as its
// own (shared L2).
struct KSpmvDot {
    static constexpr int kThreads = 256;
    struct Args {
        Params p;
        const uint8_t* A;
        const float* L;
        const float* x;
        float* y;
        const int32_t* segs;
        int64_t nseg;
        int per_thread;
        Reduce red;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const uint8_t* __restrict__ A = kargs.A;
        const float* __restrict__ L = kargs.L;
        const float* __restrict__ x = kargs.x;
        float* __restrict__ y = kargs.y;
        const int32_t* __restrict__ segs = kargs.segs;
        int64_t nseg = kargs.nseg;
        int per_thread = kargs.per_thread;
        const Reduce& red = kargs.red;

        if (red.sc->done) return;
        nseg = seg_count(nseg, red.sc);
        const int64_t nq = 2 * nseg;   // quads
        per_thread = spmv_per_thread(nq);
        const int64_t per_block = (int64_t)blockDim.x * per_thread;
        const int64_t nb = spmv_grid(nq);
        const int64_t b = xcd_block(nb);
        if (b >= nb) return;
        double v[1] = {0.0};
        for (int m = 0; m < per_thread; ++m) {
            const int64_t u = b * per_block + (int64_t)m * blockDim.x + threadIdx.x;
            if (u >= nq) break;
            const uint32_t sg = (uint32_t)segs[u >> 1];
            const int half = (int)(u & 1);
            const uint32_t NS = (uint32_t)p.N / kTile;
            const uint32_t line = sg / NS;   // i*N + j
            const int k0 = (int)(sg - line * NS) * kTile + 4 * half;
            const int j = (int)(line % (uint32_t)p.N), i = (int)(line / (uint32_t)p.N);
            spmv_quad(p, A, L, x, y, 2 * (int64_t)sg + half, i, j, k0, v[0]);
        }
        block_partials<1>(v, red.partials, b);
    }
};

// Stop decision after the residual norm of iteration `it` is known.
__device__ __forceinline__ void decide(SolverScalars* sc) {
    const double r2 = sc->rnorm2;
    if (!finite(r2)) {
        sc->status = 3;
        sc->done = 1;
    } else if (r2 <= sc->thr) {
        sc->status = 0;
        sc->done = 1;
    } else if (sc->it >= sc->max_iters) {
        sc->status = 1;
        sc->done = 1;
    }
}
__device__ __forceinline__ bool refine_due(const SolverScalars* sc) {
    return sc->refine_every > 0 && (sc->it % sc->refine_every) == 0;
}

// p += alpha s; r += (-alpha) z (Op, solver.inl:157-177); ||r||^2; stop test.
// One active segment (two float4 of each operand) per thread and iteration.
struct KAxpy2 {
    static constexpr int kThreads = kBlock;
    struct Args {
        float* P;
        const float* S;
        float* R;
        const float* Z;
        const int32_t* segs;
        int64_t nseg;
        Reduce red;
    };
    static __device__ void run(const Args& kargs) {
        float* __restrict__ P = kargs.P;
        const float* __restrict__ S = kargs.S;
        float* __restrict__ R = kargs.R;
        const float* __restrict__ Z = kargs.Z;
        const int32_t* __restrict__ segs = kargs.segs;
        int64_t nseg = kargs.nseg;
        const Reduce& red = kargs.red;

        if (red.sc->done) return;
        nseg = seg_count(nseg, red.sc);
        const int64_t nb = list_grid(nseg);
        if ((int64_t)blockIdx.x >= nb) return;
        const float af = (float)red.sc->alpha;
        const float naf = -af;
        float4* __restrict__ P4 = reinterpret_cast<float4*>(P);
        float4* __restrict__ R4 = reinterpret_cast<float4*>(R);
        const float4* __restrict__ S4 = reinterpret_cast<const float4*>(S);
        const float4* __restrict__ Z4 = reinterpret_cast<const float4*>(Z);
        double v[1] = {0.0};
        for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nseg;
             u += nb * blockDim.x) {
            const int64_t c = 2 * (int64_t)segs[u];
            float4 p0 = P4[c], s0 = S4[c], r0 = R4[c], z0 = Z4[c];
            float4 p1 = P4[c + 1], s1 = S4[c + 1], r1 = R4[c + 1], z1 = Z4[c + 1];
            p0.x = p0.x + af * s0.x; p0.y = p0.y + af * s0.y; p0.z = p0.z + af * s0.z; p0.w = p0.w + af * s0.w;
            r0.x = r0.x + naf * z0.x; r0.y = r0.y + naf * z0.y; r0.z = r0.z + naf * z0.z; r0.w = r0.w + naf * z0.w;
            p1.x = p1.x + af * s1.x; p1.y = p1.y + af * s1.y; p1.z = p1.z + af * s1.z; p1.w = p1.w + af * s1.w;
            r1.x = r1.x + naf * z1.x; r1.y = r1.y + naf * z1.y; r1.z = r1.z + naf * z1.z; r1.w = r1.w + naf * z1.w;
            P4[c] = p0;
            R4[c] = r0;
            P4[c + 1] = p1;
            R4[c + 1] = r1;
            v[0] += (double)r0.x * (double)r0.x + (double)r0.y * (double)r0.y +
                    (double)r0.z * (double)r0.z + (double)r0.w * (double)r0.w;
            v[0] += (double)r1.x * (double)r1.x + (double)r1.y * (double)r1.y +
                    (double)r1.z * (double)r1.z + (double)r1.w * (double)r1.w;
        }
        block_partials<1>(v, red.partials, blockIdx.x);
    }
};

// Reliable update, part 1: x64 += p, p = 0 (every refine_every iterations).
struct KFlush {
    static constexpr int kThreads = kBlock;
    struct Args {
        const uint8_t* A;
        float* P;
        double* P64;
        const int32_t* segs;
        int64_t nseg;
        const SolverScalars* sc;
        int final_flush;   // 1: the end of the solve (every cell, done or not)
    };
    static __device__ void run(const Args& kargs) {
        const uint8_t* __restrict__ A = kargs.A;
        float* __restrict__ P = kargs.P;
        double* __restrict__ P64 = kargs.P64;
        const int32_t* __restrict__ segs = kargs.segs;
        int64_t nseg = kargs.nseg;
        const SolverScalars* sc = kargs.sc;

        if (!kargs.final_flush && (sc->done || !refine_due(sc))) return;
        nseg = seg_count(nseg, sc);
        const int64_t nb = list_grid(nseg);
        if ((int64_t)blockIdx.x >= nb) return;
        // one cell per thread and iteration: neighbouring lanes take neighbouring cells
        for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < 8 * nseg;
             u += nb * blockDim.x) {
            const int64_t c = 8 * (int64_t)segs[u >> 3] + (u & 7);
            if (A[c] == kFluid) P64[c] += (double)P[c];
            P[c] = 0.0f;
        }
    }
};

// mode 0 (reliable update, part 2): r = b - A x64 in FP64, stored in FP32, then the
// stop test. mode 1 (report): ||b - A p||^2 of the FP32 pressure into sc->tmp.
// mode 3 (report): ||b - A x64||^2 of the FP64 solution into sc->tmp64.
struct KTrueResidual {
    static constexpr int kThreads = kBlock;
    struct Args {
        Params p;
        const uint8_t* A;
        const float* L;
        const float* P;
        const double* P64;
        const float* D;
        float* R;
        const int32_t* segs;
        int64_t nseg;
        Reduce red;
        int mode;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const uint8_t* __restrict__ A = kargs.A;
        const float* __restrict__ L = kargs.L;
        const float* __restrict__ P = kargs.P;
        const double* __restrict__ P64 = kargs.P64;
        const float* __restrict__ D = kargs.D;
        float* __restrict__ R = kargs.R;
        const int32_t* __restrict__ segs = kargs.segs;
        int64_t nseg = kargs.nseg;
        const Reduce& red = kargs.red;
        int mode = kargs.mode;

        if (mode == 0 && (red.sc->done || !refine_due(red.sc))) return;
        nseg = seg_count(nseg, red.sc);
        const int64_t nb = list_grid(nseg);
        if ((int64_t)blockIdx.x >= nb) return;
        const int N = p.N;
        double v[1] = {0.0};
        // one cell per thread and iteration: neighbouring lanes take neighbouring cells
        for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < 8 * nseg;
             u += nb * blockDim.x) {
            const int64_t c = 8 * (int64_t)segs[u >> 3] + (u & 7);
            if (A[c] != kFluid) {
                if (mode == 0) R[c] = 0.0f;
                continue;
            }
            int i, j, k;
            decode_cell(N, c, i, j, k);
            const double ax = mode != 1 ? spmv_cell_f64d(A, L, P64, N, i, j, k, c)
                                        : spmv_cell_f64(A, L, P, N, i, j, k, c);
            const double r = (double)D[c] - ax;
            v[0] += r * r;
            if (mode == 0) R[c] = (float)r;
        }
        block_partials<1>(v, red.partials, blockIdx.x);
    }
};

// End of a refined solve: p = (float)(x64 + p) on FLUID cells, 0 elsewhere.
struct KRefineFinish {
    static constexpr int kThreads = kBlock;
    struct Args {
        const uint8_t* A;
        float* P;
        const double* P64;
        const int32_t* segs;
        int64_t nseg;
        const SolverScalars* sc;
    };
    static __device__ void run(const Args& kargs) {
        const uint8_t* __restrict__ A = kargs.A;
        float* __restrict__ P = kargs.P;
        const double* __restrict__ P64 = kargs.P64;
        const int32_t* __restrict__ segs = kargs.segs;
        int64_t nseg = seg_count(kargs.nseg, kargs.sc);

        const int64_t nb = list_grid(nseg);
        if ((int64_t)blockIdx.x >= nb) return;
        for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < 8 * nseg;
             u += nb * blockDim.x) {
            const int64_t c = 8 * (int64_t)segs[u >> 3] + (u & 7);
            P[c] = A[c] == kFluid ? (float)(P64[c] + (double)P[c]) : 0.0f;
        }
    }
};

struct KUpdateS {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float* Z;
        float* S;
        const int32_t* segs;
        int64_t nseg;
        const SolverScalars* sc;
        int first;
    };
    static __device__ void run(const Args& kargs) {
        const float* __restrict__ Z = kargs.Z;
        float* __restrict__ S = kargs.S;
        const int32_t* __restrict__ segs = kargs.segs;
        int64_t nseg = kargs.nseg;
        const SolverScalars* sc = kargs.sc;
        int first = kargs.first;

        if (sc->done) return;
        nseg = seg_count(nseg, sc);
        const int64_t nb = list_grid(nseg);
        if ((int64_t)blockIdx.x >= nb) return;
        const float bf = (float)sc->beta;
        const float4* __restrict__ Z4 = reinterpret_cast<const float4*>(Z);
        float4* __restrict__ S4 = reinterpret_cast<float4*>(S);
        for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nseg;
             u += nb * blockDim.x) {
            const int64_t c = 2 * (int64_t)segs[u];
    #pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float4 z = Z4[c + h];
                if (first) {
                    S4[c + h] = z;
                } else {
                    float4 sv = S4[c + h];
                    sv.x = z.x + bf * sv.x; sv.y = z.y + bf * sv.y; sv.z = z.z + bf * sv.z; sv.w = z.w + bf * sv.w;
                    S4[c + h] = sv;
                }
            }
        }
    }
};

struct KSpmvPlain {
    static constexpr int kThreads = kBlock;
    struct Args {
        Params p;
        const uint8_t* A;
        const float* L;
        const float* x;
        float* y;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const uint8_t* __restrict__ A = kargs.A;
        const float* __restrict__ L = kargs.L;
        const float* __restrict__ x = kargs.x;
        float* __restrict__ y = kargs.y;

        const int N = p.N;
        const int64_t nc = (int64_t)N * N * N;
        const float n = (float)N;
        const float h = 1.0f / (n * n);
        int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= nc) return;
        if (A[c] != kFluid) {
            y[c] = 0.0f;
            return;
        }
        int i, j, k;
        decode_cell(N, c, i, j, k);
        y[c] = spmv_cell(A, L, x, N, i, j, k, c, h);
    }
};

// rho_new = z.r finalisation shared by the preconditioner kernels.
__device__ __forceinline__ void finish_rho(SolverScalars* sc, double rho_new) {
    if (sc->it == 0) {
        sc->rho = rho_new;
    } else {
        sc->beta = rho_new / sc->rho;
        sc->rho = rho_new;
    }
}

// ------------------------------------------------------------ finalize
// Partial counts that depend on the active sets, for finalize launches recorded
// before the counts exist (nparts < 0 selects the producer's grid rule).
enum { PARTS_LIST = -1, PARTS_SPMV = -2, PARTS_MIC_RB = -3, PARTS_MIC_ALL = -4 };
__device__ __forceinline__ int64_t parts_from_counts(int code, const SolverScalars* sc) {
    switch (code) {
        case PARTS_LIST: return list_grid(sc->nseg);
        case PARTS_SPMV: return spmv_grid(2 * sc->nseg);
        case PARTS_MIC_RB: return mic_grid(sc->ntile[1]) + mic_grid(sc->ntile[0]);
        default: return mic_grid(sc->ntile[2]);
    }
}

enum { FIN_RESID0 = 0, FIN_SIGMA = 1, FIN_RNORM = 2, FIN_TRUE = 3, FIN_REPORT = 4, FIN_RHO = 5,
       FIN_PR = 6, FIN_PR_CHECK = 7, FIN_REPORT64 = 8 };
// Flexible PCG falls back to MIC(0) when, over a window of this many
// iterations, ||r|| fell by less than 5% per iteration on average.
// Flexible PCG's progress test: over every window of kFlexWindow iterations the
// residual norm must fall by 5% per iteration on average. The CG residual norm
// is not monotone (at 256^3 it barely moves in the first 4 iterations while the
// error falls), so the window is long.
constexpr int kFlexWindow = 16;
constexpr double flex_need() {   // 0.95^(2 kFlexWindow), on ||r||^2
    double v = 1.0;
    for (int i = 0; i < kFlexWindow; ++i) v *= 0.9025;
    return v;
}

// One block: sum the partials of the preceding reducing kernel in a fixed order
// and update the scalars. Gated exactly like its producer, so a skipped producer
// (solve finished) is matched by a skipped finalize.
struct KFinalize {
    static constexpr int kThreads = 256;
    struct Args {
        Reduce red;
        int kind;
        int nparts;
        double tol;
        int stop_rule;
        int N;
    };
    static __device__ void run(const Args& kargs) {
        const Reduce& red = kargs.red;
        int kind = kargs.kind;
        int nparts = kargs.nparts;
        double tol = kargs.tol;
        int stop_rule = kargs.stop_rule;
        int N = kargs.N;

        SolverScalars* sc = red.sc;
        if ((kind == FIN_SIGMA || kind == FIN_RNORM || kind == FIN_RHO || kind == FIN_PR ||
             kind == FIN_PR_CHECK) && sc->done) return;
        if (kind == FIN_TRUE && (sc->done || !refine_due(sc))) return;
        if (kind == FIN_RHO && sc->flex) return;          // flexible PCG sets beta itself
        if (kind == FIN_PR_CHECK && (!sc->flex || sc->use_mic)) return;
        if (nparts < 0) nparts = (int)parts_from_counts(nparts, sc);
        const int NV = (kind == FIN_RESID0 || kind == FIN_PR || kind == FIN_PR_CHECK) ? 2 : 1;
        double a0 = 0.0, a1 = 0.0;
        for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
            a0 += red.partials[b * NV];
            if (NV == 2) a1 += red.partials[b * NV + 1];
        }
        __shared__ double sm[2][4];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        a0 = wave_sum(a0);
        a1 = wave_sum(a1);
        if (lane == 0) {
            sm[0][w] = a0;
            sm[1][w] = a1;
        }
        __syncthreads();
        if (threadIdx.x != 0) return;
        double s0 = 0.0, s1 = 0.0;
        for (int i = 0; i < 4; ++i) {
            s0 += sm[0][i];
            s1 += sm[1][i];
        }
        switch (kind) {
            case FIN_RESID0:
                sc->bnorm2 = s0;
                sc->rnorm2 = s1;
                sc->rnorm2_prev = s1;
                sc->rnorm2_0 = s1;
                if (stop_rule == 1) sc->thr = (double)(1.0e-2f * (float)(N * N * N));
                else sc->thr = tol * tol * s0;
                if (s1 <= sc->thr) {
                    sc->done = 1;
                    sc->status = 0;
                }
                break;
            case FIN_SIGMA:
                sc->sigma = s0;
                if (!finite(s0) || !finite(sc->rho)) {
                    sc->status = 3;
                    sc->done = 1;
                } else if (!(s0 > 0.0)) {
                    sc->status = 2;
                    sc->done = 1;
                } else {
                    sc->alpha = sc->rho / s0;
                }
                break;
            case FIN_RNORM:
                sc->rnorm2 = s0;
                sc->it += 1;
                if (sc->flex && !sc->use_mic && sc->it % kFlexWindow == 0) {
                    // average reduction per iteration over the window below 5%
                    constexpr double need = flex_need();
                    if (!(s0 <= need * sc->rnorm2_prev)) {
                        sc->use_mic = 1;
                        sc->restart = 1;
                        sc->fallback_it = sc->it;
                        sc->fallback_reason = 1;
                    }
                    sc->rnorm2_prev = s0;
                }
                if (!refine_due(sc)) decide(sc);
                break;
            case FIN_TRUE:
                sc->rnorm2 = s0;
                decide(sc);
                break;
            case FIN_REPORT:
                sc->tmp = s0;
                break;
            case FIN_REPORT64:
                sc->tmp64 = s0;
                break;
            case FIN_RHO:
                finish_rho(sc, s0);
                break;
            case FIN_PR_CHECK:
                // the network's z must be a descent direction (z.r > 0); if not,
                // fall back to MIC(0) before z is used
                if (!(s0 > 0.0) || !finite(s0)) {
                    sc->use_mic = 1;
                    sc->restart = 1;
                    sc->fallback_it = sc->it;
                    sc->fallback_reason = 2;
                }
                break;
            case FIN_PR:
                // Polak-Ribiere: beta = z.(r - r_old) / rho_old
                if (!finite(s0) || !finite(s1)) {
                    sc->status = 3;
                    sc->done = 1;
                    break;
                }
                if (sc->it == 0 || sc->restart) sc->beta = 0.0;
                else sc->beta = (s0 - s1) / sc->rho;
                sc->rho = s0;
                sc->restart = 0;
                break;
        }
    }
};

// ------------------------------------------------------------ preconditioner
// ADiag (solver.inl:81-101) and the number of FLUID upper neighbours under the
// ordering (mode 0: tiled red-black, 4: block/no coupling); Jacobi inverse diag.
struct KPrecondSetup {
    static constexpr int kThreads = kBlock;
    struct Args {
        Params p;
        const uint8_t* A;
        const float* L;
        float* diag;
        uint8_t* nup;
        float* pc;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const uint8_t* __restrict__ A = kargs.A;
        const float* __restrict__ L = kargs.L;
        float* __restrict__ diag = kargs.diag;
        uint8_t* __restrict__ nup = kargs.nup;
        float* __restrict__ pc = kargs.pc;

        const int N = p.N;
        const int64_t nc = (int64_t)N * N * N;
        int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= nc) return;
        int i, j, k;
        decode_cell(N, c, i, j, k);
        if (A[c] != kFluid) {
            diag[c] = 6.0f;
            nup[c] = 0;
            pc[c] = 0.0f;
            return;
        }
        float dg = 6.0f;
        const int q[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                             {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        int cnt = 0;
        const int color = ((i >> 3) + (j >> 3) + (k >> 3)) & 1;
    #pragma unroll
        for (int m = 0; m < 6; ++m) {
            const int a = q[m][0], b = q[m][1], cc = q[m][2];
            const bool in = a >= 0 && a < N && b >= 0 && b < N && cc >= 0 && cc < N;
            const uint8_t t = in ? A[cidx(N, a, b, cc)] : (uint8_t)kSolid;
            if (!in || t == kSolid) {
                dg -= 1.0f;
            } else if (t == kAir) {
                dg -= L[cidx(N, a, b, cc)] / gmin(1.0e-6f, L[c]);
            }
            if (in && t == kFluid) {
                const bool same = (a >> 3) == (i >> 3) && (b >> 3) == (j >> 3) && (cc >> 3) == (k >> 3);
                const bool plus = (m & 1) != 0;
                if (same) cnt += plus ? 1 : 0;
                else if (p.precond == 0 && color == 0) cnt += 1;
            }
        }
        diag[c] = dg;
        nup[c] = (uint8_t)cnt;
        pc[c] = (p.precond == 2 && dg > 0) ? 1.0f / dg : 0.0f;
    }
};

struct TileCtx {
    int N, TN;
    int ti, tj, tk;
    int il, jl, lane;
    int gi, gj, k0;
    int64_t line;   // cell index of (gi, gj, k0)
};

__device__ __forceinline__ bool tile_of(int N, int64_t t, int color, TileCtx& c) {
    const int TN = N / kTile;
    c.N = N;
    c.TN = TN;
    c.tk = (int)(t % TN);
    c.tj = (int)((t / TN) % TN);
    c.ti = (int)(t / ((int64_t)TN * TN));
    if (color >= 0 && ((c.ti + c.tj + c.tk) & 1) != color) return false;
    c.lane = threadIdx.x & 63;
    c.il = c.lane >> 3;
    c.jl = c.lane & 7;
    c.gi = c.ti * kTile + c.il;
    c.gj = c.tj * kTile + c.jl;
    c.k0 = c.tk * kTile;
    c.line = cidx(N, c.gi, c.gj, c.k0);
    return true;
}

__device__ __forceinline__ uint32_t fluid_bits(const uint8_t* __restrict__ A, int64_t line) {
    uint32_t m = 0;
#pragma unroll
    for (int kl = 0; kl < kTile; ++kl) m |= (A[line + kl] == kFluid ? 1u : 0u) << kl;
    return m;
}

// Load a neighbouring line (8 cells) into LDS slot `dst` (stride kPad) as v = a*b
// (b may be null -> v = a) together with its fluid bits.
__device__ __forceinline__ uint32_t load_halo_line(const uint8_t* __restrict__ A,
                                                   const float* __restrict__ a,
                                                   const float* __restrict__ b, int64_t line,
                                                   float* dst) {
    uint32_t m = fluid_bits(A, line);
#pragma unroll
    for (int kl = 0; kl < kTile; ++kl) {
        const float x = a[line + kl];
        dst[kl] = b ? x * b[line + kl] : x;
    }
    return m;
}

// MIC(0) factor for the tiles of one colour (mode 0) or all tiles (mode 4).
// pivot: e = diag - sum_j pc_j^2 - tau * sum_j (nup_j - 1) pc_j^2 over FLUID lower
// neighbours j in the order -x,-y,-z,+x,+y,+z; e = diag if e < sigma*diag;
// pc = 1/sqrt(e).
struct KMicBuild {
    static constexpr int kThreads = 64;
    struct Args {
        Params p;
        const uint8_t* A;
        const float* diag;
        const uint8_t* nup;
        float* pc;
        int color;
        const int32_t* tiles;
        int64_t ntl;
        const SolverScalars* sc;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const uint8_t* __restrict__ A = kargs.A;
        const float* __restrict__ diag = kargs.diag;
        const uint8_t* __restrict__ nup = kargs.nup;
        float* __restrict__ pc = kargs.pc;
        int color = kargs.color;
        const int32_t* __restrict__ tiles = kargs.tiles;
        int64_t ntl = tile_count(kargs.ntl, kargs.sc, kargs.color < 0 ? 2 : kargs.color);

        __shared__ float s_pc[64 * kPad];
        __shared__ float s_nm[64 * kPad];
        __shared__ float s_hxp[64 * kPad];   // cross-tile x neighbour line: pc
        __shared__ float s_hxn[64 * kPad];   //                              nup-1
        __shared__ float s_hyp[64 * kPad];
        __shared__ float s_hyn[64 * kPad];
        const int N = p.N;
        const int TN = N / kTile;
        (void)TN;
        const bool black = color == 1;
        const float tau = p.mic_tau, sigma = p.mic_sigma;
        for (int64_t u = blockIdx.x; u < ntl; u += gridDim.x) {
            const int64_t t = tiles[u];
            TileCtx c;
            if (!tile_of(N, t, color, c)) continue;
            const uint32_t fl = fluid_bits(A, c.line);
            const uint32_t fl_im = __shfl(fl, (c.lane - 8) & 63, 64);
            const uint32_t fl_jm = __shfl(fl, (c.lane - 1) & 63, 64);
            float dg[kTile];
    #pragma unroll
            for (int kl = 0; kl < kTile; ++kl) dg[kl] = diag[c.line + kl];
            // halos (black tiles): x line (il==0: -x, il==7: +x), y line, z singles
            uint32_t hx_fl = 0, hy_fl = 0, hzm_fl = 0, hzp_fl = 0;
            float hzm_pc = 0.f, hzm_nm = 0.f, hzp_pc = 0.f, hzp_nm = 0.f;
            if (black) {
                int xi = -1;
                if (c.il == 0 && c.gi > 0) xi = c.gi - 1;
                if (c.il == 7 && c.gi + 1 < N) xi = c.gi + 1;
                if (xi >= 0) {
                    const int64_t ln = cidx(N, xi, c.gj, c.k0);
                    hx_fl = fluid_bits(A, ln);
    #pragma unroll
                    for (int kl = 0; kl < kTile; ++kl) {
                        s_hxp[c.lane * kPad + kl] = pc[ln + kl];
                        s_hxn[c.lane * kPad + kl] = (float)((int)nup[ln + kl] - 1);
                    }
                }
                int yj = -1;
                if (c.jl == 0 && c.gj > 0) yj = c.gj - 1;
                if (c.jl == 7 && c.gj + 1 < N) yj = c.gj + 1;
                if (yj >= 0) {
                    const int64_t ln = cidx(N, c.gi, yj, c.k0);
                    hy_fl = fluid_bits(A, ln);
    #pragma unroll
                    for (int kl = 0; kl < kTile; ++kl) {
                        s_hyp[c.lane * kPad + kl] = pc[ln + kl];
                        s_hyn[c.lane * kPad + kl] = (float)((int)nup[ln + kl] - 1);
                    }
                }
                if (c.k0 > 0) {
                    const int64_t n = c.line - 1;
                    hzm_fl = A[n] == kFluid;
                    hzm_pc = pc[n];
                    hzm_nm = (float)((int)nup[n] - 1);
                }
                if (c.k0 + kTile < N) {
                    const int64_t n = c.line + kTile;
                    hzp_fl = A[n] == kFluid;
                    hzp_pc = pc[n];
                    hzp_nm = (float)((int)nup[n] - 1);
                }
            }
    #pragma unroll
            for (int kl = 0; kl < kTile; ++kl) s_nm[c.lane * kPad + kl] = (float)((int)nup[c.line + kl] - 1);
            __syncthreads();
            for (int st = 0; st < 3 * kTile - 2; ++st) {
                const int kl = st - c.il - c.jl;
                if (kl >= 0 && kl < kTile) {
                    float res = 0.0f;
                    if ((fl >> kl) & 1u) {
                        // gather the lower neighbours' (pc, nup-1) in order -x,-y,-z,+x,+y,+z
                        float pj[6], nm[6];
                        bool has[6];
                        // -x
                        has[0] = false;
                        if (c.il > 0) {
                            if ((fl_im >> kl) & 1u) { has[0] = true; pj[0] = s_pc[(c.lane - 8) * kPad + kl]; nm[0] = s_nm[(c.lane - 8) * kPad + kl]; }
                        } else if (black && ((hx_fl >> kl) & 1u)) {
                            has[0] = true; pj[0] = s_hxp[c.lane * kPad + kl]; nm[0] = s_hxn[c.lane * kPad + kl];
                        }
                        // -y
                        has[1] = false;
                        if (c.jl > 0) {
                            if ((fl_jm >> kl) & 1u) { has[1] = true; pj[1] = s_pc[(c.lane - 1) * kPad + kl]; nm[1] = s_nm[(c.lane - 1) * kPad + kl]; }
                        } else if (black && ((hy_fl >> kl) & 1u)) {
                            has[1] = true; pj[1] = s_hyp[c.lane * kPad + kl]; nm[1] = s_hyn[c.lane * kPad + kl];
                        }
                        // -z
                        has[2] = false;
                        if (kl > 0) {
                            if ((fl >> (kl - 1)) & 1u) { has[2] = true; pj[2] = s_pc[c.lane * kPad + kl - 1]; nm[2] = s_nm[c.lane * kPad + kl - 1]; }
                        } else if (black && hzm_fl) {
                            has[2] = true; pj[2] = hzm_pc; nm[2] = hzm_nm;
                        }
                        // +x, +y, +z (cross-tile, black only)
                        has[3] = black && c.il == 7 && ((hx_fl >> kl) & 1u);
                        if (has[3]) { pj[3] = s_hxp[c.lane * kPad + kl]; nm[3] = s_hxn[c.lane * kPad + kl]; }
                        has[4] = black && c.jl == 7 && ((hy_fl >> kl) & 1u);
                        if (has[4]) { pj[4] = s_hyp[c.lane * kPad + kl]; nm[4] = s_hyn[c.lane * kPad + kl]; }
                        has[5] = black && kl == 7 && hzp_fl;
                        if (has[5]) { pj[5] = hzp_pc; nm[5] = hzp_nm; }
                        const float d = dg[kl];
                        float e = d;
    #pragma unroll
                        for (int m = 0; m < 6; ++m)
                            if (has[m]) e = e - pj[m] * pj[m];
                        float ssum = 0.0f;
    #pragma unroll
                        for (int m = 0; m < 6; ++m)
                            if (has[m]) ssum = ssum + nm[m] * (pj[m] * pj[m]);
                        e = e - tau * ssum;
                        if (d > 0) {
                            if (e < sigma * d) e = d;
                            res = 1.0f / sqrtf(e);
                        }
                    }
                    s_pc[c.lane * kPad + kl] = res;
                }
                __syncthreads();
            }
    #pragma unroll
            for (int kl = 0; kl < kTile; ++kl) pc[c.line + kl] = s_pc[c.lane * kPad + kl];
            __syncthreads();
        }
    }
};

// The MIC(0) apply reads pc and its own intermediate q tile by tile. In the
// natural layout a tile's z-line is 32 bytes between lines of the other colour's
// tiles, so every read of one colour fetches whole cache lines half of which it
// does not use. The apply therefore reads a tile-major copy of pc (tile t at
// [512 t, 512 t + 512), cell (il, jl, kl) at 64 il + 8 jl + kl), made once per
// solve over the active tiles, and keeps q in the same layout.
__host__ __device__ inline int64_t tile_major(int64_t t, int il, int jl, int kl) {
    return t * 512 + (il * 8 + jl) * 8 + kl;
}
struct KPcTileMajor {
    static constexpr int kThreads = 64;
    struct Args {
        int N;
        const float* pc;
        float* pct;
        const int32_t* tiles;
        int64_t ntiles;
        const SolverScalars* sc;
    };
    static __device__ void run(const Args& a) {
        const int TN = a.N / kTile;
        const int64_t nt = tile_count(a.ntiles, a.sc, 2);
        const int il = threadIdx.x >> 3, jl = threadIdx.x & 7;
        for (int64_t u = blockIdx.x; u < nt; u += gridDim.x) {
            const int t = a.tiles[u];
            const int tk = t % TN, tj = (t / TN) % TN, ti = t / (TN * TN);
            const int64_t ln = cidx(a.N, ti * kTile + il, tj * kTile + jl, tk * kTile);
            const float4* src = reinterpret_cast<const float4*>(a.pc + ln);
            float4* dst = reinterpret_cast<float4*>(a.pct + tile_major(t, il, jl, 0));
            dst[0] = src[0];
            dst[1] = src[1];
        }
    }
};

// Fluid bits of every z-segment of 8 cells: fl8[(i*N + j)*(N/8) + k/8], bit k%8.
struct KFluidSegments {
    static constexpr int kThreads = kBlock;
    struct Args {
        int N;
        const uint8_t* A;
        uint8_t* fl8;
    };
    static __device__ void run(const Args& kargs) {
        int N = kargs.N;
        const uint8_t* __restrict__ A = kargs.A;
        uint8_t* __restrict__ fl8 = kargs.fl8;

        const int64_t nseg = (int64_t)N * N * (N / kTile);
        int64_t sgi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (sgi >= nseg) return;
        const uint2 v = *reinterpret_cast<const uint2*>(A + sgi * kTile);
        uint32_t m = 0;
    #pragma unroll
        for (int b = 0; b < 4; ++b) {
            m |= (((v.x >> (8 * b)) & 0xFFu) == (uint32_t)kFluid ? 1u : 0u) << b;
            m |= (((v.y >> (8 * b)) & 0xFFu) == (uint32_t)kFluid ? 1u : 0u) << (b + 4);
        }
        fl8[sgi] = (uint8_t)m;
    }
};

constexpr int kLine = 64 * kPad;          // floats of one per-line LDS array
constexpr int kHalo = 2 * kTile * kPad;   // two halo faces (minus / plus side) of 8 lines

// Each wave owns its tile and its LDS region, so the wavefront steps only need
// the wave's own LDS writes to be visible to its other lanes: a wavefront-scope
// release/acquire pair (waits for the wave's LDS operations) around a wave
// barrier, instead of a barrier over the whole block.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// MIC(0) apply, one pass over the tiles of one colour; 4 tiles per 256-thread
// block, one wave per tile, lane (il, jl) owning the z-line of 8 cells.
//   PH 0 (red, forward):       q = L^-1 r inside the tile          -> Q
//   PH 1 (black, fwd + bwd):   forward with red halos, backward     -> Z, z.r partials
//   PH 2 (red, backward):      backward with black halos            -> Z, z.r partials
//   PH 3 (block, fwd + bwd):   no cross-tile coupling               -> Z, z.r partials
// forward:  q_i = (r_i + sum_{j lower} pc_j q_j) pc_i   (order -x,-y,-z,+x,+y,+z)
// backward: z_i = (q_i + sum_{k upper} pc_i z_k) pc_i   (order +x,+y,+z,-x,-y,-z)
// The sweeps are wavefronts over il+jl+kl (22 steps). Every step is branch-free:
// the neighbour values kept in LDS for the sums (pc*q forward, z backward) are
// -0.0 for cells that are not FLUID, and absent neighbours (tile edges without
// a halo) read a line of -0.0, so each term is added unconditionally:
// x + (-0.0) == x for every x, and pc * (-0.0) == -0.0 for pc >= 0, which
// leaves every sum bit-identical to one that skips the term.
// Lane addresses are set up once as base - (il+jl); step st then reads base + st
// (forward) or base + 21 - st (backward) with the step in the instruction's
// immediate offset. Lanes outside their 8 cells at a step read inside the
// guarded block array and do not write.
constexpr int kSteps = 3 * kTile - 2;                // 22
constexpr int kGuardLo = 16;                          // >= 14 = max il+jl
constexpr int kWaveFloats = 3 * kLine + 2 * kHalo + kTile;
constexpr int kGuardHi = 32;                          // >= 21 past the last region

template <int PH>
struct KMicApply {
    static constexpr int kThreads = 256;
    struct Args {
        Params p;
        const uint8_t* fl8;
        const float* pc;
        const float* r;
        float* Q;
        float* Z;
        const int32_t* tiles;
        int64_t ntiles;
        Reduce red;
        int slot_base;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const uint8_t* __restrict__ fl8 = kargs.fl8;
        const float* __restrict__ pc = kargs.pc;
        const float* __restrict__ r = kargs.r;
        float* __restrict__ Q = kargs.Q;
        float* __restrict__ Z = kargs.Z;
        const int32_t* __restrict__ tiles = kargs.tiles;
        int64_t ntiles = kargs.ntiles;
        const Reduce& red = kargs.red;
        int slot_base = kargs.slot_base;

        constexpr bool do_fwd = PH != 2;
        constexpr bool do_bwd = PH != 0;
        constexpr bool halo_fwd = PH == 1;
        constexpr bool halo_bwd = PH == 2;
        constexpr bool halo = halo_fwd || halo_bwd;
        __shared__ float lds[kGuardLo + kTilesPerBlock * kWaveFloats + kGuardHi];
        const int N = p.N;
        const int TN = N / kTile;
        const int NS = N / kTile;   // segments per z-line
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        // float offsets of this wave's arrays inside lds[]
        const int o_sa = kGuardLo + wave * kWaveFloats;   // r / q, then q / z      [line*9 + kl]
        const int o_pc = o_sa + kLine;                    // pc
        const int o_sb = o_pc + kLine;                    // pc*q (forward), z (backward); -0 off-fluid
        const int o_hx = o_sb + kLine;                    // x halo [(il==7)*72 + jl*9 + kl]
        const int o_hy = o_hx + kHalo;                    // y halo [(jl==7)*72 + il*9 + kl]
        const int o_zero = o_hy + kHalo;                  // 8 x -0.0
        int ti = 0, tj = 0, tk = 0;
        ntiles = tile_count(ntiles, red.sc, PH == 1 ? 1 : (PH == 3 ? 2 : 0));
        if ((int64_t)blockIdx.x >= mic_grid(ntiles)) return;
        if (red.sc->flex && !red.sc->use_mic) return;   // flexible PCG still on the network
        if (slot_base < 0) slot_base = (int)mic_grid(tile_count(-1, red.sc, 1));
        const int64_t idx = (int64_t)blockIdx.x * kTilesPerBlock + wave;
        const bool active = idx < ntiles;
        // the tile index and the stop flag are fetched together
        const int t = active ? tiles[idx] : 0;
        if (red.sc->done) return;
        if (active) {
            tk = t % TN;
            tj = (t / TN) % TN;
            ti = t / (TN * TN);
        }
        const int il = lane >> 3, jl = lane & 7;
        const int gi = ti * kTile + il, gj = tj * kTile + jl, k0 = tk * kTile;
        uint32_t fl = 0;
        float hzm = -0.0f, hzp = -0.0f;   // z-halo terms (-0.0 when absent)
        float rreg[kTile];
        // coalesced mapping: round m covers line (m, lane>>3), element lane&7
        const int ce = lane & 7, cj = lane >> 3;
        if (lane < kTile) lds[o_zero + lane] = -0.0f;
        if (active) {
            fl = fl8[((int64_t)gi * N + gj) * NS + tk];
    #pragma unroll
            for (int m = 0; m < kTile; ++m) {
                const int64_t g = cidx(N, ti * kTile + m, tj * kTile + cj, k0 + ce);
                const int64_t gt = tile_major(t, m, cj, ce);
                const int li = (m * 8 + cj) * kPad + ce;
                lds[o_pc + li] = pc[gt];
                rreg[m] = r[g];
                lds[o_sa + li] = do_fwd ? rreg[m] : Q[gt];
            }
            if constexpr (halo) {
                // neighbour lines across the tile faces: forward halos hold pc*Q of
                // the red neighbours, backward halos their Z; -0.0 off-fluid
                // (pc and Q tile-major, Z natural)
                const float* va = halo_fwd ? pc : Z;
                const float* vb = halo_fwd ? Q : nullptr;
                const int TT = TN * TN;
                // lanes on a tile face own one halo line; at the domain boundary it
                // holds -0.0
                if (il == 0 || il == 7) {
                    float* d = lds + o_hx + (il == 7) * kTile * kPad + jl * kPad;
    #pragma unroll
                    for (int kl = 0; kl < kTile; ++kl) d[kl] = -0.0f;
                }
                if (jl == 0 || jl == 7) {
                    float* d = lds + o_hy + (jl == 7) * kTile * kPad + il * kPad;
    #pragma unroll
                    for (int kl = 0; kl < kTile; ++kl) d[kl] = -0.0f;
                }
                int xi = -1;
                if (il == 0 && gi > 0) xi = gi - 1;
                if (il == 7 && gi + 1 < N) xi = gi + 1;
                if (xi >= 0) {
                    const int64_t ln = halo_fwd ? tile_major(t + (xi < gi ? -TT : TT), xi & 7, jl, 0)
                                                : cidx(N, xi, gj, k0);
                    const uint32_t hf = fl8[((int64_t)xi * N + gj) * NS + tk];
                    float* d = lds + o_hx + (il == 7) * kTile * kPad + jl * kPad;
                    const float4 a0 = *reinterpret_cast<const float4*>(va + ln);
                    const float4 a1 = *reinterpret_cast<const float4*>(va + ln + 4);
                    float av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
                    if (vb) {
                        const float4 b0 = *reinterpret_cast<const float4*>(vb + ln);
                        const float4 b1 = *reinterpret_cast<const float4*>(vb + ln + 4);
                        float bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
    #pragma unroll
                        for (int kl = 0; kl < kTile; ++kl) av[kl] = av[kl] * bv[kl];
                    }
    #pragma unroll
                    for (int kl = 0; kl < kTile; ++kl) d[kl] = ((hf >> kl) & 1u) ? av[kl] : -0.0f;
                }
                int yj = -1;
                if (jl == 0 && gj > 0) yj = gj - 1;
                if (jl == 7 && gj + 1 < N) yj = gj + 1;
                if (yj >= 0) {
                    const int64_t ln = halo_fwd ? tile_major(t + (yj < gj ? -TN : TN), il, yj & 7, 0)
                                                : cidx(N, gi, yj, k0);
                    const uint32_t hf = fl8[((int64_t)gi * N + yj) * NS + tk];
                    float* d = lds + o_hy + (jl == 7) * kTile * kPad + il * kPad;
                    const float4 a0 = *reinterpret_cast<const float4*>(va + ln);
                    const float4 a1 = *reinterpret_cast<const float4*>(va + ln + 4);
                    float av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
                    if (vb) {
                        const float4 b0 = *reinterpret_cast<const float4*>(vb + ln);
                        const float4 b1 = *reinterpret_cast<const float4*>(vb + ln + 4);
                        float bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
    #pragma unroll
                        for (int kl = 0; kl < kTile; ++kl) av[kl] = av[kl] * bv[kl];
                    }
    #pragma unroll
                    for (int kl = 0; kl < kTile; ++kl) d[kl] = ((hf >> kl) & 1u) ? av[kl] : -0.0f;
                }
                // z neighbours: values and fluid bits loaded together, masked after
                const int64_t line = cidx(N, gi, gj, k0);
                const int64_t sgl = ((int64_t)gi * N + gj) * NS + tk;
                const bool has_m = k0 > 0, has_p = k0 + kTile < N;
                int64_t lm = has_m ? line - 1 : line, lp = has_p ? line + kTile : line;
                if (halo_fwd) {
                    lm = has_m ? tile_major(t - 1, il, jl, kTile - 1) : tile_major(t, il, jl, 0);
                    lp = has_p ? tile_major(t + 1, il, jl, 0) : tile_major(t, il, jl, 0);
                }
                const uint32_t fm = fl8[has_m ? sgl - 1 : sgl], fp = fl8[has_p ? sgl + 1 : sgl];
                const float vm = vb ? va[lm] * vb[lm] : va[lm];
                const float vp = vb ? va[lp] * vb[lp] : va[lp];
                if (has_m && ((fm >> 7) & 1u)) hzm = vm;
                if (has_p && (fp & 1u)) hzp = vp;
            }
        }
        wave_sync();
        const int sl = il + jl;
        const int me = lane * kPad;
        const uint32_t fls = fl << sl;    // bit (kl + sl) = FLUID bit of cell kl
        const int a_sa = o_sa + me - sl, a_pc = o_pc + me - sl, a_sb = o_sb + me - sl;
        const int a_zero = o_zero - sl;
        if constexpr (do_fwd) {
            // lower neighbours: -x, -y, -z (in tile), +x, +y (halo, PH 1)
            const int f_xm = il > 0 ? a_sb - 8 * kPad : (halo_fwd ? o_hx + jl * kPad - sl : a_zero);
            const int f_ym = jl > 0 ? a_sb - kPad : (halo_fwd ? o_hy + il * kPad - sl : a_zero);
            const int f_xp = (halo_fwd && il == 7) ? o_hx + kTile * kPad + jl * kPad - sl : a_zero;
            const int f_yp = (halo_fwd && jl == 7) ? o_hy + kTile * kPad + il * kPad - sl : a_zero;
    #pragma unroll
            for (int st = 0; st < kSteps; ++st) {
                float acc = lds[a_sa + st];
                acc = acc + lds[f_xm + st];
                acc = acc + lds[f_ym + st];
                const float vzm = lds[a_sb - 1 + st];         // sb[kl - 1]
                acc = acc + (sl == st ? hzm : vzm);
                if constexpr (halo_fwd) {
                    acc = acc + lds[f_xp + st];
                    acc = acc + lds[f_yp + st];
                    acc = acc + (sl + 7 == st ? hzp : -0.0f);
                }
                const float pcv = lds[a_pc + st];
                const bool fluid = (fls >> st) & 1u;
                const float q = fluid ? acc * pcv : 0.0f;
                const float pq = fluid ? pcv * q : -0.0f;
                if ((unsigned)(st - sl) < (unsigned)kTile && active) {
                    lds[a_sa + st] = q;
                    lds[a_sb + st] = pq;
                }
                wave_sync();
            }
        }
        if constexpr (do_bwd) {
            // upper neighbours: +x, +y, +z (in tile), -x, -y, -z (halo, PH 2)
            const int b_xp = il < 7 ? a_sb + 8 * kPad
                                    : (halo_bwd ? o_hx + kTile * kPad + jl * kPad - sl : a_zero);
            const int b_yp = jl < 7 ? a_sb + kPad
                                    : (halo_bwd ? o_hy + kTile * kPad + il * kPad - sl : a_zero);
            const int b_xm = (halo_bwd && il == 0) ? o_hx + jl * kPad - sl : a_zero;
            const int b_ym = (halo_bwd && jl == 0) ? o_hy + il * kPad - sl : a_zero;
            const float hzp_b = halo_bwd ? hzp : -0.0f;
    #pragma unroll
            for (int st = 0; st < kSteps; ++st) {
                constexpr int kLast = kSteps - 1;             // kl = 21 - st - sl
                const int o = kLast - st;
                const float pci = lds[a_pc + o];
                float acc = lds[a_sa + o];
                acc = acc + pci * lds[b_xp + o];
                acc = acc + pci * lds[b_yp + o];
                const float vzp = lds[a_sb + 1 + o];          // sb[kl + 1]
                acc = acc + pci * (sl + 7 == o ? hzp_b : vzp);
                if constexpr (halo_bwd) {
                    acc = acc + pci * lds[b_xm + o];
                    acc = acc + pci * lds[b_ym + o];
                    acc = acc + pci * (sl == o ? hzm : -0.0f);
                }
                const bool fluid = (fls >> o) & 1u;
                const float z = fluid ? acc * pci : 0.0f;
                if ((unsigned)(o - sl) < (unsigned)kTile && active) {
                    lds[a_sa + o] = z;
                    lds[a_sb + o] = fluid ? z : -0.0f;
                }
                wave_sync();
            }
        }
        // coalesced write-back
        double dot[1] = {0.0};
        if (active) {
    #pragma unroll
            for (int m = 0; m < kTile; ++m) {
                const int64_t g = cidx(N, ti * kTile + m, tj * kTile + cj, k0 + ce);
                const int li = (m * 8 + cj) * kPad + ce;
                const float v = lds[o_sa + li];
                if constexpr (do_bwd) {
                    Z[g] = v;
                    dot[0] += (double)v * (double)rreg[m];
                } else {
                    Q[tile_major(t, m, cj, ce)] = v;
                }
            }
        }
        if constexpr (PH == 0) return;
        block_partials<1>(dot, red.partials, slot_base + blockIdx.x);
    }
};

// r_old = r on the active segments (flexible PCG).
struct KCopyList {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float* src;
        float* dst;
        const int32_t* segs;
        int64_t nseg;
        const SolverScalars* sc;
    };
    static __device__ void run(const Args& a) {
        if (a.sc->done) return;
        const int64_t nseg = seg_count(a.nseg, a.sc);
        const int64_t nb = list_grid(nseg);
        if ((int64_t)blockIdx.x >= nb) return;
        const float4* __restrict__ S4 = reinterpret_cast<const float4*>(a.src);
        float4* __restrict__ D4 = reinterpret_cast<float4*>(a.dst);
        for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nseg;
             u += nb * blockDim.x) {
            const int64_t c = 2 * (int64_t)a.segs[u];
            D4[c] = S4[c];
            D4[c + 1] = S4[c + 1];
        }
    }
};

// z.r and z.r_old on the active segments (flexible PCG's Polak-Ribiere beta).
struct KPRDot {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float* z;
        const float* r;
        const float* r_old;
        const int32_t* segs;
        int64_t nseg;
        Reduce red;
        int check;   // 1: only while the network is in use
    };
    static __device__ void run(const Args& a) {
        const SolverScalars* sc = a.red.sc;
        if (sc->done || (a.check && sc->use_mic)) return;
        const int64_t nseg = seg_count(a.nseg, sc);
        const int64_t nb = list_grid(nseg);
        if ((int64_t)blockIdx.x >= nb) return;
        double v[2] = {0.0, 0.0};
        for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nseg;
             u += nb * blockDim.x) {
            const int64_t c0 = 8 * (int64_t)a.segs[u];
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const double zz = a.z[c0 + m];
                v[0] += zz * (double)a.r[c0 + m];
                v[1] += zz * (double)a.r_old[c0 + m];
            }
        }
        block_partials<2>(v, a.red.partials, blockIdx.x);
    }
};

// Jacobi: z = r * (1/diag), rho = z.r over the active segments
struct KJacobiApply {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float* pc;
        const float* r;
        float* Z;
        const int32_t* segs;
        int64_t nseg;
        Reduce red;
    };
    static __device__ void run(const Args& kargs) {
        const float* __restrict__ pc = kargs.pc;
        const float* __restrict__ r = kargs.r;
        float* __restrict__ Z = kargs.Z;
        const int32_t* __restrict__ segs = kargs.segs;
        int64_t nseg = kargs.nseg;
        const Reduce& red = kargs.red;

        if (red.sc->done) return;
        nseg = seg_count(nseg, red.sc);
        const int64_t nb = list_grid(nseg);
        if ((int64_t)blockIdx.x >= nb) return;
        double v[1] = {0.0};
        for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nseg;
             u += nb * blockDim.x) {
            const int64_t c0 = 8 * (int64_t)segs[u];
    #pragma unroll
            for (int m = 0; m < 8; ++m) {
                const int64_t c = c0 + m;
                const float z = r[c] * pc[c];
                Z[c] = z;
                v[0] += (double)z * (double)r[c];
            }
        }
        block_partials<1>(v, red.partials, blockIdx.x);
    }
};

// ------------------------------------------------------------ active sets
struct KSegFlags {
    static constexpr int kThreads = kBlock;
    struct Args {
        int64_t nseg;
        const uint8_t* fl8;
        int32_t* f;
    };
    static __device__ void run(const Args& kargs) {
        int64_t nseg = kargs.nseg;
        const uint8_t* __restrict__ fl8 = kargs.fl8;
        int32_t* __restrict__ f = kargs.f;

        const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (t < nseg) f[t] = fl8[t] != 0 ? 1 : 0;
    }
};

struct KSegCompact {
    static constexpr int kThreads = kBlock;
    struct Args {
        int64_t nseg;
        const int32_t* f;
        const int32_t* pos;
        int32_t* list;
        SolverScalars* sc;
    };
    static __device__ void run(const Args& kargs) {
        int64_t nseg = kargs.nseg;
        const int32_t* __restrict__ f = kargs.f;
        const int32_t* __restrict__ pos = kargs.pos;
        int32_t* __restrict__ list = kargs.list;
        SolverScalars* sc = kargs.sc;

        const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (t >= nseg) return;
        if (f[t]) list[pos[t]] = (int32_t)t;
        if (t == nseg - 1) sc->nseg = (int64_t)pos[t] + f[t];
    }
};

// any[t] = 1 when one of tile t's 64 z-segments holds a FLUID cell
struct KTileAny {
    static constexpr int kThreads = kBlock;
    struct Args {
        int N;
        const uint8_t* fl8;
        int32_t* any;
    };
    static __device__ void run(const Args& kargs) {
        int N = kargs.N;
        const uint8_t* __restrict__ fl8 = kargs.fl8;
        int32_t* __restrict__ any = kargs.any;

        const int TN = N / kTile;
        const int64_t nt = (int64_t)TN * TN * TN;
        const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (t >= nt) return;
        const int tk = (int)(t % TN), tj = (int)((t / TN) % TN), ti = (int)(t / ((int64_t)TN * TN));
        const int NS = N / kTile;
        uint32_t m = 0;
        for (int il = 0; il < kTile; ++il)
            for (int jl = 0; jl < kTile; ++jl)
                m |= fl8[((int64_t)(ti * kTile + il) * N + tj * kTile + jl) * NS + tk];
        any[t] = m != 0 ? 1 : 0;
    }
};

// One block: ascending lists of the active red, black and all tiles
// (lists + 0, + nt, + 2 nt) and their counts.
struct KTileLists {
    static constexpr int kThreads = 1024;
    struct Args {
        int N;
        const int32_t* any;
        int32_t* lists;
        SolverScalars* sc;
    };
    static __device__ void run(const Args& kargs) {
        int N = kargs.N;
        const int32_t* __restrict__ any = kargs.any;
        int32_t* __restrict__ lists = kargs.lists;
        SolverScalars* sc = kargs.sc;

        const int TN = N / kTile;
        const int nt = TN * TN * TN;
        __shared__ int wsum[3][16];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        int base[3] = {0, 0, 0};
        for (int t0 = 0; t0 < nt; t0 += blockDim.x) {
            const int t = t0 + threadIdx.x;
            int f[3] = {0, 0, 0};
            if (t < nt && any[t]) {
                const int tk = t % TN, tj = (t / TN) % TN, ti = t / (TN * TN);
                const int col = (ti + tj + tk) & 1;
                f[col] = 1;
                f[2] = 1;
            }
            int pre[3];
    #pragma unroll
            for (int c = 0; c < 3; ++c) {
                const uint64_t b = __ballot(f[c]);
                pre[c] = __popcll(b & ((1ull << lane) - 1ull));
                if (lane == 0) wsum[c][w] = __popcll(b);
            }
            __syncthreads();
    #pragma unroll
            for (int c = 0; c < 3; ++c) {
                int off = base[c];
                for (int ww = 0; ww < w; ++ww) off += wsum[c][ww];
                if (f[c]) lists[(int64_t)c * nt + off + pre[c]] = t;
                int tot = 0;
                for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) tot += wsum[c][ww];
                base[c] += tot;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            sc->ntile[0] = base[0];
            sc->ntile[1] = base[1];
            sc->ntile[2] = base[2];
        }
    }
};

int stream_blocks(int64_t n) {
    int64_t b = (n + kBlock - 1) / kBlock;
    return (int)(b < kRedBlocks ? (b < 1 ? 1 : b) : kRedBlocks);
}

}  // namespace

int reduce_blocks(int64_t n, int) { return stream_blocks(n); }

void launch_finalize(hipStream_t s, const Reduce& red, int kind, int nparts, double tol = 0.0,
                     int stop_rule = 0, int N = 0) {
    launch<KFinalize>(s, dim3(1), dim3(256), {red, kind, nparts, tol, stop_rule, N});
}

void launch_solver_init(hipStream_t s, SolverScalars* sc, int max_iters, int refine_every, int flex) {
    launch<KSolverInit>(s, dim3(1), dim3(64), {sc, max_iters, refine_every, flex});
}

void launch_precond_setup(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                          float* diag, uint8_t* nup, float* pc) {
    int64_t nc = (int64_t)p.N * p.N * p.N;
    launch<KPrecondSetup>(s, dim3(grid_for(nc, kBlock)), dim3(kBlock), {p, A, L, diag, nup, pc});
}

void launch_precond_build(hipStream_t s, const Params& p, const uint8_t* A, const float* diag,
                          const uint8_t* nup, float* pc, const Active& act, SolverScalars* sc) {
    if (p.precond == 2) return;   // Jacobi: built by setup
    auto nb = [&](int w) {
        const int64_t n = act.ntiles[w] >= 0 ? act.ntiles[w] : act.tile_cap[w];
        return (unsigned)std::max<int64_t>(1, std::min<int64_t>(n, 4096));
    };
    if (p.precond == 0) {
        launch<KMicBuild>(s, dim3(nb(0)), dim3(64), {p, A, diag, nup, pc, 0, act.tiles[0], act.ntiles[0], sc});
        launch<KMicBuild>(s, dim3(nb(1)), dim3(64), {p, A, diag, nup, pc, 1, act.tiles[1], act.ntiles[1], sc});
    } else {
        launch<KMicBuild>(s, dim3(nb(2)), dim3(64), {p, A, diag, nup, pc, -1, act.tiles[2], act.ntiles[2], sc});
    }
    const int64_t nc = (int64_t)p.N * p.N * p.N;
    launch<KPcTileMajor>(s, dim3(nb(2)), dim3(64), {p.N, pc, pc + nc, act.tiles[2], act.ntiles[2], sc});
}

void launch_fluid_segments(hipStream_t s, int N, const uint8_t* A, uint8_t* fl8) {
    const int64_t nseg = (int64_t)N * N * (N / kTile);
    launch<KFluidSegments>(s, dim3(grid_for(nseg, kBlock)), dim3(kBlock), {N, A, fl8});
}

void launch_active_sets(hipStream_t s, int N, const uint8_t* fl8, int32_t* seg_flag,
                        int32_t* seg_scan, int32_t* seg_list, int32_t* tile_any,
                        int32_t* tile_lists, int32_t* scan_sums, SolverScalars* sc) {
    const int64_t nseg = (int64_t)N * N * (N / kTile);
    launch<KSegFlags>(s, dim3(grid_for(nseg, kBlock)), dim3(kBlock), {nseg, fl8, seg_flag});
    launch_scan(s, seg_flag, seg_scan, nseg, scan_sums);
    launch<KSegCompact>(s, dim3(grid_for(nseg, kBlock)), dim3(kBlock), {nseg, (const int32_t*)seg_flag, (const int32_t*)seg_scan, seg_list, sc});
    const int64_t TN = N / kTile;
    const int64_t nt = TN * TN * TN;
    launch<KTileAny>(s, dim3(grid_for(nt, kBlock)), dim3(kBlock), {N, fl8, tile_any});
    launch<KTileLists>(s, dim3(1), dim3(1024), {N, (const int32_t*)tile_any, tile_lists, sc});
}

// Launch geometry for a count that may be -1 (device-side): size by the cap;
// the matching finalize then derives its partial count on the device.
static int64_t seg_or_cap(const Active& a) { return a.nseg >= 0 ? a.nseg : a.seg_cap; }
static int64_t tiles_or_cap(const Active& a, int w) { return a.ntiles[w] >= 0 ? a.ntiles[w] : a.tile_cap[w]; }

void launch_precond_apply(hipStream_t s, const Params& p, const uint8_t* fl8, const float* pc,
                          const float* r, float* q, float* z, const Reduce& red,
                          const Active& act) {
    const bool dev = act.nseg < 0;
    if (p.precond == 2) {
        const int nb = (int)list_grid(seg_or_cap(act));
        launch<KJacobiApply>(s, dim3(nb), dim3(kBlock), {pc, r, z, act.segs, act.nseg, red});
        launch_finalize(s, red, FIN_RHO, dev ? PARTS_LIST : nb);
        return;
    }
    const float* pct = pc + (int64_t)p.N * p.N * p.N;   // the tile-major copy (MIC(0) only)
    if (p.precond == 0) {
        const int nr = (int)mic_grid(tiles_or_cap(act, 0)), nbk = (int)mic_grid(tiles_or_cap(act, 1));
        launch<KMicApply<0>>(s, dim3(nr), dim3(256), {p, fl8, pct, r, q, z, act.tiles[0], act.ntiles[0], red, 0});
        launch<KMicApply<1>>(s, dim3(nbk), dim3(256), {p, fl8, pct, r, q, z, act.tiles[1], act.ntiles[1], red, 0});
        launch<KMicApply<2>>(s, dim3(nr), dim3(256), {p, fl8, pct, r, q, z, act.tiles[0], act.ntiles[0], red, dev ? -1 : nbk});
        launch_finalize(s, red, FIN_RHO, dev ? PARTS_MIC_RB : nbk + nr);
    } else {
        const int nb = (int)mic_grid(tiles_or_cap(act, 2));
        launch<KMicApply<3>>(s, dim3(nb), dim3(256), {p, fl8, pct, r, q, z, act.tiles[2], act.ntiles[2], red, 0});
        launch_finalize(s, red, FIN_RHO, dev ? PARTS_MIC_ALL : nb);
    }
}

void launch_residual0(hipStream_t s, const Params& p, const uint8_t* A, const float* L, float* P,
                      const float* D, float* R, const Reduce& red, double tol, int stop_rule,
                      double* P64, int refine) {
    const int64_t nc = (int64_t)p.N * p.N * p.N;
    const int nb = (int)list_grid(nc);
    launch<KResidual0>(s, dim3(nb), dim3(kBlock), {p, A, L, P, D, R, red, tol, stop_rule, P64, refine});
    launch_finalize(s, red, FIN_RESID0, nb, tol, stop_rule, p.N);
}

void launch_spmv_dot(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                     const float* x, float* y, const Reduce& red, const Active& act) {
    const int64_t nq = 2 * seg_or_cap(act);   // quads
    const int64_t nb = spmv_grid(nq);
    launch<KSpmvDot>(s, dim3((unsigned)xcd_grid(nb)), dim3(kBlock), {p, A, L, x, y, act.segs, act.nseg, spmv_per_thread(nq), red});
    launch_finalize(s, red, FIN_SIGMA, act.nseg < 0 ? PARTS_SPMV : (int)nb);
}

void launch_axpy2(hipStream_t s, float* P, const float* S, float* R, const float* Z,
                  const Reduce& red, const Active& act) {
    const int nb = (int)list_grid(seg_or_cap(act));
    launch<KAxpy2>(s, dim3(nb), dim3(kBlock), {P, S, R, Z, act.segs, act.nseg, red});
    launch_finalize(s, red, FIN_RNORM, act.nseg < 0 ? PARTS_LIST : nb);
}

void launch_reliable_update(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                            float* P, double* P64, const float* D, float* R, const Reduce& red,
                            const Active& act) {
    const int nb = (int)list_grid(seg_or_cap(act));
    launch<KFlush>(s, dim3(nb), dim3(kBlock), {A, P, P64, act.segs, act.nseg, (const SolverScalars*)red.sc});
    launch<KTrueResidual>(s, dim3(nb), dim3(kBlock), {p, A, L, (const float*)P, (const double*)P64, D, R, act.segs, act.nseg, red, 0});
    launch_finalize(s, red, FIN_TRUE, act.nseg < 0 ? PARTS_LIST : nb);
}

void launch_refine_finish(hipStream_t s, const uint8_t* A, float* P, const double* P64,
                          const Active& act, const SolverScalars* sc) {
    launch<KRefineFinish>(s, dim3((unsigned)list_grid(seg_or_cap(act))), dim3(kBlock), {A, P, P64, act.segs, act.nseg, sc});
}

void launch_true_residual_norm(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                               const float* P, const float* D, const Reduce& red,
                               const Active& act) {
    const int nb = (int)list_grid(seg_or_cap(act));
    launch<KTrueResidual>(s, dim3(nb), dim3(kBlock), {p, A, L, P, (const double*)nullptr, D, (float*)nullptr, act.segs, act.nseg, red, 1});
    launch_finalize(s, red, FIN_REPORT, act.nseg < 0 ? PARTS_LIST : nb);
}

void launch_true_residual_norm64(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                                 float* P, double* P64, const float* D, const Reduce& red,
                                 const Active& act) {
    const int nb = (int)list_grid(seg_or_cap(act));
    launch<KFlush>(s, dim3(nb), dim3(kBlock), {A, P, P64, act.segs, act.nseg, (const SolverScalars*)red.sc, 1});
    launch<KTrueResidual>(s, dim3(nb), dim3(kBlock), {p, A, L, (const float*)P, (const double*)P64, D, (float*)nullptr, act.segs, act.nseg, red, 3});
    launch_finalize(s, red, FIN_REPORT64, act.nseg < 0 ? PARTS_LIST : nb);
}

void launch_update_s(hipStream_t s, const float* Z, float* S, const Reduce& red, int first,
                     const Active& act) {
    launch<KUpdateS>(s, dim3((unsigned)list_grid(seg_or_cap(act))), dim3(kBlock), {Z, S, act.segs, act.nseg, (const SolverScalars*)red.sc, first});
}

void launch_copy_list(hipStream_t s, const float* src, float* dst, const Active& act,
                      const SolverScalars* sc) {
    launch<KCopyList>(s, dim3((unsigned)list_grid(seg_or_cap(act))), dim3(kBlock), {src, dst, act.segs, act.nseg, sc});
}

void launch_pr_dot(hipStream_t s, const float* z, const float* r, const float* r_old,
                   const Reduce& red, const Active& act, bool check) {
    const int nb = (int)list_grid(seg_or_cap(act));
    launch<KPRDot>(s, dim3(nb), dim3(kBlock), {z, r, r_old, act.segs, act.nseg, red, check ? 1 : 0});
    launch_finalize(s, red, check ? FIN_PR_CHECK : FIN_PR, act.nseg < 0 ? PARTS_LIST : nb);
}

void launch_spmv_plain(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                       const float* x, float* y) {
    const int64_t nc = (int64_t)p.N * p.N * p.N;
    launch<KSpmvPlain>(s, dim3(grid_for(nc, kBlock)), dim3(kBlock), {p, A, L, x, y});
}

}  // namespace ariel
