// kernels_pcg.hip — pressure solve for gfx950: matrix-free 7-point operator,
// FP32 vectors with FP64 reductions, and MIC(0) under a tiled red-black ordering.
//
// PCG state never leaves the device inside an iteration: every reduction writes
// per-block FP64 partials, and a one-block finalize kernel sums them in a fixed
// order and updates the scalars (alpha, beta, residual, stop flag) in
// SolverScalars. Later kernels read those scalars and return early once `done`
// is set, so the host can queue iterations in chunks without synchronising.
// The iterations run over ascending lists of the active z-segments and 8^3 tiles
// (those holding a FLUID cell), built once per solve. The iteration vectors (r,
// z, s, y = A s, the iterate p and the preconditioner's intermediate q) are kept
// tile-major: tile t = (ti*TN + tj)*TN + tk holds its 512 cells contiguously,
// cell (il, jl, kl) at 512 t + 64 il + 8 jl + kl, so every pass over the tiles of
// one colour reads and writes whole 2 KB runs. The right-hand side, the cell
// classes, the level set and the pressure handed back stay in the natural layout.
//
// MIC(0) (Bridson, "Fluid Simulation for Computer Graphics", ch. 5; the
// reference's BuildPreconditioner/ApplyPreconditioner, solver.inl:104-286) is
// factored under an ordering where 8^3 tiles are coloured red/black by
// (ti+tj+tk)&1, red tiles precede black ones and cells inside a tile follow the
// natural order. A black tile's cross-tile neighbours are all red (lower), a red
// tile's all black (upper). One 64-lane wave owns a tile, lane = (il,jl) owns the
// z-line of 8 cells, and the in-tile triangular solves run as a 22-step
// wavefront over i+j+k through LDS.
#include <type_traits>

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
// Counts passed as -1 are read from the tank's device scalars (batched launches,
// which are recorded before the counts exist).
__device__ __forceinline__ int64_t seg_count(int64_t n, const SolverScalars* sc) {
    return n >= 0 ? n : sc->nseg;
}
__device__ __forceinline__ int64_t tile_count(int64_t n, const SolverScalars* sc, int which) {
    return n >= 0 ? n : (int64_t)sc->ntile[which];
}

// Tile-major index of tile t's cell (il, jl, kl), and of the natural cell (i, j, k).
__host__ __device__ inline int64_t tile_major(int64_t t, int il, int jl, int kl) {
    return t * 512 + (il * 8 + jl) * 8 + kl;
}
__device__ __forceinline__ int64_t tm_cell(int N, int i, int j, int k) {
    const int TN = N / kTile;
    const int64_t t = ((int64_t)(i / kTile) * TN + j / kTile) * TN + k / kTile;
    return tile_major(t, i % kTile, j % kTile, k % kTile);
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

// One value per block into component `comp` of the two-value layout
// (part[2 slot + comp]); the other component is left to another kernel.
__device__ void block_partial_comp(double v, double* part, int slot, int comp) {
    double t[1] = {v};
    __shared__ double sm1[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    const double ws = wave_sum(t[0]);
    if (lane == 0) sm1[w] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int ww = 0; ww < nw; ++ww) a += sm1[ww];
        part[2 * slot + comp] = a;
    }
    __syncthreads();
}

__device__ __forceinline__ bool finite(double x) { return isfinite(x); }

// Sum of n one-value block partials by a block of 256 threads, in KFinalize's
// order (thread t adds part[t], part[t + 256], ... in turn, then the waves' trees,
// then the four waves in order); every thread gets it. Sixteen loads per thread
// are issued before their adds (adding 0.0 past the end changes no sum: the
// running sum starts at +0.0 and never becomes -0.0).
__device__ double block_partial_sum(const double* __restrict__ part, int64_t n) {
    __shared__ double sm[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double a = 0.0;
    for (int64_t b0 = 0; b0 < n; b0 += 16 * 256) {
        double v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int64_t b = b0 + threadIdx.x + k * 256;
            v[k] = b < n ? part[b] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) a += v[k];
    }
    a = wave_sum(a);
    if (lane == 0) sm[w] = a;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < 4; ++i) t += sm[i];
    __syncthreads();
    return t;
}

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
// The residual goes to the tile-major Rt, the iterate to the tile-major Pt (0 off
// FLUID). With reliable updates (refine): x64 = P, r = b - A x64 in FP64, and Pt
// (the correction to x64) is cleared by a separate fill.
struct KResidual0 {
    static constexpr int kThreads = kBlock;
    struct Args {
        Params p;
        const uint8_t* A;
        const float* L;
        float* P;
        const float* D;
        float* R;      // tile-major
        Reduce red;
        double* P64;
        int refine;
        float* Pt;     // tile-major
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const uint8_t* __restrict__ A = kargs.A;
        const float* __restrict__ L = kargs.L;
        float* __restrict__ P = kargs.P;
        const float* __restrict__ D = kargs.D;
        float* __restrict__ R = kargs.R;
        const Reduce& red = kargs.red;
        double* __restrict__ P64 = kargs.P64;
        int refine = kargs.refine;
        float* __restrict__ Pt = kargs.Pt;

        const int N = p.N;
        const int64_t nc = (int64_t)N * N * N;
        const float n = (float)N;
        const float h = 1.0f / (n * n);
        const int64_t nb = list_grid(nc);
        if ((int64_t)blockIdx.x >= nb) return;
        double v[2] = {0.0, 0.0};
        for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
             c += nb * blockDim.x) {
            int i, j, k;
            decode_cell(N, c, i, j, k);
            const int64_t ct = tm_cell(N, i, j, k);
            // P is read below only at FLUID cells, so clearing it off FLUID here
            // races with no reader
            if (A[c] != kFluid) {
                P[c] = 0.0f;
                if (refine) P64[c] = 0.0;
                else Pt[ct] = 0.0f;
                R[ct] = 0.0f;
                continue;
            }
            const float d = D[c];
            v[0] += (double)d * (double)d;
            if (refine) {
                const double r = (double)d - spmv_cell_f64(A, L, P, N, i, j, k, c);
                P64[c] = (double)P[c];
                R[ct] = (float)r;
                v[1] += r * r;
            } else {
                const float ax = spmv_cell(A, L, P, N, i, j, k, c, h);
                const float r = d + (-1.0f) * ax;
                R[ct] = r;
                Pt[ct] = P[c];
                v[1] += (double)r * (double)r;
            }
        }
        block_partials<2>(v, red.partials, blockIdx.x);
    }
};

// Waits for the wave's own LDS writes and makes them visible to its other lanes
// (each wave owns its LDS region): a wavefront-scope release/acquire pair around
// a wave barrier, instead of a barrier over the whole block.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kTilesPerBlock = 4;
__host__ __device__ inline int64_t mic_grid(int64_t ntiles) {
    const int64_t nb = (ntiles + kTilesPerBlock - 1) / kTilesPerBlock;
    return nb < 1 ? 1 : nb;
}

// One XRef term (solver.inl:138-154) of FLUID cell c with value xc: out of range
// -> xc (the clamped index is c itself); FLUID neighbour -> its value xq; SOLID
// -> xc; AIR -> L_q / min(1e-6, L_c) * xc. q = c + dq in the natural layout.
// The same term from the ghost level set G (L, with NaN at SOLID cells): the
// neighbour's gq and the cell's gc = L_c (a FLUID cell is never SOLID).
__device__ __forceinline__ float xref_term(bool in_range, bool nb_fluid, float xq, float gq,
                                           float gc, float xc) {
    if (!in_range) return xc;
    if (nb_fluid) return xq;
    if (gq != gq) return xc;   // SOLID
    return gq / gmin(1.0e-6f, gc) * xc;
}

// 8 floats of a tile-major line: s = z + b * s_prev, b = 0 for the first
// direction (s = z up to the sign of a zero z: -0 + 0 gives +0, which changes
// no value the solve computes) and beta after. s_prev is finite (zero-filled
// at creation, then written by earlier solves).
__device__ __forceinline__ void search_line(const float* __restrict__ Z, const float* __restrict__ Sp,
                                            int64_t off, float b, float (&x)[8]) {
    const float4 z0 = *reinterpret_cast<const float4*>(Z + off);
    const float4 z1 = *reinterpret_cast<const float4*>(Z + off + 4);
    const float4 s0 = *reinterpret_cast<const float4*>(Sp + off);
    const float4 s1 = *reinterpret_cast<const float4*>(Sp + off + 4);
    const float z[8] = {z0.x, z0.y, z0.z, z0.w, z1.x, z1.y, z1.z, z1.w};
    const float sp[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
#pragma unroll
    for (int kl = 0; kl < 8; ++kl) x[kl] = z[kl] + b * sp[kl];
}
__device__ __forceinline__ float search_cell(const float* __restrict__ Z, const float* __restrict__ Sp,
                                             int64_t off, float b) {
    return Z[off] + b * Sp[off];
}
__device__ __forceinline__ void load_line(const float* __restrict__ p, float (&v)[8]) {
    const float4 a = *reinterpret_cast<const float4*>(p);
    const float4 b = *reinterpret_cast<const float4*>(p + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

// The ghost level set of a solve, tile-major over every tile: G = L, except NaN
// at SOLID cells, so that one float tells the operator's ghost term of a
// non-FLUID neighbour (SOLID: s_c; AIR: L_q / min(1e-6, L_c) s_c) and gives a
// FLUID cell its own L_c; and the SOLID bits of every tile-major line (byte
// off/8 for the line at tile-major offset off). One z-line of 8 cells per thread.
struct KGhostLevelSet {
    static constexpr int kThreads = kBlock;
    struct Args {
        int N;
        const uint8_t* A;
        const float* L;
        float* G;
        uint8_t* sb;
    };
    static __device__ void run(const Args& a) {
        const int N = a.N, TN = N / kTile;
        const int64_t nlines = (int64_t)N * N * TN;
        const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= nlines) return;
        // e enumerates tile-major lines: tile e/64, line e%64 = (il, jl)
        const int64_t t = e >> 6;
        const int line = (int)(e & 63), il = line >> 3, jl = line & 7;
        const int tk = (int)(t % TN), tj = (int)((t / TN) % TN), ti = (int)(t / ((int64_t)TN * TN));
        const int64_t nat = cidx(N, ti * kTile + il, tj * kTile + jl, tk * kTile);
        const uint2 cls = *reinterpret_cast<const uint2*>(a.A + nat);
        float v[8];
        load_line(a.L + nat, v);
        const float nan = __builtin_nanf("");
        uint32_t solid = 0;
#pragma unroll
        for (int kl = 0; kl < 8; ++kl) {
            const uint32_t c = ((kl < 4 ? cls.x : cls.y) >> (8 * (kl & 3))) & 0xffu;
            if (c == (uint32_t)kSolid) {
                v[kl] = nan;
                solid |= 1u << kl;
            }
        }
        a.sb[e] = (uint8_t)solid;
        float* g = a.G + e * 8;
        *reinterpret_cast<float4*>(g) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4*>(g + 4) = make_float4(v[4], v[5], v[6], v[7]);
    }
};

// The search direction and the operator over the active tiles (one wave per
// tile, lane (il, jl) owning the z-line of 8 cells): s = z on the first
// iteration, else s = z + beta s_prev (the update of s, fused here); y = A s
// (ComputeAx, solver.inl:196-230: y_c = (6 s_c - sum XRef_q) / h, h = 1/(n n),
// terms in the order +x, -x, +y, -y, +z, -z; y = 0 off FLUID); sigma = y.s over
// FLUID cells into the block partials.
// S holds two tile-major buffers: the iteration after `it` completed ones reads
// s_prev from buffer it&1 and writes s to buffer (it+1)&1, so the neighbour
// values that a wave recomputes from z and s_prev never meet another wave's
// stores. Neighbour lines inside the tile come from the wave's LDS; lines across
// tile faces and the z-neighbours at the line ends from the neighbouring tiles.
// A neighbour's value is used only when its FLUID bit is set, and FLUID cells
// lie in active tiles, so values of inactive tiles are never used.
struct KTileSpmv {
    static constexpr int kThreads = 256;
    struct Args {
        Params p;
        const float* G;      // ghost level set (KGhostLevelSet), tile-major
        const uint8_t* sb;   // SOLID bits per tile-major line
        const uint8_t* fl8;
        const float* Z;
        float* S;
        float* Y;
        const int32_t* tiles;
        int64_t ntiles;
        Reduce red;
    };
    static __device__ void run(const Args& a) {
        const SolverScalars* sc = a.red.sc;
        if (sc->done) return;
        const int64_t ntiles = tile_count(a.ntiles, sc, 2);
        // each XCD takes a contiguous run of the list, so the neighbouring tiles
        // whose faces a tile reads were mostly read through the same L2
        const int64_t nbk = mic_grid(ntiles);
        const int64_t bl = xcd_block(nbk);
        if (bl >= nbk) return;
        const int N = a.p.N;
        const int TN = N / kTile, NS = N / kTile;
        const int64_t nc = (int64_t)N * N * N, NN = (int64_t)N * N;
        const int it = sc->it;
        const float b = it == 0 ? 0.0f : (float)sc->beta;
        const float* __restrict__ Z = a.Z;
        const float* __restrict__ Sp = a.S + (int64_t)(it & 1) * nc;
        float* __restrict__ Sn = a.S + (int64_t)((it + 1) & 1) * nc;
        const float* __restrict__ G = a.G;
        const uint8_t* __restrict__ fl8 = a.fl8;
        __shared__ float lds[kTilesPerBlock][64 * 8];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int64_t idx = bl * kTilesPerBlock + wave;
        double dot[1] = {0.0};
        if (idx < ntiles) {
            const int64_t t = a.tiles[idx];
            const int tk = (int)(t % TN), tj = (int)((t / TN) % TN), ti = (int)(t / ((int64_t)TN * TN));
            const int il = lane >> 3, jl = lane & 7;
            const int gi = ti * kTile + il, gj = tj * kTile + jl, k0 = tk * kTile;
            const int64_t own = tile_major(t, il, jl, 0);
            const int64_t sg = ((int64_t)gi * N + gj) * NS + tk;
            const int64_t TT = (int64_t)TN * TN;
            const bool r_xp = gi + 1 < N, r_xm = gi > 0, r_yp = gj + 1 < N, r_ym = gj > 0;
            const bool r_zp = k0 + kTile < N, r_zm = k0 > 0;
            // every global load is issued before the first use: the own line, one
            // x-halo line (il == 7: +x, il == 0: -x), one y-halo line (jl == 7: +y,
            // jl == 0: -y) and the two z neighbours at the line's ends, with their
            // FLUID bits; lanes without a halo line on an axis read their own line
            // instead and discard it
            // branch-free offsets: the +x line (0, jl) of tile t + TN^2 lies
            // 512 TN^2 - 7*64 floats past the own line (7, jl), the -x line the
            // same distance before (0, jl); likewise +-y with 512 TN - 7*8
            const int hx = (il == 7 && r_xp) ? 1 : ((il == 0 && r_xm) ? -1 : 0);
            const int hy = (jl == 7 && r_yp) ? 1 : ((jl == 0 && r_ym) ? -1 : 0);
            const int dzp = r_zp ? 1 : 0, dzm = r_zm ? 1 : 0;
            const int64_t lnx = own + (int64_t)hx * (512 * TT - 7 * 64);
            const int64_t lny = own + (int64_t)hy * (512 * (int64_t)TN - 7 * 8);
            const int64_t lzp = own + (int64_t)dzp * 512;
            const int64_t lzm = own - (int64_t)dzm * (512 - 7);
            const uint32_t fl = fl8[sg];
            const uint32_t fhx = fl8[sg + (int64_t)hx * (NN / kTile)];
            const uint32_t fhy = fl8[sg + (int64_t)hy * NS];
            const uint32_t f_zp = (fl8[sg + dzp] & 1u) & (uint32_t)dzp;
            const uint32_t f_zm = ((fl8[sg - dzm] >> 7) & 1u) & (uint32_t)dzm;
            const uint32_t sl = a.sb[own >> 3];
            const uint32_t shx = a.sb[lnx >> 3], shy = a.sb[lny >> 3];
            const uint32_t s_zp = (a.sb[lzp >> 3] & 1u) & (uint32_t)dzp;
            const uint32_t s_zm = ((a.sb[lzm >> 3] >> 7) & 1u) & (uint32_t)dzm;
            float x[8], hxl[8], hyl[8];
            search_line(Z, Sp, own, b, x);
            search_line(Z, Sp, lnx, b, hxl);
            search_line(Z, Sp, lny, b, hyl);
            const float zp = search_cell(Z, Sp, lzp, b);
            const float zm = search_cell(Z, Sp, lzm, b);
            *reinterpret_cast<float4*>(Sn + own) = make_float4(x[0], x[1], x[2], x[3]);
            *reinterpret_cast<float4*>(Sn + own + 4) = make_float4(x[4], x[5], x[6], x[7]);
            float* w = &lds[wave][0];
            *reinterpret_cast<float4*>(w + lane * 8) = make_float4(x[0], x[1], x[2], x[3]);
            *reinterpret_cast<float4*>(w + lane * 8 + 4) = make_float4(x[4], x[5], x[6], x[7]);
            // FLUID and SOLID bits of the in-tile neighbour lines
            const uint32_t f_xp_in = __shfl(fl, (lane + 8) & 63, 64), f_xm_in = __shfl(fl, (lane - 8) & 63, 64);
            const uint32_t f_yp_in = __shfl(fl, (lane + 1) & 63, 64), f_ym_in = __shfl(fl, (lane - 1) & 63, 64);
            const uint32_t s_xp_in = __shfl(sl, (lane + 8) & 63, 64), s_xm_in = __shfl(sl, (lane - 8) & 63, 64);
            const uint32_t s_yp_in = __shfl(sl, (lane + 1) & 63, 64), s_ym_in = __shfl(sl, (lane - 1) & 63, 64);
            wave_sync();
            float xp[8], xm[8], yp[8], ym[8];
#pragma unroll
            for (int kl = 0; kl < 8; ++kl) {
                xp[kl] = il < 7 ? w[((lane + 8) & 63) * 8 + kl] : hxl[kl];
                xm[kl] = il > 0 ? w[((lane - 8) & 63) * 8 + kl] : hxl[kl];
                yp[kl] = jl < 7 ? w[((lane + 1) & 63) * 8 + kl] : hyl[kl];
                ym[kl] = jl > 0 ? w[((lane - 1) & 63) * 8 + kl] : hyl[kl];
            }
            const uint32_t f_xp = il < 7 ? f_xp_in : (hx == 1 ? fhx : 0u);
            const uint32_t f_xm = il > 0 ? f_xm_in : (hx == -1 ? fhx : 0u);
            const uint32_t f_yp = jl < 7 ? f_yp_in : (hy == 1 ? fhy : 0u);
            const uint32_t f_ym = jl > 0 ? f_ym_in : (hy == -1 ? fhy : 0u);
            const uint32_t s_xp = il < 7 ? s_xp_in : (hx == 1 ? shx : 0u);
            const uint32_t s_xm = il > 0 ? s_xm_in : (hx == -1 ? shx : 0u);
            const uint32_t s_yp = jl < 7 ? s_yp_in : (hy == 1 ? shy : 0u);
            const uint32_t s_ym = jl > 0 ? s_ym_in : (hy == -1 ? shy : 0u);
            const float n = (float)N;
            const float h = 1.0f / (n * n);
            // FLUID or SOLID bits of the z-neighbours along the line
            const uint32_t nzp = ((fl | sl) >> 1 | (f_zp | s_zp) << 7) & 0xffu;
            const uint32_t nzm = ((fl | sl) << 1 | (f_zm | s_zm)) & 0xffu;
            // FLUID cells with an in-range AIR neighbour need the level-set ghost
            // term L_q / min(1e-6, L_c) s_c; every other term is s_q (FLUID
            // neighbour) or s_c (SOLID or out of range)
            const uint32_t need =
                fl & ((r_xp ? ~(f_xp | s_xp) : 0u) | (r_xm ? ~(f_xm | s_xm) : 0u) |
                      (r_yp ? ~(f_yp | s_yp) : 0u) | (r_ym ? ~(f_ym | s_ym) : 0u) |
                      (~nzp & (r_zp ? 0xffu : 0x7fu)) | (~nzm & (r_zm ? 0xffu : 0xfeu))) & 0xffu;
            const bool ghosts = __ballot(need != 0u) != 0ull;
            // y_c = (6 s_c - t1 - ... - t6) / h; dividing by h = 2^-2k (N a power
            // of two) is the exact product with 2^2k
            const bool p2 = is_pow2_float(h);
            const float inv_h = 1.0f / h;
            float y[8];
            auto cells = [&](auto p2_tag, auto ghost_tag) {
                constexpr bool P2 = decltype(p2_tag)::value;
                constexpr bool GHOST = decltype(ghost_tag)::value;
                // the ghost level set of the own line, the in-tile neighbour lines
                // (through the wave's LDS, whose s values are already in
                // registers), the halo lines and the z-ends: only in waves with
                // ghost terms, at the same tile-major offsets as s
                float gcl[8], gxp[8], gxm[8], gyp[8], gym[8], gzp = 0.0f, gzm = 0.0f;
                if constexpr (GHOST) {
                    float ghx[8], ghy[8];
                    load_line(G + own, gcl);
                    load_line(G + lnx, ghx);
                    load_line(G + lny, ghy);
                    gzp = G[lzp];
                    gzm = G[lzm];
                    *reinterpret_cast<float4*>(w + lane * 8) = make_float4(gcl[0], gcl[1], gcl[2], gcl[3]);
                    *reinterpret_cast<float4*>(w + lane * 8 + 4) = make_float4(gcl[4], gcl[5], gcl[6], gcl[7]);
                    wave_sync();
#pragma unroll
                    for (int kl = 0; kl < 8; ++kl) {
                        gxp[kl] = il < 7 ? w[((lane + 8) & 63) * 8 + kl] : ghx[kl];
                        gxm[kl] = il > 0 ? w[((lane - 8) & 63) * 8 + kl] : ghx[kl];
                        gyp[kl] = jl < 7 ? w[((lane + 1) & 63) * 8 + kl] : ghy[kl];
                        gym[kl] = jl > 0 ? w[((lane - 1) & 63) * 8 + kl] : ghy[kl];
                    }
                }
#pragma unroll
                for (int kl = 0; kl < 8; ++kl) {
                    const bool fc = (fl >> kl) & 1u;
                    const float xc = x[kl];
                    const int ku = kl < 7 ? kl + 1 : 7, kd = kl > 0 ? kl - 1 : 0;
                    float t1, t2, t3, t4, t5, t6;
                    if constexpr (GHOST) {
                        if (!fc) {
                            y[kl] = 0.0f;
                            continue;
                        }
                        const float gc = gcl[kl];
                        t1 = xref_term(r_xp, (f_xp >> kl) & 1u, xp[kl], gxp[kl], gc, xc);
                        t2 = xref_term(r_xm, (f_xm >> kl) & 1u, xm[kl], gxm[kl], gc, xc);
                        t3 = xref_term(r_yp, (f_yp >> kl) & 1u, yp[kl], gyp[kl], gc, xc);
                        t4 = xref_term(r_ym, (f_ym >> kl) & 1u, ym[kl], gym[kl], gc, xc);
                        t5 = kl < 7 ? xref_term(true, (fl >> ku) & 1u, x[ku], gcl[ku], gc, xc)
                                    : xref_term(r_zp, f_zp, zp, gzp, gc, xc);
                        t6 = kl > 0 ? xref_term(true, (fl >> kd) & 1u, x[kd], gcl[kd], gc, xc)
                                    : xref_term(r_zm, f_zm, zm, gzm, gc, xc);
                    } else {
                        // no AIR neighbour: s_q for a FLUID neighbour, else s_c
                        t1 = (f_xp >> kl) & 1u ? xp[kl] : xc;
                        t2 = (f_xm >> kl) & 1u ? xm[kl] : xc;
                        t3 = (f_yp >> kl) & 1u ? yp[kl] : xc;
                        t4 = (f_ym >> kl) & 1u ? ym[kl] : xc;
                        t5 = kl < 7 ? ((fl >> ku) & 1u ? x[ku] : xc) : (f_zp ? zp : xc);
                        t6 = kl > 0 ? ((fl >> kd) & 1u ? x[kd] : xc) : (f_zm ? zm : xc);
                    }
                    const float acc = 6.0f * xc - t1 - t2 - t3 - t4 - t5 - t6;
                    const float yc = P2 ? acc * inv_h : acc / h;
                    y[kl] = fc ? yc : 0.0f;
                    dot[0] += fc ? (double)yc * (double)xc : 0.0;
                }
            };
            if (ghosts) {
                if (p2) cells(std::true_type{}, std::true_type{});
                else cells(std::false_type{}, std::true_type{});
            } else {
                if (p2) cells(std::true_type{}, std::false_type{});
                else cells(std::false_type{}, std::false_type{});
            }
            *reinterpret_cast<float4*>(a.Y + own) = make_float4(y[0], y[1], y[2], y[3]);
            *reinterpret_cast<float4*>(a.Y + own + 4) = make_float4(y[4], y[5], y[6], y[7]);
        }
        block_partials<1>(dot, a.red.sig, (int)bl);
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

// Elementwise passes over the active tiles: work item e is line e%64 of the
// list's tile e/64 (8 tile-major floats, two float4), so consecutive threads
// touch consecutive 32-byte lines.
__host__ __device__ inline int64_t tile_lines_grid(int64_t ntiles) { return list_grid(64 * ntiles); }
__device__ __forceinline__ int64_t line_offset(const int32_t* __restrict__ tiles, int64_t e) {
    return (int64_t)tiles[e >> 6] * 512 + (e & 63) * 8;
}
// The search direction of the iteration after `it` completed ones (KTileSpmv).
__device__ __forceinline__ const float* search_dir(const float* S, const SolverScalars* sc, int64_t nc) {
    return S + (int64_t)((sc->it + 1) & 1) * nc;
}

// p += alpha s; r += (-alpha) y (Op, solver.inl:157-177); ||r||^2; stop test.
// Cells of active tiles off FLUID hold zero in s, y, p and r and stay zero.
struct KAxpy2 {
    static constexpr int kThreads = kBlock;
    struct Args {
        float* P;
        const float* S;   // both search-direction buffers
        float* R;
        const float* Y;
        const int32_t* tiles;
        int64_t ntiles;
        int64_t nc;
        Reduce red;
    };
    static __device__ void run(const Args& kargs) {
        const Reduce& red = kargs.red;
        if (red.sc->done) return;
        const int64_t ntl = tile_count(kargs.ntiles, red.sc, 2);
        const int64_t nb = tile_lines_grid(ntl);
        if ((int64_t)blockIdx.x >= nb) return;
        const float af = (float)red.sc->alpha;
        const float naf = -af;
        float* __restrict__ P = kargs.P;
        float* __restrict__ R = kargs.R;
        const float* __restrict__ S = search_dir(kargs.S, red.sc, kargs.nc);
        const float* __restrict__ Y = kargs.Y;
        double v[1] = {0.0};
        for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 64 * ntl;
             e += nb * blockDim.x) {
            const int64_t o = line_offset(kargs.tiles, e);
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                float4 p = *reinterpret_cast<float4*>(P + o + 4 * hh);
                float4 r = *reinterpret_cast<float4*>(R + o + 4 * hh);
                const float4 sv = *reinterpret_cast<const float4*>(S + o + 4 * hh);
                const float4 y = *reinterpret_cast<const float4*>(Y + o + 4 * hh);
                p.x = p.x + af * sv.x; p.y = p.y + af * sv.y; p.z = p.z + af * sv.z; p.w = p.w + af * sv.w;
                r.x = r.x + naf * y.x; r.y = r.y + naf * y.y; r.z = r.z + naf * y.z; r.w = r.w + naf * y.w;
                *reinterpret_cast<float4*>(P + o + 4 * hh) = p;
                *reinterpret_cast<float4*>(R + o + 4 * hh) = r;
                v[0] += (double)r.x * (double)r.x + (double)r.y * (double)r.y +
                        (double)r.z * (double)r.z + (double)r.w * (double)r.w;
            }
        }
        block_partials<1>(v, red.partials, blockIdx.x);
    }
};

// Reliable update, part 1: x64 += p, p = 0 (every refine_every iterations);
// x64 natural, p tile-major.
struct KFlush {
    static constexpr int kThreads = kBlock;
    struct Args {
        int N;
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
            int i, j, k;
            decode_cell(kargs.N, c, i, j, k);
            const int64_t ct = tm_cell(kargs.N, i, j, k);
            if (A[c] == kFluid) P64[c] += (double)P[ct];
            P[ct] = 0.0f;
        }
    }
};

// mode 0 (reliable update, part 2): r = b - A x64 in FP64, stored in FP32 into the
// tile-major R, then the stop test. mode 1 (report): ||b - A p||^2 of the FP32
// pressure (natural) into sc->tmp. mode 3 (report): ||b - A x64||^2 of the FP64
// solution into sc->tmp64.
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
            int i, j, k;
            decode_cell(N, c, i, j, k);
            if (A[c] != kFluid) {
                if (mode == 0) R[tm_cell(N, i, j, k)] = 0.0f;
                continue;
            }
            const double ax = mode != 1 ? spmv_cell_f64d(A, L, P64, N, i, j, k, c)
                                        : spmv_cell_f64(A, L, P, N, i, j, k, c);
            const double r = (double)D[c] - ax;
            v[0] += r * r;
            if (mode == 0) R[tm_cell(N, i, j, k)] = (float)r;
        }
        block_partials<1>(v, red.partials, blockIdx.x);
    }
};

// End of a solve: the pressure (natural) from the iterate (tile-major) on the
// active segments: P = (float)(x64 + p) with reliable updates, else P = p; 0 off
// FLUID.
struct KRefineFinish {
    static constexpr int kThreads = kBlock;
    struct Args {
        int N;
        const uint8_t* A;
        float* P;
        const float* Pt;
        const double* P64;   // null: no reliable updates
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
            int i, j, k;
            decode_cell(kargs.N, c, i, j, k);
            const float pt = kargs.Pt[tm_cell(kargs.N, i, j, k)];
            if (A[c] != kFluid) P[c] = 0.0f;
            else P[c] = P64 ? (float)(P64[c] + (double)pt) : pt;
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
enum { PARTS_LIST = -1, PARTS_TILE_LINES = -2, PARTS_MIC_RB = -3, PARTS_MIC_ALL = -4, PARTS_TILES = -5 };
__device__ __forceinline__ int64_t parts_from_counts(int code, const SolverScalars* sc) {
    switch (code) {
        case PARTS_LIST: return list_grid(sc->nseg);
        case PARTS_TILE_LINES: return tile_lines_grid(sc->ntile[2]);
        case PARTS_TILES: return sc->ntile[2] > 0 ? sc->ntile[2] : 1;
        case PARTS_MIC_RB: return mic_grid(sc->ntile[1]) + mic_grid(sc->ntile[0]);
        default: return mic_grid(sc->ntile[2]);
    }
}

enum { FIN_RESID0 = 0, FIN_SIGMA = 1, FIN_RNORM = 2, FIN_TRUE = 3, FIN_REPORT = 4, FIN_RHO = 5,
       FIN_PR = 6, FIN_PR_CHECK = 7, FIN_REPORT64 = 8, FIN_RHO_RNORM = 9 };
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

// sigma = s.As: alpha = rho / sigma, or the breakdown statuses (3: not finite,
// 2: sigma <= 0) that end the solve.
__device__ __forceinline__ bool sigma_breakdown(const SolverScalars* sc, double sigma) {
    return !finite(sigma) || !finite(sc->rho) || !(sigma > 0.0);
}
__device__ void set_alpha(SolverScalars* sc, double sigma) {
    sc->sigma = sigma;
    if (!finite(sigma) || !finite(sc->rho)) {
        sc->status = 3;
        sc->done = 1;
    } else if (!(sigma > 0.0)) {
        sc->status = 2;
        sc->done = 1;
    } else {
        sc->alpha = sc->rho / sigma;
    }
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
             kind == FIN_PR_CHECK || kind == FIN_RHO_RNORM) && sc->done) return;
        if (kind == FIN_TRUE && (sc->done || !refine_due(sc))) return;
        if (kind == FIN_RHO && sc->flex) return;          // flexible PCG sets beta itself
        if (kind == FIN_PR_CHECK && (!sc->flex || sc->use_mic)) return;
        if (nparts < 0) nparts = (int)parts_from_counts(nparts, sc);
        if (kind == FIN_SIGMA) {
            // shared with the first MIC(0) pass, which folds this finalize in
            const double sg = block_partial_sum(red.partials, nparts);
            if (threadIdx.x == 0) set_alpha(sc, sg);
            return;
        }
        const int NV = (kind == FIN_RESID0 || kind == FIN_PR || kind == FIN_PR_CHECK ||
                        kind == FIN_RHO || kind == FIN_RHO_RNORM) ? 2 : 1;
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
                finish_rho(sc, s1);
                break;
            case FIN_RHO_RNORM:
                // the AXPY rode on the apply: ||r||^2 and the stop test of the
                // iteration, then rho_new = z.r and beta (a converged iteration's
                // z is never used)
                sc->rnorm2 = s0;
                sc->it += 1;
                decide(sc);
                finish_rho(sc, s1);
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

// The preconditioner applies read a tile-major copy of pc, made once per solve
// over the active tiles (in the natural layout a tile's z-line is 32 bytes
// between lines of the other colour's tiles, so a pass over one colour would
// fetch whole cache lines half of which it does not use).
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

// MIC(0) apply, one pass over the tiles of one colour; 4 tiles per 256-thread
// block, one wave per tile, lane (il, jl) owning the z-line of 8 cells. r, q, z
// and the factor are tile-major.
// AXPY (passes 0 and 1, the first pass over each colour): the iteration's
// p += alpha s and r += (-alpha) y (Op, solver.inl:157-177) for the tile's cells
// as they are loaded, r written back, and ||r||^2 into component 0 of the block
// partials; the stop test then follows the apply (FIN_RHO_RNORM).
// Partials use the two-value layout: component 0 ||r||^2 (AXPY passes),
// component 1 z.r (passes 1, 2, 3); black blocks take slots [0, nb), red
// blocks [nb, nb + nr).
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

template <int PH, bool AXPY = false>
struct KMicApply {
    static_assert(!AXPY || PH == 0 || PH == 1, "the AXPY rides on the first pass over a colour");
    static constexpr int kThreads = 256;
    struct Args {
        Params p;
        const uint8_t* fl8;
        const float* pc;   // tile-major factor
        float* r;
        float* Q;
        float* Z;
        const int32_t* tiles;
        int64_t ntiles;
        Reduce red;
        int slot_base;
        float* P;          // AXPY: iterate, both search-direction buffers, y
        const float* S;
        const float* Y;
        int64_t nc;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const uint8_t* __restrict__ fl8 = kargs.fl8;
        const float* __restrict__ pc = kargs.pc;
        float* __restrict__ r = kargs.r;
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
        // logical block: each XCD takes a contiguous run of the list, so the
        // neighbouring tiles whose faces a tile reads were mostly handled by the
        // same XCD (its L2); the grid is xcd_grid of the pass's block count
        const int64_t nb_arg = mic_grid(tile_count(ntiles, red.sc, PH == 1 ? 1 : (PH == 3 ? 2 : 0)));
        const int64_t bl = xcd_block(nb_arg);
        const int64_t idx = bl * kTilesPerBlock + wave;
        // the tile index is fetched together with the scalars that gate the pass
        // (every list has room for all TN^3 tiles, so the clamped read is in bounds)
        const int64_t nt_all = (int64_t)TN * TN * TN;
        const int t_pre = tiles[idx < nt_all ? idx : nt_all - 1];
        ntiles = tile_count(ntiles, red.sc, PH == 1 ? 1 : (PH == 3 ? 2 : 0));
        if (bl >= nb_arg) return;
        if (red.sc->flex && !red.sc->use_mic) return;   // flexible PCG still on the network
        if (slot_base < 0) slot_base = (int)mic_grid(tile_count(-1, red.sc, 1));
        const bool active = idx < ntiles;
        const int t = active ? t_pre : 0;
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
        double rr = 0.0;                  // ||r||^2 of the tile after the AXPY
        // the first pass with the AXPY takes alpha from the SpMV's partials itself
        // (no finalize launch between them): every block sums them in KFinalize's
        // order, so all get the same sigma; logical block 0 records it, and on a
        // breakdown every block returns
        constexpr bool kFoldSigma = AXPY && PH == 0;
        float af = 0.0f;
        float pcv[kTile], sav[kTile], pv[kTile], sv[kTile], yv[kTile];
        float hxv[kTile], hyv[kTile], vm = -0.0f, vp = -0.0f;
        uint32_t hfx = 0, hfy = 0, fm = 0, fp = 0;
        int xi = -1, yj = -1;
        bool has_m = false, has_p = false;
        // coalesced mapping: round m covers line (m, lane>>3), element lane&7
        const int ce = lane & 7, cj = lane >> 3;
        if (lane < kTile) lds[o_zero + lane] = -0.0f;
        if (active) {
            // every global load of the pass is issued before the first LDS store:
            // own data, then one x-halo line, one y-halo line and the two z
            // neighbours per lane (lanes without a halo line on an axis read their
            // own line instead and discard it)
            const int64_t sgl = ((int64_t)gi * N + gj) * NS + tk;
            fl = fl8[sgl];
            const float* __restrict__ Sd = AXPY ? search_dir(kargs.S, red.sc, kargs.nc) : nullptr;
    #pragma unroll
            for (int m = 0; m < kTile; ++m) {
                const int64_t gt = tile_major(t, m, cj, ce);
                pcv[m] = pc[gt];
                rreg[m] = r[gt];
                sav[m] = do_fwd ? rreg[m] : Q[gt];
                if constexpr (AXPY) {
                    pv[m] = kargs.P[gt];
                    sv[m] = Sd[gt];
                    yv[m] = kargs.Y[gt];
                }
            }
            if constexpr (halo) {
                // neighbour lines across the tile faces: forward halos hold pc*Q of
                // the red neighbours, backward halos their Z; -0.0 off-fluid
                const float* va = halo_fwd ? pc : Z;
                const float* vb = halo_fwd ? Q : nullptr;
                const int TT = TN * TN;
                const int64_t own = tile_major(t, il, jl, 0);
                // branch-free offsets (dx, dy in {-1, 0, 1}): the +x line (0, jl) of
                // tile t + TN^2 lies 512 TN^2 - 7*64 floats past the own line (7, jl),
                // the -x line the same distance before (0, jl); likewise +-y
                const int dx = (il == 7 && gi + 1 < N) ? 1 : ((il == 0 && gi > 0) ? -1 : 0);
                const int dy = (jl == 7 && gj + 1 < N) ? 1 : ((jl == 0 && gj > 0) ? -1 : 0);
                xi = dx != 0 ? gi + dx : -1;
                yj = dy != 0 ? gj + dy : -1;
                has_m = k0 > 0;
                has_p = k0 + kTile < N;
                const int dzm = has_m ? 1 : 0, dzp = has_p ? 1 : 0;
                const int64_t lnx = own + (int64_t)dx * (512 * (int64_t)TT - 7 * 64);
                const int64_t lny = own + (int64_t)dy * (512 * (int64_t)TN - 7 * 8);
                const int64_t lm = own - (int64_t)dzm * (512 - (kTile - 1));
                const int64_t lp = own + (int64_t)dzp * 512;
                hfx = fl8[sgl + (int64_t)dx * ((int64_t)N * NS)];
                hfy = fl8[sgl + (int64_t)dy * NS];
                fm = fl8[sgl - dzm];
                fp = fl8[sgl + dzp];
                const float4 x0 = *reinterpret_cast<const float4*>(va + lnx);
                const float4 x1 = *reinterpret_cast<const float4*>(va + lnx + 4);
                const float4 y0 = *reinterpret_cast<const float4*>(va + lny);
                const float4 y1 = *reinterpret_cast<const float4*>(va + lny + 4);
                const float zmv = va[lm], zpv = va[lp];
                hxv[0] = x0.x; hxv[1] = x0.y; hxv[2] = x0.z; hxv[3] = x0.w;
                hxv[4] = x1.x; hxv[5] = x1.y; hxv[6] = x1.z; hxv[7] = x1.w;
                hyv[0] = y0.x; hyv[1] = y0.y; hyv[2] = y0.z; hyv[3] = y0.w;
                hyv[4] = y1.x; hyv[5] = y1.y; hyv[6] = y1.z; hyv[7] = y1.w;
                vm = zmv;
                vp = zpv;
                if (vb) {
                    const float4 bx0 = *reinterpret_cast<const float4*>(vb + lnx);
                    const float4 bx1 = *reinterpret_cast<const float4*>(vb + lnx + 4);
                    const float4 by0 = *reinterpret_cast<const float4*>(vb + lny);
                    const float4 by1 = *reinterpret_cast<const float4*>(vb + lny + 4);
                    const float bzm = vb[lm], bzp = vb[lp];
                    const float bxv[8] = {bx0.x, bx0.y, bx0.z, bx0.w, bx1.x, bx1.y, bx1.z, bx1.w};
                    const float byv[8] = {by0.x, by0.y, by0.z, by0.w, by1.x, by1.y, by1.z, by1.w};
    #pragma unroll
                    for (int kl = 0; kl < kTile; ++kl) {
                        hxv[kl] = hxv[kl] * bxv[kl];
                        hyv[kl] = hyv[kl] * byv[kl];
                    }
                    vm = vm * bzm;
                    vp = vp * bzp;
                }
            }
        }
        if constexpr (kFoldSigma) {
            // after the pass's loads are issued, by every thread of the block
            const double sg = block_partial_sum(red.sig, mic_grid(tile_count(-1, red.sc, 2)));
            if (bl == 0 && threadIdx.x == 0) set_alpha(red.sc, sg);
            if (sigma_breakdown(red.sc, sg)) return;
            af = (float)(red.sc->rho / sg);
        } else if constexpr (AXPY) {
            af = (float)red.sc->alpha;
        }
        if (active) {
            if constexpr (AXPY) {
                const float naf = -af;
    #pragma unroll
                for (int m = 0; m < kTile; ++m) {
                    const int64_t gt = tile_major(t, m, cj, ce);
                    pv[m] = pv[m] + af * sv[m];
                    rreg[m] = rreg[m] + naf * yv[m];
                    sav[m] = rreg[m];
                    kargs.P[gt] = pv[m];
                    r[gt] = rreg[m];
                    rr += (double)rreg[m] * (double)rreg[m];
                }
            }
    #pragma unroll
            for (int m = 0; m < kTile; ++m) {
                const int li = (m * 8 + cj) * kPad + ce;
                lds[o_pc + li] = pcv[m];
                lds[o_sa + li] = sav[m];
            }
            if constexpr (halo) {
                // lanes on a tile face own one halo line; at the domain boundary it
                // holds -0.0
                if (il == 0 || il == 7) {
                    float* d = lds + o_hx + (il == 7) * kTile * kPad + jl * kPad;
    #pragma unroll
                    for (int kl = 0; kl < kTile; ++kl) d[kl] = (xi >= 0 && ((hfx >> kl) & 1u)) ? hxv[kl] : -0.0f;
                }
                if (jl == 0 || jl == 7) {
                    float* d = lds + o_hy + (jl == 7) * kTile * kPad + il * kPad;
    #pragma unroll
                    for (int kl = 0; kl < kTile; ++kl) d[kl] = (yj >= 0 && ((hfy >> kl) & 1u)) ? hyv[kl] : -0.0f;
                }
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
                const int64_t gt = tile_major(t, m, cj, ce);
                const int li = (m * 8 + cj) * kPad + ce;
                const float v = lds[o_sa + li];
                if constexpr (do_bwd) {
                    Z[gt] = v;
                    dot[0] += (double)v * (double)rreg[m];
                } else {
                    Q[gt] = v;
                }
            }
        }
        if constexpr (PH == 0) {
            if constexpr (AXPY) block_partial_comp(rr, red.partials, slot_base + (int)bl, 0);
            return;
        }
        if constexpr (PH == 2) {
            block_partial_comp(dot[0], red.partials, slot_base + (int)bl, 1);
        } else {
            double both[2] = {rr, dot[0]};
            block_partials<2>(both, red.partials, slot_base + (int)bl);
        }
    }
};

// r_old = r on the active tiles (flexible PCG).
struct KCopyList {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float* src;
        float* dst;
        const int32_t* tiles;
        int64_t ntiles;
        const SolverScalars* sc;
    };
    static __device__ void run(const Args& a) {
        if (a.sc->done) return;
        const int64_t ntl = tile_count(a.ntiles, a.sc, 2);
        const int64_t nb = tile_lines_grid(ntl);
        if ((int64_t)blockIdx.x >= nb) return;
        for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 64 * ntl;
             e += nb * blockDim.x) {
            const int64_t o = line_offset(a.tiles, e);
            *reinterpret_cast<float4*>(a.dst + o) = *reinterpret_cast<const float4*>(a.src + o);
            *reinterpret_cast<float4*>(a.dst + o + 4) = *reinterpret_cast<const float4*>(a.src + o + 4);
        }
    }
};

// z.r and z.r_old on the active tiles (flexible PCG's Polak-Ribiere beta).
struct KPRDot {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float* z;
        const float* r;
        const float* r_old;
        const int32_t* tiles;
        int64_t ntiles;
        Reduce red;
        int check;   // 1: only while the network is in use
    };
    static __device__ void run(const Args& a) {
        const SolverScalars* sc = a.red.sc;
        if (sc->done || (a.check && sc->use_mic)) return;
        const int64_t ntl = tile_count(a.ntiles, sc, 2);
        const int64_t nb = tile_lines_grid(ntl);
        if ((int64_t)blockIdx.x >= nb) return;
        double v[2] = {0.0, 0.0};
        for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 64 * ntl;
             e += nb * blockDim.x) {
            const int64_t o = line_offset(a.tiles, e);
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const double zz = a.z[o + m];
                v[0] += zz * (double)a.r[o + m];
                v[1] += zz * (double)a.r_old[o + m];
            }
        }
        block_partials<2>(v, a.red.partials, blockIdx.x);
    }
};

// Jacobi: z = r * (1/diag), rho = z.r over the active tiles (pc tile-major).
struct KJacobiApply {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float* pc;
        const float* r;
        float* Z;
        const int32_t* tiles;
        int64_t ntiles;
        Reduce red;
    };
    static __device__ void run(const Args& kargs) {
        const Reduce& red = kargs.red;
        if (red.sc->done) return;
        const int64_t ntl = tile_count(kargs.ntiles, red.sc, 2);
        const int64_t nb = tile_lines_grid(ntl);
        if ((int64_t)blockIdx.x >= nb) return;
        double v[1] = {0.0};
        for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 64 * ntl;
             e += nb * blockDim.x) {
            const int64_t o = line_offset(kargs.tiles, e);
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const float r = kargs.r[o + m];
                const float z = r * kargs.pc[o + m];
                kargs.Z[o + m] = z;
                v[0] += (double)z * (double)r;
            }
        }
        double both[2] = {0.0, v[0]};
        block_partials<2>(both, red.partials, blockIdx.x);
    }
};

// Layout conversion over the active tiles: to_tiles 1: dst (tile-major) = src
// (natural); 0: dst (natural) = src (tile-major). gate 1: only while flexible
// PCG is on the network and not done (the network's input and output).
struct KTileConvert {
    static constexpr int kThreads = kBlock;
    struct Args {
        int N;
        const float* src;
        float* dst;
        const int32_t* tiles;
        int64_t ntiles;
        const SolverScalars* sc;
        int to_tiles;
        int gate;
    };
    static __device__ void run(const Args& a) {
        if (a.gate && (a.sc->done || !a.sc->flex || a.sc->use_mic)) return;
        const int64_t ntl = tile_count(a.ntiles, a.sc, 2);
        const int64_t nb = tile_lines_grid(ntl);
        if ((int64_t)blockIdx.x >= nb) return;
        const int TN = a.N / kTile;
        for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 64 * ntl;
             e += nb * blockDim.x) {
            const int64_t t = a.tiles[e >> 6];
            const int line = (int)(e & 63), il = line >> 3, jl = line & 7;
            const int tk = (int)(t % TN), tj = (int)((t / TN) % TN), ti = (int)(t / ((int64_t)TN * TN));
            const int64_t nat = cidx(a.N, ti * kTile + il, tj * kTile + jl, tk * kTile);
            const int64_t tmo = t * 512 + line * 8;
            const float* src = a.src + (a.to_tiles ? nat : tmo);
            float* dst = a.dst + (a.to_tiles ? tmo : nat);
            *reinterpret_cast<float4*>(dst) = *reinterpret_cast<const float4*>(src);
            *reinterpret_cast<float4*>(dst + 4) = *reinterpret_cast<const float4*>(src + 4);
        }
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


// ------------------------------------------------------------ multigrid
// A multigrid V(1,1) cycle as the preconditioner of PCG (ARIEL_PRECOND_MG; no
// reference counterpart: the reference's preconditioner is an incomplete
// factorisation, solver.inl:104-135). Level 0 is the solve's grid on its active
// tiles (tile-major r and z, the exact operator with its ghost-fluid diagonal);
// coarser levels are dense grids of edge M_l = ceil(M_{l-1} / 2) down to an
// edge of at most 8. A coarse cell is FLUID when any of its 2x2x2 children is,
// else AIR when any child is AIR, else SOLID; its operator is the 7-point
// Laplacian on spacing 2^l / N with Dirichlet AIR and Neumann SOLID (and
// out-of-range) neighbours. Transfers are piecewise constant: a fine cell takes
// its parent's value, and a coarse cell's right-hand side is the mean of its
// children's residuals (restriction = prolongation transposed / 8). Each level
// smooths once before and once after its coarse correction by damped Jacobi;
// the coarsest level runs symmetric red-black Gauss-Seidel sweeps in one
// workgroup. Jacobi is symmetric and the restriction is the prolongation's
// transpose, so the cycle is a symmetric positive definite operator and plain
// PCG applies. (Trilinear transfers took as many PCG iterations at 8 times the
// memory traffic.)
//
// With z1 = w r / d (Jacobi from zero, d the diagonal scaled by 1/h2), the
// residual r - A z1 = (1 - w_c) r_c + sum_{FLUID q} w_q r_q / d_q needs no
// stored z1; pre-smoothing and restriction are one pass, one work item per
// coarse cell over its 8 children. On level 0 the Jacobi weight is 1 where the
// ghost-fluid terms raise the diagonal above 6 (free-surface cells, diagonals
// up to ~10^3): there a full local solve keeps the coarse correction out of the
// stiff cell, which otherwise makes A z large there and the FP32 recursive
// residual drift from the true one. Both weights keep 2 M - A diagonally
// dominant, so the cycle stays positive definite.
constexpr float kMgOmega = 6.0f / 7.0f;
constexpr int kMgCoarsestSweeps = 6;
__device__ __forceinline__ float mg_omega0(float dg) { return dg > 6.0f ? 1.0f : kMgOmega; }

struct MgLevelArgs {
    int M;                  // edge of the level's grid
    float h2;               // squared spacing, 4^l / N^2
    const uint8_t* type;    // cell classes
    const float* d;         // diagonal count (6 - Neumann neighbours) on FLUID cells
    float* r;               // right-hand side (restricted residual)
    float* z;               // pre-smoothed z (FLUID cells)
    float* zo;              // z after the coarse correction and post-smoothing
};

__device__ __forceinline__ void mg_decode(int M, int64_t c, int& i, int& j, int& k) {
    k = (int)(c % M);
    const int64_t q = c / M;
    j = (int)(q % M);
    i = (int)(q / M);
}

// cell classes of level l from its finer level (A of the solve for l = 1)
struct KMgTypes {
    static constexpr int kThreads = kBlock;
    struct Args {
        const uint8_t* fine;
        int Mf;
        uint8_t* coarse;
        int Mc;
    };
    static __device__ void run(const Args& a) {
        const int64_t n = (int64_t)a.Mc * a.Mc * a.Mc;
        const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= n) return;
        int I, J, K;
        mg_decode(a.Mc, c, I, J, K);
        bool fluid = false, air = false;
        for (int u = 0; u < 8; ++u) {
            const int x = 2 * I + (u >> 2), y = 2 * J + ((u >> 1) & 1), z = 2 * K + (u & 1);
            if (x >= a.Mf || y >= a.Mf || z >= a.Mf) continue;
            const uint8_t t = a.fine[((int64_t)x * a.Mf + y) * a.Mf + z];
            fluid |= t == kFluid;
            air |= t == kAir;
        }
        a.coarse[c] = fluid ? (uint8_t)kFluid : (air ? (uint8_t)kAir : (uint8_t)kSolid);
    }
};

// diagonal counts of a coarse level: 6 less the SOLID or out-of-range neighbours
struct KMgDiag {
    static constexpr int kThreads = kBlock;
    struct Args {
        const uint8_t* type;
        int M;
        float* d;
    };
    static __device__ void run(const Args& a) {
        const int M = a.M;
        const int64_t n = (int64_t)M * M * M;
        const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= n) return;
        int i, j, k;
        mg_decode(M, c, i, j, k);
        float d = 6.0f;
        const int q[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
        for (int m = 0; m < 6; ++m) {
            const int x = q[m][0], y = q[m][1], z = q[m][2];
            const bool in = x >= 0 && x < M && y >= 0 && y < M && z >= 0 && z < M;
            if (!in || a.type[((int64_t)x * M + y) * M + z] == kSolid) d -= 1.0f;
        }
        a.d[c] = a.type[c] == kFluid ? d : 1.0f;
    }
};

// Sum over the 8 consecutive lanes that hold a coarse cell's children (lanes
// 8m .. 8m+7), in a fixed order (pairs, then quads, then the halves); every lane
// of the group gets it.
__device__ __forceinline__ float mg_group8_sum(float v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    return v;
}

// Level 0 works one 8^3 tile per block of 512 threads: each thread computes a
// value of its own cell, threads 0..383 the same value of the cells across the
// tile's 6 faces (64 each: +x, -x, +y, -y, +z, -z, in-plane coordinates in
// (y, z), (y, z), (x, z), (x, z), (x, y), (x, y) order), all into LDS, and then
// each cell reads its 6 neighbours' values from LDS.
constexpr int kMgTileThreads = 512;
__device__ __forceinline__ void mg_halo_cell(int h, int ti, int tj, int tk, int& i, int& j, int& k) {
    const int f = h >> 6, a = (h >> 3) & 7, b = h & 7;
    const int x0 = ti * kTile, y0 = tj * kTile, z0 = tk * kTile;
    switch (f) {
        case 0: i = x0 + kTile; j = y0 + a; k = z0 + b; break;
        case 1: i = x0 - 1; j = y0 + a; k = z0 + b; break;
        case 2: i = x0 + a; j = y0 + kTile; k = z0 + b; break;
        case 3: i = x0 + a; j = y0 - 1; k = z0 + b; break;
        case 4: i = x0 + a; j = y0 + b; k = z0 + kTile; break;
        default: i = x0 + a; j = y0 + b; k = z0 - 1; break;
    }
}
// sum of the 6 neighbours' values of in-tile cell (il, jl, kl): own-tile values
// v[il*64 + jl*8 + kl], face values hv[face*64 + a*8 + b] (0 outside the grid)
__device__ __forceinline__ float mg_nbr_sum(const float* v, const float* hv, int il, int jl, int kl) {
    const int li = il * 64 + jl * 8 + kl;
    float s = il < 7 ? v[li + 64] : hv[0 * 64 + jl * 8 + kl];
    s += il > 0 ? v[li - 64] : hv[1 * 64 + jl * 8 + kl];
    s += jl < 7 ? v[li + 8] : hv[2 * 64 + il * 8 + kl];
    s += jl > 0 ? v[li - 8] : hv[3 * 64 + il * 8 + kl];
    s += kl < 7 ? v[li + 1] : hv[4 * 64 + il * 8 + jl];
    s += kl > 0 ? v[li - 1] : hv[5 * 64 + il * 8 + jl];
    return s;
}

// Level 0, pre-smoothing and restriction, one tile per block: u = w_c r / d on
// FLUID cells (0 elsewhere), res = (1 - w_c) r + sum of the neighbours' u, and
// each coarse cell of the tile's 4^3 block of level 1 takes the mean of its 8
// children's res (its children in 8 consecutive lanes: thread t is child t % 8
// of coarse cell t / 8). Coarse cells off the active tiles have no FLUID child
// and are not written (every reader of r checks the class).
struct KMgSmoothRestrict0 {
    static constexpr int kThreads = kMgTileThreads;
    struct Args {
        int N;
        const uint8_t* A;
        const float* diag;
        const float* R;     // tile-major r
        float* rc;          // level 1 right-hand side, dense
        int Mc;
        const int32_t* tiles;
        int64_t ntiles;
        const SolverScalars* sc;
    };
    static __device__ void run(const Args& a) {
        if (a.sc->done) return;
        const int64_t ntl = tile_count(a.ntiles, a.sc, 2);
        if ((int64_t)blockIdx.x >= ntl) return;
        __shared__ float uv[512], hu[384];
        const int N = a.N, TN = N / kTile, Mc = a.Mc;
        const int t = a.tiles[blockIdx.x];
        const int tk = t % TN, tj = (t / TN) % TN, ti = t / (TN * TN);
        const int tid = threadIdx.x;
        const int lc = tid >> 3, u = tid & 7;
        const int ci = lc >> 4, cj = (lc >> 2) & 3, ck = lc & 3;
        const int il = 2 * ci + (u >> 2), jl = 2 * cj + ((u >> 1) & 1), kl = 2 * ck + (u & 1);
        const int i = ti * kTile + il, j = tj * kTile + jl, k = tk * kTile + kl;
        const int64_t c = cidx(N, i, j, k);
        const bool fl = a.A[c] == kFluid;
        const float r = fl ? a.R[(int64_t)t * 512 + il * 64 + jl * 8 + kl] : 0.0f;
        const float d = a.diag[c];
        const float w = mg_omega0(d);
        uv[il * 64 + jl * 8 + kl] = fl ? w * r / d : 0.0f;
        if (tid < 384) {
            int x, y, z;
            mg_halo_cell(tid, ti, tj, tk, x, y, z);
            float hv = 0.0f;
            if (x >= 0 && x < N && y >= 0 && y < N && z >= 0 && z < N) {
                const int64_t cq = cidx(N, x, y, z);
                if (a.A[cq] == kFluid) {
                    const float dq = a.diag[cq];
                    hv = mg_omega0(dq) * a.R[tm_cell(N, x, y, z)] / dq;
                }
            }
            hu[tid] = hv;
        }
        __syncthreads();
        const float res = fl ? (1.0f - w) * r + mg_nbr_sum(uv, hu, il, jl, kl) : 0.0f;
        const float acc = mg_group8_sum(res);
        if (u == 0) a.rc[((int64_t)(ti * 4 + ci) * Mc + (tj * 4 + cj)) * Mc + (tk * 4 + ck)] = acc * 0.125f;
    }
};

// Levels >= 1, pre-smoothing and restriction to level l+1: work item e is child
// e % 8 of coarse cell e / 8 (8 consecutive lanes); FLUID children get
// z = w h2 r / d, and the mean of their residuals r - A z is the coarse cell's
// right-hand side (0 on non-FLUID coarse cells).
struct KMgSmoothRestrict {
    static constexpr int kThreads = kBlock;
    struct Args {
        MgLevelArgs f;
        MgLevelArgs c;
        const SolverScalars* sc;
    };
    static __device__ void run(const Args& a) {
        if (a.sc->done) return;
        const int Mc = a.c.M, M = a.f.M;
        const int64_t n = 8 * (int64_t)Mc * Mc * Mc;
        const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const bool live = e < n;
        const int64_t cc = live ? e >> 3 : 0;
        const int u = (int)(e & 7);
        int I, J, K;
        mg_decode(Mc, cc, I, J, K);
        const int i = 2 * I + (u >> 2), j = 2 * J + ((u >> 1) & 1), k = 2 * K + (u & 1);
        float res = 0.0f;
        if (live && i < M && j < M && k < M) {
            const int64_t c = ((int64_t)i * M + j) * M + k;
            if (a.f.type[c] == kFluid) {
                const float r = a.f.r[c];
                float sum = 0.0f;
                const int q[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
#pragma unroll
                for (int m = 0; m < 6; ++m) {
                    const int x = q[m][0], y = q[m][1], z = q[m][2];
                    if (x < 0 || x >= M || y < 0 || y >= M || z < 0 || z >= M) continue;
                    const int64_t cq = ((int64_t)x * M + y) * M + z;
                    if (a.f.type[cq] != kFluid) continue;
                    sum += a.f.r[cq] / a.f.d[cq];
                }
                a.f.z[c] = kMgOmega * a.f.h2 * r / a.f.d[c];
                res = (1.0f - kMgOmega) * r + kMgOmega * sum;
            }
        }
        const float acc = mg_group8_sum(res);
        if (live && u == 0) a.c.r[cc] = a.c.type[cc] == kFluid ? acc * 0.125f : 0.0f;
    }
};

// The coarse correction P e at fine cell (x, y, z): its parent's value in the
// coarse level's final z (dense, edge Mc; 0 where the parent is not FLUID).
__device__ __forceinline__ float mg_prolong(const float* __restrict__ e, int Mc, int x, int y, int z) {
    return Mc > 0 ? e[((int64_t)(x >> 1) * Mc + (y >> 1)) * Mc + (z >> 1)] : 0.0f;
}

// Levels >= 1, coarse correction and post-smoothing in one pass: zt = z + P e
// (recomputed for the FLUID neighbours), zo = zt + w (h2 r + sum zt_q - d zt) / d,
// as (1 - w) zt + w (h2 r + sum zt_q) / d.
struct KMgProlongPost {
    static constexpr int kThreads = kBlock;
    struct Args {
        MgLevelArgs l;
        const float* e;     // coarse level's final z
        int Mc;
        const SolverScalars* sc;
    };
    static __device__ void run(const Args& a) {
        if (a.sc->done) return;
        const int M = a.l.M;
        const int64_t n = (int64_t)M * M * M;
        const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= n) return;
        if (a.l.type[c] != kFluid) {
            a.l.zo[c] = 0.0f;
            return;
        }
        int i, j, k;
        mg_decode(M, c, i, j, k);
        const float ztc = a.l.z[c] + mg_prolong(a.e, a.Mc, i, j, k);
        float sum = 0.0f;
        const int q[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const int x = q[m][0], y = q[m][1], z = q[m][2];
            if (x < 0 || x >= M || y < 0 || y >= M || z < 0 || z >= M) continue;
            const int64_t cq = ((int64_t)x * M + y) * M + z;
            if (a.l.type[cq] != kFluid) continue;
            sum += a.l.z[cq] + mg_prolong(a.e, a.Mc, x, y, z);
        }
        const float d = a.l.d[c];
        a.l.zo[c] = (1.0f - kMgOmega) * ztc + kMgOmega * ((a.l.h2 * a.l.r[c] + sum) / d);
    }
};

// Coarsest level (M^3 <= 512 cells): z from r by kMgCoarsestSweeps symmetric
// red-black Gauss-Seidel sweeps (red, then black and red again per sweep, from
// z = 0), in LDS, one thread per cell.
struct KMgCoarsest {
    static constexpr int kThreads = 512;
    struct Args {
        MgLevelArgs l;
        const SolverScalars* sc;
    };
    static __device__ void run(const Args& a) {
        if (a.sc->done) return;
        __shared__ float z[512];
        __shared__ uint8_t ty[512];
        const int M = a.l.M;
        const int n = M * M * M;
        const int c = threadIdx.x;
        int i = 0, j = 0, k = 0;
        float rc = 0.0f, d = 1.0f;
        if (c < n) {
            mg_decode(M, c, i, j, k);
            ty[c] = a.l.type[c];
            rc = a.l.r[c];
            d = a.l.d[c];
            z[c] = 0.0f;
        }
        __syncthreads();
        const bool mine = c < n && ty[c] == kFluid;
        const int colour = (i + j + k) & 1;
        auto half = [&](int col) {
            float v = 0.0f;
            if (mine && colour == col) {
                float sum = 0.0f;
                const int q[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
                for (int m = 0; m < 6; ++m) {
                    const int x = q[m][0], y = q[m][1], zz = q[m][2];
                    if (x < 0 || x >= M || y < 0 || y >= M || zz < 0 || zz >= M) continue;
                    const int cq = (x * M + y) * M + zz;
                    if (ty[cq] == kFluid) sum += z[cq];
                }
                v = (a.l.h2 * rc + sum) / d;
            }
            __syncthreads();
            if (mine && colour == col) z[c] = v;
            __syncthreads();
        };
        half(0);
        for (int s = 0; s < kMgCoarsestSweeps; ++s) {
            half(1);
            half(0);
        }
        if (c < n) a.l.zo[c] = mine ? z[c] : 0.0f;
    }
};

// Level 0, coarse correction and post-smoothing into the tile-major z of PCG (0
// off FLUID), one tile per block: zt = w_c h2 r / d + P e on FLUID cells (0
// elsewhere) for the tile and its faces into LDS, then
// z = zt + w_c (h2 r + sum_{FLUID q} zt_q - d zt) / d, as (1 - w_c) zt +
// w_c (h2 r + sum zt_q) / d; z.r into component 1 of the block's partial
// (two-value layout; component 0 is 0) for FIN_RHO.
struct KMgPost0 {
    static constexpr int kThreads = kMgTileThreads;
    struct Args {
        int N;
        float h2;
        const uint8_t* A;
        const float* diag;
        const float* R;
        const float* e;     // level 1's final z (null with no coarse level)
        int Mc;
        float* Z;
        const int32_t* tiles;
        int64_t ntiles;
        Reduce red;
    };
    static __device__ void run(const Args& a) {
        if (a.red.sc->done) return;
        const int64_t ntl = tile_count(a.ntiles, a.red.sc, 2);
        if ((int64_t)blockIdx.x >= ntl) return;
        __shared__ float zv[512], hz[384];
        const int N = a.N, TN = N / kTile;
        const float h2 = a.h2;
        const int t = a.tiles[blockIdx.x];
        const int tk = t % TN, tj = (t / TN) % TN, ti = t / (TN * TN);
        const int li = threadIdx.x;
        const int il = li >> 6, jl = (li >> 3) & 7, kl = li & 7;
        const int i = ti * kTile + il, j = tj * kTile + jl, k = tk * kTile + kl;
        const int64_t c = cidx(N, i, j, k);
        const int64_t o = (int64_t)t * 512 + li;
        const bool fl = a.A[c] == kFluid;
        const float rc = fl ? a.R[o] : 0.0f;
        const float d = a.diag[c];
        const float w = mg_omega0(d);
        const float ztc = fl ? w * h2 * rc / d + mg_prolong(a.e, a.Mc, i, j, k) : 0.0f;
        zv[li] = ztc;
        if (li < 384) {
            int x, y, z;
            mg_halo_cell(li, ti, tj, tk, x, y, z);
            float hv = 0.0f;
            if (x >= 0 && x < N && y >= 0 && y < N && z >= 0 && z < N) {
                const int64_t cq = cidx(N, x, y, z);
                if (a.A[cq] == kFluid) {
                    const float dq = a.diag[cq];
                    hv = mg_omega0(dq) * h2 * a.R[tm_cell(N, x, y, z)] / dq + mg_prolong(a.e, a.Mc, x, y, z);
                }
            }
            hz[li] = hv;
        }
        __syncthreads();
        double v[2] = {0.0, 0.0};
        float out = 0.0f;
        if (fl) {
            out = (1.0f - w) * ztc + w * ((h2 * rc + mg_nbr_sum(zv, hz, il, jl, kl)) / d);
            v[1] = (double)out * (double)rc;
        }
        a.Z[o] = out;
        block_partials<2>(v, a.red.partials, blockIdx.x);
    }
};
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
    auto nb = [&](int w) {
        const int64_t n = act.ntiles[w] >= 0 ? act.ntiles[w] : act.tile_cap[w];
        return (unsigned)std::max<int64_t>(1, std::min<int64_t>(n, 4096));
    };
    if (p.precond == 0) {
        launch<KMicBuild>(s, dim3(nb(0)), dim3(64), {p, A, diag, nup, pc, 0, act.tiles[0], act.ntiles[0], sc});
        launch<KMicBuild>(s, dim3(nb(1)), dim3(64), {p, A, diag, nup, pc, 1, act.tiles[1], act.ntiles[1], sc});
    } else if (p.precond != 2) {   // Jacobi's 1/diag comes from the set-up
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
                          float* r, float* q, float* z, const Reduce& red,
                          const Active& act, const IterAxpy* ax) {
    const bool dev = act.nseg < 0;
    const float* pct = pc + (int64_t)p.N * p.N * p.N;   // the tile-major copy
    if (p.precond == 2) {
        const int nb = (int)tile_lines_grid(tiles_or_cap(act, 2));
        launch<KJacobiApply>(s, dim3(nb), dim3(kBlock), {pct, r, z, act.tiles[2], act.ntiles[2], red});
        launch_finalize(s, red, FIN_RHO, dev ? PARTS_TILE_LINES : nb);
        return;
    }
    if (p.precond == 0) {
        const int nr = (int)mic_grid(tiles_or_cap(act, 0)), nbk = (int)mic_grid(tiles_or_cap(act, 1));
        const int red_base = dev ? -1 : nbk;
        if (ax) {
            launch<KMicApply<0, true>>(s, dim3((unsigned)xcd_grid(nr)), dim3(256), {p, fl8, pct, r, q, z, act.tiles[0], act.ntiles[0], red, red_base, ax->P, ax->S, ax->Y, ax->nc});
            launch<KMicApply<1, true>>(s, dim3((unsigned)xcd_grid(nbk)), dim3(256), {p, fl8, pct, r, q, z, act.tiles[1], act.ntiles[1], red, 0, ax->P, ax->S, ax->Y, ax->nc});
        } else {
            launch<KMicApply<0>>(s, dim3((unsigned)xcd_grid(nr)), dim3(256), {p, fl8, pct, r, q, z, act.tiles[0], act.ntiles[0], red, red_base, nullptr, nullptr, nullptr, 0});
            launch<KMicApply<1>>(s, dim3((unsigned)xcd_grid(nbk)), dim3(256), {p, fl8, pct, r, q, z, act.tiles[1], act.ntiles[1], red, 0, nullptr, nullptr, nullptr, 0});
        }
        launch<KMicApply<2>>(s, dim3((unsigned)xcd_grid(nr)), dim3(256), {p, fl8, pct, r, q, z, act.tiles[0], act.ntiles[0], red, red_base, nullptr, nullptr, nullptr, 0});
        launch_finalize(s, red, ax ? FIN_RHO_RNORM : FIN_RHO, dev ? PARTS_MIC_RB : nbk + nr);
    } else {
        const int nb = (int)mic_grid(tiles_or_cap(act, 2));
        launch<KMicApply<3>>(s, dim3((unsigned)xcd_grid(nb)), dim3(256), {p, fl8, pct, r, q, z, act.tiles[2], act.ntiles[2], red, 0, nullptr, nullptr, nullptr, 0});
        launch_finalize(s, red, FIN_RHO, dev ? PARTS_MIC_ALL : nb);
    }
}

void launch_residual0(hipStream_t s, const Params& p, const uint8_t* A, const float* L, float* P,
                      const float* D, float* Rt, const Reduce& red, double tol, int stop_rule,
                      double* P64, int refine, float* Pt) {
    const int64_t nc = (int64_t)p.N * p.N * p.N;
    const int nb = (int)list_grid(nc);
    launch<KResidual0>(s, dim3(nb), dim3(kBlock), {p, A, L, P, D, Rt, red, P64, refine, Pt});
    launch_finalize(s, red, FIN_RESID0, nb, tol, stop_rule, p.N);
}

void launch_ghost_levelset(hipStream_t s, int N, const uint8_t* A, const float* L, float* G,
                           uint8_t* sb) {
    const int64_t nlines = (int64_t)N * N * (N / kTile);
    launch<KGhostLevelSet>(s, dim3(grid_for(nlines, kBlock)), dim3(kBlock), {N, A, L, G, sb});
}

void launch_spmv_dot(hipStream_t s, const Params& p, const float* G, const uint8_t* sb,
                     const uint8_t* fl8, const float* Zt, float* St, float* Yt, const Reduce& red,
                     const Active& act, bool fold_sigma) {
    const int nb = (int)mic_grid(tiles_or_cap(act, 2));
    launch<KTileSpmv>(s, dim3((unsigned)xcd_grid(nb)), dim3(256), {p, G, sb, fl8, Zt, St, Yt, act.tiles[2], act.ntiles[2], red});
    if (fold_sigma) return;
    Reduce rs = red;
    rs.partials = red.sig;
    launch_finalize(s, rs, FIN_SIGMA, act.ntiles[2] < 0 ? PARTS_MIC_ALL : nb);
}

int64_t partial_capacity(int N) {
    const int64_t tn = N / kTile;
    const int64_t nb = tn * tn * tn + 1;   // one block per tile (multigrid level 0) at most
    return nb > kMaxPartials ? nb : kMaxPartials;
}

void launch_axpy2(hipStream_t s, float* Pt, const float* St, float* Rt, const float* Yt, int64_t nc,
                  const Reduce& red, const Active& act) {
    const int nb = (int)tile_lines_grid(tiles_or_cap(act, 2));
    launch<KAxpy2>(s, dim3(nb), dim3(kBlock), {Pt, St, Rt, Yt, act.tiles[2], act.ntiles[2], nc, red});
    launch_finalize(s, red, FIN_RNORM, act.ntiles[2] < 0 ? PARTS_TILE_LINES : nb);
}

void launch_reliable_update(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                            float* Pt, double* P64, const float* D, float* Rt, const Reduce& red,
                            const Active& act) {
    const int nb = (int)list_grid(seg_or_cap(act));
    launch<KFlush>(s, dim3(nb), dim3(kBlock), {p.N, A, Pt, P64, act.segs, act.nseg, (const SolverScalars*)red.sc, 0});
    launch<KTrueResidual>(s, dim3(nb), dim3(kBlock), {p, A, L, (const float*)nullptr, (const double*)P64, D, Rt, act.segs, act.nseg, red, 0});
    launch_finalize(s, red, FIN_TRUE, act.nseg < 0 ? PARTS_LIST : nb);
}

void launch_refine_finish(hipStream_t s, int N, const uint8_t* A, float* P, const float* Pt,
                          const double* P64, const Active& act, const SolverScalars* sc) {
    launch<KRefineFinish>(s, dim3((unsigned)list_grid(seg_or_cap(act))), dim3(kBlock), {N, A, P, Pt, P64, act.segs, act.nseg, sc});
}

void launch_true_residual_norm(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                               const float* P, const float* D, const Reduce& red,
                               const Active& act) {
    const int nb = (int)list_grid(seg_or_cap(act));
    launch<KTrueResidual>(s, dim3(nb), dim3(kBlock), {p, A, L, P, (const double*)nullptr, D, (float*)nullptr, act.segs, act.nseg, red, 1});
    launch_finalize(s, red, FIN_REPORT, act.nseg < 0 ? PARTS_LIST : nb);
}

void launch_true_residual_norm64(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                                 float* Pt, double* P64, const float* D, const Reduce& red,
                                 const Active& act) {
    const int nb = (int)list_grid(seg_or_cap(act));
    launch<KFlush>(s, dim3(nb), dim3(kBlock), {p.N, A, Pt, P64, act.segs, act.nseg, (const SolverScalars*)red.sc, 1});
    launch<KTrueResidual>(s, dim3(nb), dim3(kBlock), {p, A, L, (const float*)nullptr, (const double*)P64, D, (float*)nullptr, act.segs, act.nseg, red, 3});
    launch_finalize(s, red, FIN_REPORT64, act.nseg < 0 ? PARTS_LIST : nb);
}

void launch_copy_list(hipStream_t s, const float* src, float* dst, const Active& act,
                      const SolverScalars* sc) {
    launch<KCopyList>(s, dim3((unsigned)tile_lines_grid(tiles_or_cap(act, 2))), dim3(kBlock), {src, dst, act.tiles[2], act.ntiles[2], sc});
}

void launch_pr_dot(hipStream_t s, const float* z, const float* r, const float* r_old,
                   const Reduce& red, const Active& act, bool check) {
    const int nb = (int)tile_lines_grid(tiles_or_cap(act, 2));
    launch<KPRDot>(s, dim3(nb), dim3(kBlock), {z, r, r_old, act.tiles[2], act.ntiles[2], red, check ? 1 : 0});
    launch_finalize(s, red, check ? FIN_PR_CHECK : FIN_PR, act.ntiles[2] < 0 ? PARTS_TILE_LINES : nb);
}

void launch_tile_convert(hipStream_t s, int N, const float* src, float* dst, const Active& act,
                         const SolverScalars* sc, bool to_tiles, bool gate) {
    launch<KTileConvert>(s, dim3((unsigned)tile_lines_grid(tiles_or_cap(act, 2))), dim3(kBlock),
                         {N, src, dst, act.tiles[2], act.ntiles[2], sc, to_tiles ? 1 : 0, gate ? 1 : 0});
}

void launch_spmv_plain(hipStream_t s, const Params& p, const uint8_t* A, const float* L,
                       const float* x, float* y) {
    const int64_t nc = (int64_t)p.N * p.N * p.N;
    launch<KSpmvPlain>(s, dim3(grid_for(nc, kBlock)), dim3(kBlock), {p, A, L, x, y});
}


void mg_plan(int N, MgLevels& mg) {
    mg.M[0] = N;
    int l = 0;
    while (mg.M[l] > 8 && l + 1 < kMgMaxLevels) {
        mg.M[l + 1] = (mg.M[l] + 1) / 2;
        ++l;
    }
    mg.L = l;
}

static MgLevelArgs mg_level(const MgLevels& mg, int N, int l) {
    MgLevelArgs a{};
    a.M = mg.M[l];
    a.h2 = (float)((double)(1u << (2 * l)) / ((double)N * (double)N));
    a.type = mg.type[l];
    a.d = mg.d[l];
    a.r = mg.r[l];
    a.z = mg.z[l];
    a.zo = mg.zo[l];
    return a;
}

static unsigned mg_cells_grid(int M) { return grid_for((int64_t)M * M * M, kBlock); }

void launch_mg_build(hipStream_t s, int N, const uint8_t* A, const MgLevels& mg) {
    for (int l = 1; l <= mg.L; ++l) {
        launch<KMgTypes>(s, dim3(mg_cells_grid(mg.M[l])), dim3(kBlock),
                         {l == 1 ? A : mg.type[l - 1], mg.M[l - 1], mg.type[l], mg.M[l]});
        launch<KMgDiag>(s, dim3(mg_cells_grid(mg.M[l])), dim3(kBlock), {mg.type[l], mg.M[l], mg.d[l]});
    }
}

void launch_mg_apply(hipStream_t s, int N, const uint8_t* A, const float* diag, const MgLevels& mg,
                     const float* Rt, float* Zt, const Reduce& red, const Active& act) {
    const SolverScalars* sc = red.sc;
    const int64_t ntl = tiles_or_cap(act, 2);
    const float h2 = (float)(1.0 / ((double)N * (double)N));
    if (mg.L > 0) {
        launch<KMgSmoothRestrict0>(s, dim3((unsigned)std::max<int64_t>(1, ntl)), dim3(kMgTileThreads),
                                   {N, A, diag, Rt, mg.r[1], mg.M[1], act.tiles[2], act.ntiles[2], sc});
        for (int l = 1; l < mg.L; ++l)
            launch<KMgSmoothRestrict>(s, dim3(grid_for(8 * (int64_t)mg.M[l + 1] * mg.M[l + 1] * mg.M[l + 1], kBlock)),
                                      dim3(kBlock), {mg_level(mg, N, l), mg_level(mg, N, l + 1), sc});
        launch<KMgCoarsest>(s, dim3(1), dim3(512), {mg_level(mg, N, mg.L), sc});
        for (int l = mg.L - 1; l >= 1; --l)
            launch<KMgProlongPost>(s, dim3(mg_cells_grid(mg.M[l])), dim3(kBlock),
                                   {mg_level(mg, N, l), mg.zo[l + 1], mg.M[l + 1], sc});
    }
    const unsigned nb0 = (unsigned)std::max<int64_t>(1, ntl);
    launch<KMgPost0>(s, dim3(nb0), dim3(kMgTileThreads),
                     {N, h2, A, diag, Rt, mg.L > 0 ? mg.zo[1] : nullptr, mg.L > 0 ? mg.M[1] : 0, Zt,
                      act.tiles[2], act.ntiles[2], red});
    launch_finalize(s, red, FIN_RHO, act.ntiles[2] < 0 ? PARTS_TILES : (int)nb0);
}

}  // namespace ariel
