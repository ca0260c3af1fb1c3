// kernels_grid.hip — MAC-grid kernels of the projection for gfx950.
//
// Each kernel is one streaming pass over faces or cells (z-fastest, coalesced);
// the reference's separate loops over x/y/z faces (flip.cpp:324-680) are fused
// where the result of a face depends only on that face and its two cells.
#include "ariel_device.h"
#include "ariel_internal.h"

namespace ariel {

namespace {

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// Faces in rows along k: row r (one wave each) of component comp holds the
// faces (i, j, 0..kext-1), with rows of component 0, then 1, then 2, each in
// (i, j) order, so the row's first face is fidx(N, comp, i, j, 0) = rl * kext
// for the row's index rl within its component. The row index is the same for
// the whole wave, so its decode is scalar code instead of a division per lane.
inline int64_t face_rows(int N) { return 2 * (int64_t)(N + 1) * N + (int64_t)N * N; }
__device__ __forceinline__ bool face_row(int N, int64_t r, int& comp, int& i, int& j, int& kext,
                                         int64_t& f0) {
    const int64_t r01 = (int64_t)(N + 1) * N;   // rows of component 0, and of 1
    if (r < r01) {
        comp = 0; i = (int)(r / N); j = (int)(r % N); kext = N; f0 = r * kext;
    } else if (r < 2 * r01) {
        r -= r01;
        comp = 1; i = (int)(r / (N + 1)); j = (int)(r % (N + 1)); kext = N; f0 = r * kext;
    } else if (r < 2 * r01 + (int64_t)N * N) {
        r -= 2 * r01;
        comp = 2; i = (int)(r / N); j = (int)(r % N); kext = N + 1; f0 = r * kext;
    } else {
        return false;
    }
    return true;
}
constexpr int kRowThreads = 64;

// Cell indices stay below 2^31 for N <= 1024: decode in 32-bit.
__device__ __forceinline__ void decode_cell32(int N, int64_t c64, int& i, int& j, int& k) {
    const uint32_t c = (uint32_t)c64, n = (uint32_t)N;
    k = (int)(c % n);
    const uint32_t q = c / n;
    j = (int)(q % n);
    i = (int)(q / n);
}

// Project, divergence per cell (flip.cpp:417-437) and FlipGrid (solver.inl:42-56):
// D = -((ux(i+1)-ux(i) + uy(j+1)-uy(j) + uz(k+1)-uz(k)) / h), h = 1/maxd.
struct KDivergence {
    static constexpr int kThreads = kBlock;
    struct Args {
        Params p;
        const float* ux;
        const float* uy;
        const float* uz;
        float* D;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const float* __restrict__ ux = kargs.ux;
        const float* __restrict__ uy = kargs.uy;
        const float* __restrict__ uz = kargs.uz;
        float* __restrict__ D = kargs.D;

        const int N = p.N;
        const int64_t nc = (int64_t)N * N * N;
        int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= nc) return;
        int i, j, k;
        decode_cell32(N, c, i, j, k);
        const float h = 1.0f / p.maxd;
        float dv = (ux[fidx(N, 0, i + 1, j, k)] - ux[fidx(N, 0, i, j, k)] + uy[fidx(N, 1, i, j + 1, k)] -
                    uy[fidx(N, 1, i, j, k)] + uz[fidx(N, 2, i, j, k + 1)] - uz[fidx(N, 2, i, j, k)]) / h;
        D[c] = -dv;
    }
};

// SubtractPressureGradient (flip.cpp:568-680, subcell = 1) + EnforceBoundaryVelocity
// (particlegridoperations.inl:45-104) per face.
struct KGradient {
    static constexpr int kThreads = kRowThreads;
    struct Args {
        Params p;
        const float* P;
        const float* L;
        const uint8_t* A;
        float* ux;
        float* uy;
        float* uz;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const float* __restrict__ P = kargs.P;
        const float* __restrict__ L = kargs.L;
        const uint8_t* __restrict__ A = kargs.A;
        float* __restrict__ ux = kargs.ux;
        float* __restrict__ uy = kargs.uy;
        float* __restrict__ uz = kargs.uz;

        const int N = p.N;
        int comp, i, j, kext;
        int64_t f0;
        if (!face_row(N, blockIdx.x, comp, i, j, kext, f0)) return;
        float* u = comp == 0 ? ux : (comp == 1 ? uy : uz);
        for (int k = threadIdx.x; k < kext; k += kRowThreads) face(p, P, L, A, u, comp, i, j, k, f0 + k);
    }
    static __device__ void face(const Params& p, const float* __restrict__ P, const float* __restrict__ L,
                                const uint8_t* __restrict__ A, float* __restrict__ u, int comp, int i,
                                int j, int k, int64_t f) {
        const int N = p.N;
        const int a = comp == 0 ? i : (comp == 1 ? j : k);
        if (a == 0 || a == N) {
            u[f] = 0.0f;
            return;
        }
        const float h = 1.0f / p.maxd;
        const int bi = i - (comp == 0), bj = j - (comp == 1), bk = k - (comp == 2);
        const int64_t cf = cidx(N, i, j, k), cb = cidx(N, bi, bj, bk);
        float pf = P[cf], pb = P[cb];
        const float lf = L[cf], lb = L[cb];
        if (lf * lb < 0.0f) {
            if (lf < 0.0f) pf = P[cf];
            else pf = lf / gmin(1.0e-3f, lb) * P[cb];
            if (lb < 0.0f) pb = P[cb];
            else pb = lb / gmin(1.0e-6f, lf) * P[cf];
        }
        float val = u[f];
        val -= (pf - pb) / h;
        const bool s0 = A[cf] == kSolid, s1 = A[cb] == kSolid;
        u[f] = (s0 != s1) ? 0.0f : val;
    }
};

// ExtrapolateVelocity (flip.cpp:452-566) + SubtractPreviousGrid (365-410). A face
// that touches no FLUID cell but has SOLID (or out-of-domain) cells on both sides
// takes the mean of its same-component neighbours that touch FLUID. Only such
// faces are written and only fluid-touching faces are read, so the update is in
// place without a hazard. Then du = u - prev for every face.
__device__ __forceinline__ bool face_mark(const uint8_t* A, int N, int comp, int i, int j, int k) {
    const int a = comp == 0 ? i : (comp == 1 ? j : k);
    const int bi = i - (comp == 0), bj = j - (comp == 1), bk = k - (comp == 2);
    return (a > 0 && A[cidx(N, bi, bj, bk)] == kFluid) || (a < N && A[cidx(N, i, j, k)] == kFluid);
}
__device__ __forceinline__ bool face_wallmark(const uint8_t* A, int N, int comp, int i, int j, int k) {
    const int a = comp == 0 ? i : (comp == 1 ? j : k);
    const int bi = i - (comp == 0), bj = j - (comp == 1), bk = k - (comp == 2);
    return (a <= 0 || A[cidx(N, bi, bj, bk)] == kSolid) && (a >= N || A[cidx(N, i, j, k)] == kSolid);
}

struct KExtrapolate {
    static constexpr int kThreads = kRowThreads;
    struct Args {
        Params p;
        const uint8_t* A;
        float* ux;
        float* uy;
        float* uz;
        float* dux;
        float* duy;
        float* duz;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const uint8_t* __restrict__ A = kargs.A;
        float* ux = kargs.ux;
        float* uy = kargs.uy;
        float* uz = kargs.uz;
        float* dux = kargs.dux;
        float* duy = kargs.duy;
        float* duz = kargs.duz;

        const int N = p.N;
        int comp, i, j, kext;
        int64_t f0;
        if (!face_row(N, blockIdx.x, comp, i, j, kext, f0)) return;
        float* u = comp == 0 ? ux : (comp == 1 ? uy : uz);
        float* du = comp == 0 ? dux : (comp == 1 ? duy : duz);
        for (int k = threadIdx.x; k < kext; k += kRowThreads) face(A, N, u, du, comp, i, j, k, f0 + k);
    }
    static __device__ void face(const uint8_t* __restrict__ A, int N, float* u, float* du, int comp, int i,
                                int j, int k, int64_t f) {
        float val = u[f];
        if (!face_mark(A, N, comp, i, j, k) && face_wallmark(A, N, comp, i, j, k)) {
            unsigned wsum = 0;
            float sum = 0.0f;
            const int q[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                                 {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
            const int ex = N + (comp == 0), ey = N + (comp == 1), ez = N + (comp == 2);
    #pragma unroll
            for (int m = 0; m < 6; ++m) {
                const int a = q[m][0], b = q[m][1], c = q[m][2];
                if (a >= 0 && a < ex && b >= 0 && b < ey && c >= 0 && c < ez) {
                    if (face_mark(A, N, comp, a, b, c)) {
                        wsum++;
                        sum += u[fidx(N, comp, a, b, c)];
                    }
                }
            }
            if (wsum) {
                val = sum / (float)wsum;
                u[f] = val;
            }
        }
        du[f] = val - du[f];
    }
};

// Post-projection diagnostics: max |div u| over FLUID cells and the FLUID count.
// Grid-stride over a capped grid; one pair of atomics per block.
struct KGridStats {
    static constexpr int kThreads = kBlock;
    struct Args {
        Params p;
        const float* ux;
        const float* uy;
        const float* uz;
        const uint8_t* A;
        SolverScalars* sc;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const float* __restrict__ ux = kargs.ux;
        const float* __restrict__ uy = kargs.uy;
        const float* __restrict__ uz = kargs.uz;
        const uint8_t* __restrict__ A = kargs.A;
        SolverScalars* sc = kargs.sc;

        const int N = p.N;
        const int64_t nc = (int64_t)N * N * N;
        const float h = 1.0f / p.maxd;
        float mx = 0.0f;
        int cnt = 0;
        for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
             c += (int64_t)gridDim.x * blockDim.x) {
            if (A[c] != kFluid) continue;
            int i, j, k;
            decode_cell32(N, c, i, j, k);
            float dv = (ux[fidx(N, 0, i + 1, j, k)] - ux[fidx(N, 0, i, j, k)] +
                        uy[fidx(N, 1, i, j + 1, k)] - uy[fidx(N, 1, i, j, k)] +
                        uz[fidx(N, 2, i, j, k + 1)] - uz[fidx(N, 2, i, j, k)]) / h;
            mx = fmaxf(mx, fabsf(dv));
            cnt += 1;
        }
    #pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mx = fmaxf(mx, __shfl_down(mx, off, 64));
            cnt += __shfl_down(cnt, off, 64);
        }
        __shared__ float smx[16];
        __shared__ int scnt[16];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
        if (lane == 0) {
            smx[w] = mx;
            scnt[w] = cnt;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 1; i < nw; ++i) {
                mx = fmaxf(mx, smx[i]);
                cnt += scnt[i];
            }
            if (cnt) atomicAdd(&sc->fluid_cells, cnt);
            if (mx > 0.0f) atomicMax(&sc->maxdiv_bits, __float_as_uint(mx));
        }
    }
};

struct KStatsReset {
    static constexpr int kThreads = kBlock;
    struct Args {
        SolverScalars* sc;
    };
    static __device__ void run(const Args& kargs) {
        SolverScalars* sc = kargs.sc;

        if (threadIdx.x == 0) {
            sc->fluid_cells = 0;
            sc->maxdiv_bits = 0u;
        }
    }
};

}  // namespace

void launch_divergence(hipStream_t s, const Params& p, float* const u[3], float* D) {
    int64_t nc = (int64_t)p.N * p.N * p.N;
    launch<KDivergence>(s, dim3(grid_for(nc, kBlock)), dim3(kBlock), {p, u[0], u[1], u[2], D});
}

void launch_gradient(hipStream_t s, const Params& p, const float* P, const float* L,
                     const uint8_t* A, float* const u[3]) {
    launch<KGradient>(s, dim3((unsigned)face_rows(p.N)), dim3(kRowThreads), {p, P, L, A, u[0], u[1], u[2]});
}

void launch_extrapolate(hipStream_t s, const Params& p, const uint8_t* A, float* const u[3],
                        float* const du[3]) {
    launch<KExtrapolate>(s, dim3((unsigned)face_rows(p.N)), dim3(kRowThreads),
                         {p, A, u[0], u[1], u[2], du[0], du[1], du[2]});
}

void launch_grid_stats(hipStream_t s, const Params& p, float* const u[3], const uint8_t* A,
                       SolverScalars* sc) {
    int64_t nc = (int64_t)p.N * p.N * p.N;
    launch<KStatsReset>(s, dim3(1), dim3(64), {sc});
    const int64_t nb = grid_for(nc, kBlock) < 1024 ? grid_for(nc, kBlock) : 1024;
    launch<KGridStats>(s, dim3((unsigned)nb), dim3(kBlock), {p, u[0], u[1], u[2], A, sc});
}

}  // namespace ariel
