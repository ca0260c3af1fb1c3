// kernels_particles.hip — particle-side kernels of the FLIP step for gfx950.
//
// Particles live sorted by (cell, id) in float4 pairs (pos.xyz + id bits,
// vel.xyz + density). A cell's particles are contiguous and cells are laid out
// z-fastest, so every neighbour query of the reference — 3x3x3 cells visited
// sx -> sy -> sz, each cell's list in particle-vector order
// (ParticleGrid::GetCellNeighbors, particlegrid.cpp:28-50) — becomes 9 (or 8)
// contiguous index ranges, walked in that same order. Sums are therefore formed
// in the reference's order and agree bit for bit with a serial evaluation.
#include "ariel_device.h"
#include "ariel_internal.h"

namespace ariel {

namespace {

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// ------------------------------------------------------------------ sort
// ParticleGrid::Sort (particlegrid.cpp:140-172) as a counting sort:
// bin (histogram) -> exclusive scan -> scatter -> per-cell order by id -> gather.
// Runs of equal cells among consecutive lanes (the input is the previous sorted
// order, so most of a wave shares a handful of cells): the lowest lane of a run
// is its leader; returns the leader lane and the run length (valid in leaders).
__device__ __forceinline__ void wave_runs(int32_t c, bool valid, int& leader, int& len) {
    const int lane = threadIdx.x & 63;
    const int32_t prev = __shfl_up(c, 1, 64);
    const bool head = valid && (lane == 0 || prev != c);
    const uint64_t heads = __ballot(head);
    const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    leader = 63 - __clzll(heads & upto);
    const uint64_t after = (lane == 63) ? 0ull : (heads >> (lane + 1));
    const uint64_t vmask = __ballot(valid);
    const int nvalid = __popcll(vmask);
    const int next = after ? lane + 1 + __ffsll((unsigned long long)after) - 1 : nvalid;
    len = next - lane;
}

struct KBin {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float4* pos;
        int64_t np;
        int N;
        float maxd;
        int32_t* cell_of;
        int32_t* counts;
    };
    static __device__ void run(const Args& kargs) {
        const float4* __restrict__ pos = kargs.pos;
        int64_t np = kargs.np;
        int N = kargs.N;
        float maxd = kargs.maxd;
        int32_t* __restrict__ cell_of = kargs.cell_of;
        int32_t* __restrict__ counts = kargs.counts;

        int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const bool valid = s < np;
        int32_t c = -1;
        if (valid) {
            float4 p = pos[s];
            int cx = cell_coord(p.x, maxd), cy = cell_coord(p.y, maxd), cz = cell_coord(p.z, maxd);
            c = (int32_t)cidx(N, cx, cy, cz);
            cell_of[s] = c;
        }
        int leader, len;
        wave_runs(c, valid, leader, len);
        if (valid && leader == (int)(threadIdx.x & 63)) atomicAdd(&counts[c], len);
    }
};

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ int32_t block_exclusive_scan(int32_t v, int32_t* total) {
    // 256 threads = 4 waves
    __shared__ int32_t wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int32_t before = 0, tot = 0;
    for (int i = 0; i < kScanThreads / 64; ++i) {
        if (i < w) before += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return before + x - v;
}

struct KScanReduce {
    static constexpr int kThreads = kBlock;
    struct Args {
        const int32_t* in;
        int64_t n;
        int32_t* bsum;
    };
    static __device__ void run(const Args& kargs) {
        const int32_t* __restrict__ in = kargs.in;
        int64_t n = kargs.n;
        int32_t* __restrict__ bsum = kargs.bsum;

        if (blockIdx.x > 0 && (int64_t)blockIdx.x * kScanTile >= n) return;   // beyond this grid
        int64_t base = (int64_t)blockIdx.x * kScanTile;
        int32_t acc = 0;
        for (int t = 0; t < kScanItems; ++t) {
            int64_t i = base + (int64_t)t * kScanThreads + threadIdx.x;
            if (i < n) acc += in[i];
        }
        int32_t tot;
        block_exclusive_scan(acc, &tot);
        if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
    }
};

// Single block: exclusive scan of nb block sums in place, bsum[nb] = total.
struct KScanBsums {
    static constexpr int kThreads = kBlock;
    struct Args {
        int32_t* bsum;
        int nb;
    };
    static __device__ void run(const Args& kargs) {
        int32_t* __restrict__ bsum = kargs.bsum;
        int nb = kargs.nb;

        const int per = (nb + kScanThreads - 1) / kScanThreads;
        const int lo = threadIdx.x * per;
        int32_t acc = 0;
        for (int i = 0; i < per; ++i)
            if (lo + i < nb) acc += bsum[lo + i];
        int32_t tot;
        int32_t ex = block_exclusive_scan(acc, &tot);
        for (int i = 0; i < per; ++i) {
            if (lo + i < nb) {
                int32_t v = bsum[lo + i];
                bsum[lo + i] = ex;
                ex += v;
            }
        }
        if (threadIdx.x == 0) bsum[nb] = tot;
    }
};

struct KScanApply {
    static constexpr int kThreads = kBlock;
    struct Args {
        const int32_t* in;
        int64_t n;
        const int32_t* bsum;
        int32_t* out;
        int nb;
    };
    static __device__ void run(const Args& kargs) {
        const int32_t* __restrict__ in = kargs.in;
        int64_t n = kargs.n;
        const int32_t* __restrict__ bsum = kargs.bsum;
        int32_t* __restrict__ out = kargs.out;
        int nb = kargs.nb;

        if ((int)blockIdx.x >= nb) return;   // beyond this grid
        int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
        int32_t v[kScanItems];
        int32_t acc = 0;
    #pragma unroll
        for (int t = 0; t < kScanItems; ++t) {
            int64_t i = base + t;
            v[t] = (i < n) ? in[i] : 0;
            acc += v[t];
        }
        int32_t tot;
        int32_t ex = block_exclusive_scan(acc, &tot) + bsum[blockIdx.x];
    #pragma unroll
        for (int t = 0; t < kScanItems; ++t) {
            int64_t i = base + t;
            if (i < n) out[i] = ex;
            ex += v[t];
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = bsum[nb];
    }
};

// A run leader reserves len slots at the top of its cell with one atomic
// (counts return to 0 for the next sort); members take consecutive slots in lane
// order. The particle id is written next to its slot for the per-cell ordering.
struct KScatter {
    static constexpr int kThreads = kBlock;
    struct Args {
        const int32_t* cell_of;
        int64_t np;
        const int32_t* offsets;
        int32_t* counts;
        int32_t* slots;
        const float4* pos;
        int32_t* sid;
    };
    static __device__ void run(const Args& kargs) {
        const int32_t* __restrict__ cell_of = kargs.cell_of;
        int64_t np = kargs.np;
        const int32_t* __restrict__ offsets = kargs.offsets;
        int32_t* __restrict__ counts = kargs.counts;
        int32_t* __restrict__ slots = kargs.slots;
        const float4* __restrict__ pos = kargs.pos;
        int32_t* __restrict__ sid = kargs.sid;

        int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const bool valid = s < np;
        const int32_t c = valid ? cell_of[s] : -1;
        int leader, len;
        wave_runs(c, valid, leader, len);
        const int lane = threadIdx.x & 63;
        int32_t base = 0;
        if (valid && leader == lane) base = atomicSub(&counts[c], len) - len;
        base = __shfl(base, leader, 64);
        if (!valid) return;
        const int32_t dst = offsets[c] + base + (lane - leader);
        slots[dst] = (int32_t)s;
        sid[dst] = __float_as_int(pos[s].w);
    }
};

// Per-cell insertion sort of (id, slot) pairs by id: restores the reference's
// per-cell order (push_back in particle-vector order). Particles that stayed in
// their cell arrive in order already, so most segments need no move.
// One wave owns 64 consecutive cells, whose particles are contiguous: it stages
// their (id, slot) pairs in LDS with coalesced loads, each lane sorts its cell's
// segment there, and the wave writes the pairs back only if a lane moved one.
// A wave whose cells hold more than kCap particles sorts in global memory.
struct KCellsort {
    static constexpr int kThreads = 64;
    static constexpr int kCap = 1024;
    struct Args {
        const int32_t* offsets;
        int64_t nc;
        int32_t* sid;
        int32_t* slots;
    };
    static __device__ void sort_segment(int32_t* __restrict__ id, int32_t* __restrict__ sl, int32_t o0,
                                        int32_t o1, bool& moved) {
        for (int32_t a = o0 + 1; a < o1; ++a) {
            const int32_t key = id[a];
            if (id[a - 1] <= key) continue;
            moved = true;
            const int32_t sv = sl[a];
            int32_t b = a - 1;
            while (b >= o0 && id[b] > key) {
                id[b + 1] = id[b];
                sl[b + 1] = sl[b];
                --b;
            }
            id[b + 1] = key;
            sl[b + 1] = sv;
        }
    }
    static __device__ void run(const Args& a) {
        __shared__ int32_t s_id[kCap], s_sl[kCap];
        const int lane = threadIdx.x;
        const int64_t c_first = (int64_t)blockIdx.x * kThreads;
        if (c_first >= a.nc) return;   // beyond this grid (a launch sized for a larger tank)
        const int64_t c_end = c_first + kThreads < a.nc ? c_first + kThreads : a.nc;
        const int64_t c = c_first + lane;
        const int32_t base = a.offsets[c_first], n = a.offsets[c_end] - base;
        bool moved = false;
        if (n > kCap) {
            if (c < a.nc) sort_segment(a.sid, a.slots, a.offsets[c], a.offsets[c + 1], moved);
            return;
        }
        for (int32_t u = lane; u < n; u += kThreads) {
            s_id[u] = a.sid[base + u];
            s_sl[u] = a.slots[base + u];
        }
        __syncthreads();
        if (c < a.nc) sort_segment(s_id, s_sl, a.offsets[c] - base, a.offsets[c + 1] - base, moved);
        __syncthreads();
        if (!__any(moved)) return;
        for (int32_t u = lane; u < n; u += kThreads) {
            a.sid[base + u] = s_id[u];
            a.slots[base + u] = s_sl[u];
        }
    }
};

struct KGather {
    static constexpr int kThreads = kBlock;
    struct Args {
        const int32_t* slots;
        int64_t np;
        const float4* pin;
        const float4* vin;
        float4* pout;
        float4* vout;
    };
    static __device__ void run(const Args& kargs) {
        const int32_t* __restrict__ slots = kargs.slots;
        int64_t np = kargs.np;
        const float4* __restrict__ pin = kargs.pin;
        const float4* __restrict__ vin = kargs.vin;
        float4* __restrict__ pout = kargs.pout;
        float4* __restrict__ vout = kargs.vout;

        int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (s >= np) return;
        int32_t src = slots[s];
        pout[s] = pin[src];
        vout[s] = vin[src];
    }
};

// ------------------------------------------------ neighbourhood staging
// A block of 256 consecutive sorted particles sits in a few cells of one or two
// grid columns. Its particles are split into groups, the runs of equal query
// columns (lanes whose windows start at the same (x, y) cell); each group's box
// is the union of its lanes' windows. When the groups' columns number at most
// kStageCols and their particles fit kStageCap, each column's z-range of its
// group's box is copied (x, y, z) into LDS with coalesced loads, all in flight
// at once, and the neighbour loops read LDS instead of issuing a dependent
// global load per candidate. A block whose particles end one grid column and
// begin the next thus stages two short boxes instead of one box spanning the
// whole column height. Candidate order and values are unchanged.
constexpr int kStageBlock = 256;   // threads (particles) per staging block
constexpr int kStageCap = 2560;
constexpr int kStageGroups = 4;
constexpr int kStageCols = 36;
constexpr int kStagePer = kStageCap / kStageBlock;   // staged particles per thread

struct StageLds {
    float x[kStageCap], y[kStageCap], z[kStageCap];
    int q0[kStageCols];          // first particle of each column's staged range
    int base[kStageCols + 1];    // its LDS index; base[ncols] = total
    int gbox[kStageGroups][6];   // per group: x0, x1, y0, y1, z0, z1 (cells, inclusive)
    int gcol[kStageGroups];      // index of the group's first column
    int wstart[kStageBlock / 64];
    int ncols, ngroups, ok, total;
};

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Every thread of the block calls this; valid threads pass their query box.
// Returns the thread's group (-1 if not valid) with st.ok set (block-uniform)
// and, if set, the columns staged.
__device__ int stage_neighbourhood(StageLds& st, bool valid, int x0, int x1, int y0, int y1,
                                   int z0, int z1, int N, const int32_t* __restrict__ offsets,
                                   const float4* __restrict__ pos) {
    const int tid = (int)threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid < kStageGroups) {
        st.gbox[tid][0] = st.gbox[tid][2] = st.gbox[tid][4] = 1 << 30;
        st.gbox[tid][1] = st.gbox[tid][3] = st.gbox[tid][5] = -1;
    }
    // group starts: a valid lane whose window start differs from the previous
    // thread's (or which follows an invalid thread or begins the block)
    const int key = valid ? x0 * 65536 + y0 : -1;
    const int kprev = __shfl_up(key, 1, 64);
    __shared__ int s_lastkey[kStageBlock / 64];
    if (lane == 63) s_lastkey[w] = key;
    __syncthreads();
    const int before = lane > 0 ? kprev : (w > 0 ? s_lastkey[w - 1] : -1);
    const bool start = valid && key != before;
    const unsigned long long ball = __ballot(start);
    if (lane == 0) st.wstart[w] = __popcll(ball);
    __syncthreads();
    int g = __popcll(ball & ((2ull << lane) - 1ull)) - 1;   // starts up to this lane
    int ng = 0;
#pragma unroll
    for (int ww = 0; ww < kStageBlock / 64; ++ww) {
        if (ww < w) g += st.wstart[ww];
        ng += st.wstart[ww];
    }
    if (!valid) g = -1;
    if (ng > kStageGroups) {
        if (tid == 0) {
            st.ok = 0;
            st.ngroups = ng;
        }
        __syncthreads();
        return g;
    }
    if (g >= 0) {
        atomicMin(&st.gbox[g][0], x0);
        atomicMax(&st.gbox[g][1], x1);
        atomicMin(&st.gbox[g][2], y0);
        atomicMax(&st.gbox[g][3], y1);
        atomicMin(&st.gbox[g][4], z0);
        atomicMax(&st.gbox[g][5], z1);
    }
    __syncthreads();
    if (tid == 0) {
        int nc = 0;
        for (int gg = 0; gg < ng; ++gg) {
            st.gcol[gg] = nc;
            nc += (st.gbox[gg][1] - st.gbox[gg][0] + 1) * (st.gbox[gg][3] - st.gbox[gg][2] + 1);
        }
        st.ncols = nc;
        st.ngroups = ng;
    }
    __syncthreads();
    const int ncols = st.ncols;
    if (ncols <= 0 || ncols > kStageCols) {
        __syncthreads();
        if (tid == 0) st.ok = 0;
        __syncthreads();
        return g;
    }
    if (tid < ncols) {
        int gg = 0;
        while (gg + 1 < ng && st.gcol[gg + 1] <= tid) ++gg;
        const int* bx = st.gbox[gg];
        const int c = tid - st.gcol[gg], ny = bx[3] - bx[2] + 1;
        const int cx = bx[0] + c / ny, cy = bx[2] + c % ny;
        const int64_t col = cidx(N, cx, cy, 0);
        const int q0 = offsets[col + bx[4]];
        st.q0[tid] = q0;
        st.base[tid] = offsets[col + bx[5] + 1] - q0;   // lengths, turned into bases below
    }
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int c = 0; c < ncols; ++c) {
            const int l = st.base[c];
            st.base[c] = acc;
            acc += l;
        }
        st.base[ncols] = acc;
        st.ok = acc <= kStageCap;
        st.total = acc;
    }
    __syncthreads();
    if (!st.ok) return g;
    const int total = st.total;
    float4 t[kStagePer];
    // a thread's slots u increase with m, so its column search resumes where the
    // previous one stopped
    int c = 0;
#pragma unroll
    for (int m = 0; m < kStagePer; ++m) {
        const int u = tid + m * kStageBlock;
        if (u < total) {
            while (c + 1 < ncols && st.base[c + 1] <= u) ++c;
            t[m] = pos[st.q0[c] + (u - st.base[c])];
        }
    }
#pragma unroll
    for (int m = 0; m < kStagePer; ++m) {
        const int u = tid + m * kStageBlock;
        if (u < total) {
            st.x[u] = t[m].x;
            st.y[u] = t[m].y;
            st.z[u] = t[m].z;
        }
    }
    __syncthreads();
    return g;
}

// Particle index of LDS slot l of a staged block.
__device__ __forceinline__ int staged_particle(const StageLds& st, int l) {
    int c = 0;
    while (c + 1 < st.ncols && st.base[c + 1] <= l) ++c;
    return st.q0[c] + (l - st.base[c]);
}

// Whether column (cx, cy), cells [za, zb], lies in group g's staged box.
__device__ __forceinline__ bool staged_has(const StageLds& st, int g, int cx, int cy, int za, int zb) {
    const int* bx = st.gbox[g];
    return cx >= bx[0] && cx <= bx[1] && cy >= bx[2] && cy <= bx[3] && za >= bx[4] && zb <= bx[5];
}

// LDS index of particle q of column (cx, cy) inside group g's staged box.
__device__ __forceinline__ int staged_index(const StageLds& st, int g, int cx, int cy, int q) {
    const int* bx = st.gbox[g];
    const int c = st.gcol[g] + (cx - bx[0]) * (bx[3] - bx[2] + 1) + (cy - bx[2]);
    return st.base[c] + (q - st.q0[c]);
}

// ---------------------------------------------------------- ComputeDensity
// flip.cpp:697-733 + ApplyExternalForces (flip.cpp:682-695) fused: each particle
// gathers Smooth weights (h = 4*density/maxd) over the 3x3x3 cells around its
// cell; the 3 cells along z of each (sx,sy) column are one contiguous range.
// ComputeDensity's sum for one particle over its staged (or, when the
// neighbourhood did not fit, global) 3x3x3 cells, in the reference's order.
// POW2: the smoothing radius squared is a power of two (smooth_pow2).
template <bool POW2>
__device__ __forceinline__ float density_sum(const StageLds& st, int g, const float4& me, int x0, int x1,
                                             int y0, int y1, int z0, int z1, int N,
                                             const int32_t* __restrict__ offsets,
                                             const float4* __restrict__ pos, float hh, float yhh) {
    float weightsum = 0.0f;
    if (st.ok) {
        // the nine columns' ranges are fetched together before any is used
        int32_t qa[9], qb[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const int sx = x0 + c / 3, sy = y0 + c % 3;
            qa[c] = qb[c] = 0;
            if (sx <= x1 && sy <= y1) {
                const int64_t col = cidx(N, sx, sy, 0);
                qa[c] = offsets[col + z0];
                qb[c] = offsets[col + z1 + 1];
            }
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const int sx = x0 + c / 3, sy = y0 + c % 3;
            if (sx > x1 || sy > y1) continue;
            {
                const int l0 = staged_index(st, g, sx, sy, qa[c]);
                const int l1 = l0 + (qb[c] - qa[c]);
#pragma unroll 4
                for (int l = l0; l < l1; ++l) {
                    float sqd = sqrlength(st.x[l], st.y[l], st.z[l], me.x, me.y, me.z);
                    float weight = 1.0f * (POW2 ? smooth_pow2(sqd, yhh) : smooth_fixed(sqd, hh, yhh));
                    weightsum = weightsum + weight;
                }
            }
        }
    } else {
        for (int sx = x0; sx <= x1; ++sx) {
            for (int sy = y0; sy <= y1; ++sy) {
                const int64_t col = cidx(N, sx, sy, 0);
                const int32_t q0 = offsets[col + z0], q1 = offsets[col + z1 + 1];
                for (int32_t q = q0; q < q1; ++q) {
                    const float4 o = pos[q];
                    float sqd = sqrlength(o.x, o.y, o.z, me.x, me.y, me.z);
                    float weight = 1.0f * (POW2 ? smooth_pow2(sqd, yhh) : smooth_fixed(sqd, hh, yhh));
                    weightsum = weightsum + weight;
                }
            }
        }
    }
    return weightsum;
}

struct KDensity {
    static constexpr int kThreads = kStageBlock;
    struct Args {
        Params p;
        const float4* pos;
        float4* vel;
        int64_t np;
        const int32_t* offsets;
        int apply_forces;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const float4* __restrict__ pos = kargs.pos;
        float4* __restrict__ vel = kargs.vel;
        int64_t np = kargs.np;
        const int32_t* __restrict__ offsets = kargs.offsets;
        int apply_forces = kargs.apply_forces;

        __shared__ StageLds st;
        const int64_t s = xcd_block((np + blockDim.x - 1) / blockDim.x) * blockDim.x + threadIdx.x;
        const bool valid = s < np;
        const int N = p.N;
        const float maxd = p.maxd;
        const float4 me = valid ? pos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
        const int cx = cell_coord(me.x, maxd), cy = cell_coord(me.y, maxd), cz = cell_coord(me.z, maxd);
        const float hsm = 4.0f * p.density / maxd;
        const float hh = hsm * hsm, yhh = 1.0f / hh;
        const int x0 = max(cx - 1, 0), x1 = min(cx + 1, N - 1);
        const int y0 = max(cy - 1, 0), y1 = min(cy + 1, N - 1);
        const int z0 = max(cz - 1, 0), z1 = min(cz + 1, N - 1);
        const int g = stage_neighbourhood(st, valid, x0, x1, y0, y1, z0, z1, N, offsets, pos);
        if (!valid) return;
        // the radius is the same for every particle: a uniform branch
        const float weightsum =
            is_pow2_float(hh) ? density_sum<true>(st, g, me, x0, x1, y0, y1, z0, z1, N, offsets, pos, hh, yhh)
                              : density_sum<false>(st, g, me, x0, x1, y0, y1, z0, z1, N, offsets, pos, hh, yhh);
        float4 v = vel[s];
        v.w = weightsum / p.max_density;
        if (apply_forces) {
            for (int f = 0; f < p.n_forces; ++f) {
                v.x = v.x + p.forces[f][0] * p.dt;
                v.y = v.y + p.forces[f][1] * p.dt;
                v.z = v.z + p.forces[f][2] * p.dt;
            }
        }
        vel[s] = v;
    }
};

// ------------------------------------------------- MarkCellTypes + BuildSDF
// CellSDF (particlegrid.cpp:75-90): L = 0.2/density^3 - sum of the cell's
// densities in list order; A = SOLID on the static walls, else FLUID iff L < 0
// (MarkCellTypes, particlegrid.cpp:108-138). BuildSDF (92-106) stores L.
struct KMark {
    static constexpr int kThreads = kBlock;
    struct Args {
        Params p;
        const float4* vel;
        const int32_t* offsets;
        uint8_t* A;
        float* L;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const float4* __restrict__ vel = kargs.vel;
        const int32_t* __restrict__ offsets = kargs.offsets;
        uint8_t* __restrict__ A = kargs.A;
        float* __restrict__ L = kargs.L;

        const int N = p.N;
        const int64_t nc = (int64_t)N * N * N;
        int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= nc) return;
        const uint32_t cu = (uint32_t)c, nu = (uint32_t)N;
        const int k = (int)(cu % nu), j = (int)((cu / nu) % nu), i = (int)(cu / nu / nu);
        float accm = 0.0f;
        const int32_t o0 = offsets[c], o1 = offsets[c + 1];
        for (int32_t a = o0; a < o1; ++a) accm += vel[a].w;
        const float dd = p.density;
        const float n0 = 1.0f / (dd * dd * dd);
        const float lc = 0.2f * n0 - accm;
        const int w = p.wall;
        const bool wall = w > 0 && (i < w || j < w || k < w || i >= N - w || j >= N - w || k >= N - w);
        uint8_t a = kAir;
        if (wall) a = kSolid;
        else a = (lc < 0.0f) ? kFluid : kAir;
        A[c] = a;
        L[c] = lc;
    }
};

// -------------------------------------------------- SplatParticlesToMACGrid
// particlegridoperations.inl:146-249: each face gathers the Sharpen-weighted
// (RE = 1.4 cells), mass-weighted mean of particle velocities over the cells of
// GetWallNeighbors' asymmetric window (particlegrid.cpp:52-73): 2 cells along the
// face normal, 4 along the others. Fused: StorePreviousGrid (flip.cpp:324-363,
// copy before the boundary pass) and EnforceBoundaryVelocity
// (particlegridoperations.inl:45-104).
//
// P2G through LDS. A 64-lane block owns the cells (i, j, k0..k0+63) and computes
// the x-, y- and z-faces of each (faces at i = N, j = N, k = N included). The
// union of the three reference windows is the 4x4 block of particle columns
// (sx, sy) in [i-2, i+1] x [j-2, j+1], cells k-2..k+1 along z. Columns are visited
// in (sx, sy) order; each column's particles (contiguous in sorted order) are
// staged in LDS with their clamped grid positions, and every face adds the
// particles of its own window in ascending order: per face the additions happen
// in exactly the reference's order (sx -> sy -> sz -> list), so the sums are the
// same as a per-face gather's. Weights beyond the kernel radius are exactly
// zero and skipped (adding +0 to a sum that starts at +0 never changes it).
constexpr int kP2GCap = 256;   // particles staged per pass (small: occupancy hides the staging)

// Weight of one particle at offset (dx, dy, dz) from a face, branch-free: the
// Sharpen or hat weight, 0 outside its support. Terms of weight 0 are added to
// the sums like the others (as the reference does): they add +-0 to sums that
// start at +0 and so never change a bit.
template <bool TRI>
__device__ __forceinline__ float p2g_weight(float dx, float dy, float dz, float hh) {
    if constexpr (TRI) {
        // a factor is +0 once its |d| >= 1, so the product is too
        const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
        return (gmax_num(1.0f - ax, 0.0f) * gmax_num(1.0f - ay, 0.0f)) * gmax_num(1.0f - az, 0.0f);
    } else {
        // r2 >= hh gives hh / max(r2, 1e-5) <= 1 (a correctly rounded quotient
        // of a smaller by a larger number), so the weight is +0 there already
        const float r2 = dx * dx + dy * dy + dz * dz;
        return 1.0f * sharpen_nr(r2, hh);
    }
}

// One staged span [q0, q1) of a particle column added to the block's faces, in
// list order. DX / DY: the column lies in the x / y faces' window (uniform over
// the block); the z faces take every column, the cells of their window being
// [rb0, rb1). Every operand of an iteration is read from LDS before any weight
// is computed and no branch separates the components, so the LDS latency of one
// iteration is paid once. Lanes whose face does not exist accumulate too and
// never store. RE14: the Sharpen radius is the reference's 1.4 (sharpen_re14).
template <bool TRI, bool RE14, bool DX, bool DY>
__device__ __forceinline__ void p2g_span(const float (*__restrict__ sp)[kP2GCap], int32_t q0, int32_t q1,
                                         int32_t c0, int32_t rb0, int32_t rb1, float fi, float fih,
                                         float fj, float fjh, float fk, float fkh, float hh, float& sux,
                                         float& swx, float& suy, float& swy, float& suz, float& swz) {
    auto sharpen_w = [&](float r2) { return RE14 ? sharpen_re14(r2) : sharpen_nr(r2, hh); };
    for (int32_t q = q0; q < q1; ++q) {
        const int t = q - c0;
        const float pz = sp[2][t];
        const float s0 = sp[0][t], s1 = sp[1][t], s6 = TRI ? 0.0f : sp[6][t];
        const float ux = DX ? sp[3][t] : 0.0f, uy = DY ? sp[4][t] : 0.0f, uz = sp[5][t];
        // offsets to the faces (x face at (i, j+.5, k+.5), ...); for Sharpen r^2
        // in the reference's order (dx^2 + dy^2) + dz^2, the first two terms staged
        const float az = pz - fk, bz = pz - fkh;
        const float bz2 = bz * bz;
        if constexpr (DX) {
            const float w = TRI ? p2g_weight<true>(s0 - fi, s1 - fjh, bz, hh) : 1.0f * sharpen_w(s0 + bz2);
            sux += w * ux;
            swx += w;
        }
        if constexpr (DY) {
            const float w = TRI ? p2g_weight<true>(s0 - fih, s1 - fj, bz, hh) : 1.0f * sharpen_w(s1 + bz2);
            suy += w * uy;
            swy += w;
        }
        const float wz0 = TRI ? p2g_weight<true>(s0 - fih, s1 - fjh, az, hh) : 1.0f * sharpen_w(s6 + az * az);
        const float wz = (q >= rb0 && q < rb1) ? wz0 : 0.0f;
        suz += wz * uz;
        swz += wz;
    }
}

// The weighted means of the three faces at (i, j, k) (0 where no weight reached a
// face), stored after StorePreviousGrid's copy (flip.cpp:324-363) into du and,
// with EnforceBoundaryVelocity (particlegridoperations.inl:45-104) applied, into u.
__device__ __forceinline__ void p2g_store(int N, const uint8_t* __restrict__ A, int i, int j, int k,
                                          bool vx, bool vy, bool vz, float sux, float swx, float suy,
                                          float swy, float suz, float swz, float* __restrict__ ux,
                                          float* __restrict__ uy, float* __restrict__ uz,
                                          float* __restrict__ dux, float* __restrict__ duy,
                                          float* __restrict__ duz) {
    if (vx) {
        const int64_t f = fidx(N, 0, i, j, k);
        const float val = swx > 0 ? sux / swx : 0.0f;
        dux[f] = val;
        bool zero = (i == 0 || i == N);
        if (!zero) zero = (A[cidx(N, i, j, k)] == kSolid) != (A[cidx(N, i - 1, j, k)] == kSolid);
        ux[f] = zero ? 0.0f : val;
    }
    if (vy) {
        const int64_t f = fidx(N, 1, i, j, k);
        const float val = swy > 0 ? suy / swy : 0.0f;
        duy[f] = val;
        bool zero = (j == 0 || j == N);
        if (!zero) zero = (A[cidx(N, i, j, k)] == kSolid) != (A[cidx(N, i, j - 1, k)] == kSolid);
        uy[f] = zero ? 0.0f : val;
    }
    if (vz) {
        const int64_t f = fidx(N, 2, i, j, k);
        const float val = swz > 0 ? suz / swz : 0.0f;
        duz[f] = val;
        bool zero = (k == 0 || k == N);
        if (!zero) zero = (A[cidx(N, i, j, k)] == kSolid) != (A[cidx(N, i, j, k - 1)] == kSolid);
        uz[f] = zero ? 0.0f : val;
    }
}

template <bool TRI>
struct KP2gCols {
    static constexpr int kThreads = 64;
    struct Args {
        Params p;
        const float4* pos;
        const float4* vel;
        const int32_t* offsets;
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
        const float4* __restrict__ pos = kargs.pos;
        const float4* __restrict__ vel = kargs.vel;
        const int32_t* __restrict__ offsets = kargs.offsets;
        const uint8_t* __restrict__ A = kargs.A;
        float* __restrict__ ux = kargs.ux;
        float* __restrict__ uy = kargs.uy;
        float* __restrict__ uz = kargs.uz;
        float* __restrict__ dux = kargs.dux;
        float* __restrict__ duy = kargs.duy;
        float* __restrict__ duz = kargs.duz;

        // staged particles: the Sharpen kernel keeps, per particle, the first two
        // terms of r^2 = (dx^2 + dy^2) + dz^2 for each face component (dx and dy
        // are the same for all 64 faces of the block's column) and pz; the
        // trilinear kernel keeps px, py, pz; then the velocity
        __shared__ float sp[7][kP2GCap];
        const int N = p.N;
        const int nkc = (N + 1 + 63) / 64;
        // Blocks run over 8x8 tiles of face columns (i, j), k-chunks innermost, so
        // the particle columns that neighbouring face columns share are staged while
        // still in L2; each XCD takes a contiguous run of tiles. The tile edge is
        // a launch parameter (autotuned per grid size for batched tanks).
        const int kT = p.p2g_tile;
        const int nti = (N + 1 + kT - 1) / kT;
        const int64_t nbl = (int64_t)nti * nti * kT * kT * nkc;
        int64_t b = xcd_block(nbl);
        if (b >= nbl) return;
        const int kc = (int)(b % nkc);
        b /= nkc;
        const int jj = (int)(b % kT);
        b /= kT;
        const int ii = (int)(b % kT);
        b /= kT;
        const int j = (int)(b % nti) * kT + jj;
        const int i = (int)(b / nti) * kT + ii;
        if (i > N || j > N) return;
        const int lane = threadIdx.x;
        const int k0 = kc * 64;
        const int k = k0 + lane;
        const bool kval = k <= N;
        const bool vx = kval && j < N && k < N;
        const bool vy = kval && i < N && k < N;
        const bool vz = kval && i < N && j < N;
        const float maxd = (float)N;
        const float RE = p.sharpen_re;
        const float hh = RE * RE;
        const bool tri = TRI;
        const bool re14 = !TRI && hh == kSharpenHH;
        const float fi = (float)i, fj = (float)j, fk = (float)k;
        const float fih = fi + 0.5f, fjh = fj + 0.5f, fkh = fk + 0.5f;
        float swx = 0.0f, sux = 0.0f, swy = 0.0f, suy = 0.0f, swz = 0.0f, suz = 0.0f;
        // Along a face's tangential axis the window's first cell ([f-2.5, f-1.5]
        // from the face coordinate f) is at least 1.5 cells away: every weight
        // there is exactly 0 for the trilinear kernel and for Sharpen with
        // RE^2 <= 2.25 (the weight is 0 once r2 >= RE^2), so those cells are not
        // visited. The window then starts one cell later on every axis.
        const int lo = (tri || hh <= 2.25f) ? 1 : 2;
        const int zlo_blk = max(0, k0 - lo), zhi_blk = min(N - 1, k0 + 64);
        const int a0 = max(k - lo, 0), a1 = min(k + 1, N - 1);
        const int b0 = max(k - 1, 0), b1 = min(k, N - 1);
        for (int sx = max(i - lo, 0); sx <= min(i + 1, N - 1); ++sx) {
            for (int sy = max(j - lo, 0); sy <= min(j + 1, N - 1); ++sy) {
                const bool colx = (sx >= i - 1 && sx <= i) && j < N;
                const bool coly = (sy >= j - 1 && sy <= j) && i < N;
                const bool colz = i < N && j < N;
                if (!(colx || coly || colz)) continue;
                const int64_t col = cidx(N, sx, sy, 0);
                const int32_t c_begin = offsets[col + zlo_blk], c_end = offsets[col + zhi_blk + 1];
                int32_t ra0 = 0, ra1 = 0, rb0 = 0, rb1 = 0;
                if (kval && a0 <= a1) { ra0 = offsets[col + a0]; ra1 = offsets[col + a1 + 1]; }
                if (kval && b0 <= b1) { rb0 = offsets[col + b0]; rb1 = offsets[col + b1 + 1]; }
                const bool inx = sx >= i - 1 && sx <= i, iny = sy >= j - 1 && sy <= j;
                for (int32_t c0 = c_begin; c0 < c_end; c0 += kP2GCap) {
                    const int32_t c1 = min(c_end, c0 + kP2GCap);
                    __syncthreads();
                    for (int32_t q = c0 + lane; q < c1; q += 64) {
                        const float4 o = pos[q];
                        const float4 v = vel[q];
                        const int t = q - c0;
                        const float px = gmax(0.0f, gmin(maxd, maxd * o.x));
                        const float py = gmax(0.0f, gmin(maxd, maxd * o.y));
                        if constexpr (TRI) {
                            sp[0][t] = px;
                            sp[1][t] = py;
                        } else {
                            // x face (i, j+.5, k+.5), y face (i+.5, j, k+.5), z face (i+.5, j+.5, k)
                            const float ax = px - fi, bx = px - fih, ay = py - fj, by = py - fjh;
                            sp[0][t] = ax * ax + by * by;
                            sp[1][t] = bx * bx + ay * ay;
                            sp[6][t] = bx * bx + by * by;
                        }
                        sp[2][t] = gmax(0.0f, gmin(maxd, maxd * o.z));
                        sp[3][t] = v.x;
                        sp[4][t] = v.y;
                        sp[5][t] = v.z;
                    }
                    __syncthreads();
                    const int32_t q0 = max(ra0, c0), q1 = min(ra1, c1);
                    const int32_t b0c = rb0, b1c = rb1;
                    if (re14) {
                        if (inx && iny) p2g_span<TRI, true, true, true>(sp, q0, q1, c0, b0c, b1c, fi, fih, fj, fjh, fk, fkh, hh, sux, swx, suy, swy, suz, swz);
                        else if (inx) p2g_span<TRI, true, true, false>(sp, q0, q1, c0, b0c, b1c, fi, fih, fj, fjh, fk, fkh, hh, sux, swx, suy, swy, suz, swz);
                        else if (iny) p2g_span<TRI, true, false, true>(sp, q0, q1, c0, b0c, b1c, fi, fih, fj, fjh, fk, fkh, hh, sux, swx, suy, swy, suz, swz);
                        else p2g_span<TRI, true, false, false>(sp, q0, q1, c0, b0c, b1c, fi, fih, fj, fjh, fk, fkh, hh, sux, swx, suy, swy, suz, swz);
                    } else {
                        if (inx && iny) p2g_span<TRI, false, true, true>(sp, q0, q1, c0, b0c, b1c, fi, fih, fj, fjh, fk, fkh, hh, sux, swx, suy, swy, suz, swz);
                        else if (inx) p2g_span<TRI, false, true, false>(sp, q0, q1, c0, b0c, b1c, fi, fih, fj, fjh, fk, fkh, hh, sux, swx, suy, swy, suz, swz);
                        else if (iny) p2g_span<TRI, false, false, true>(sp, q0, q1, c0, b0c, b1c, fi, fih, fj, fjh, fk, fkh, hh, sux, swx, suy, swy, suz, swz);
                        else p2g_span<TRI, false, false, false>(sp, q0, q1, c0, b0c, b1c, fi, fih, fj, fjh, fk, fkh, hh, sux, swx, suy, swy, suz, swz);
                    }
                }
            }
        }
        p2g_store(N, A, i, j, k, vx, vy, vz, sux, swx, suy, swy, suz, swz, ux, uy, uz, dux, duy, duz);
    }
};

// ------------------------------------------------ P2G, tiled scatter (ARIEL_P2G_TILED)
// The same sums as KP2gCols (particlegridoperations.inl:146-249 over the window of
// GetWallNeighbors, particlegrid.cpp:52-73), gathered the other way round. A
// particle in cell (cx, cy, cz) lies in the window of the x faces i in {cx, cx+1},
// j in {cy-1, cy, cy+1}, k in {cz-1, cz, cz+1} (the window's first cell along a
// tangential axis, 1.5 cells or more from the face, carries only zero weights
// while RE <= 1.5), and likewise for y and z: 54 candidate faces.
//
// Pass 1 (KP2gTiles): a 512-lane block owns an 8^3 tile of cells and takes the
// three components in turn. Four lanes share a cell: they walk its particles
// (contiguous in sorted order) four apart, so a wave's loads cover a few
// contiguous runs of particles, and each lane keeps the (weight, weighted
// velocity) sums of the cell's 18 candidate faces of the component in
// registers. The four lanes' sums are added within the quad (DPP), stored in
// LDS per cell, and each face the tile's cells reach (900 per component) adds
// the sums of the tile's cells in its window in a fixed order and goes to HBM as
// the tile's partial. No atomics: the sums are the same on every run. Pass 2
// (KP2gCombine): each face adds the partials of the <= 8 tiles that hold cells
// of its window, in tile order, and stores the mean (p2g_store).
//
// Sharpen's quotient uses the hardware reciprocal: the sums differ from the
// ordered gather's by rounding only (per-cell partial sums, a reciprocal within
// 1 ulp), far inside the 1e-5 relative bound the face velocities are held to.
struct KP2gTilesBase {
    // local index of a face in its component's region of the tile: x faces
    // (i', j', k') in 9 x 10 x 10 (i' = i - 8 tx, j' = j - 8 ty + 1, ...), y in
    // 10 x 9 x 10, z in 10 x 10 x 9
    static __device__ __forceinline__ int xi(int i, int j, int k) { return (i * 10 + j) * 10 + k; }
    static __device__ __forceinline__ int yi(int i, int j, int k) { return (i * 9 + j) * 10 + k; }
    static __device__ __forceinline__ int zi(int i, int j, int k) { return (i * 10 + j) * 9 + k; }
};

template <bool TRI>
struct KP2gTiles : KP2gTilesBase {
    static constexpr int kThreads = 512;
    static constexpr int kMinWavesPerEU = 4;   // two blocks per CU
    struct Args {
        Params p;
        const float4* pos;
        const float4* vel;
        const int32_t* offsets;
        float* part;
        uint8_t* tflag;
    };
    // Factors of one axis for the faces normal to it (E: distances to the integer
    // faces c, c+1) and along it (H: to the half-integer centres c-0.5, c+0.5,
    // c+1.5), from f = p - c; squares for Sharpen, hat values for trilinear.
    static __device__ __forceinline__ float fac(float d) {
        return TRI ? gmax_num(1.0f - fabsf(d), 0.0f) : d * d;
    }
    // Sum of a value over the four lanes of a quad (DPP quad permutations).
    static __device__ __forceinline__ float quad_sum(float v) {
        v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));
        v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, true));
        return v;
    }
    // One component's 18 candidate sums over the particles q0 + slot, q0 + slot +
    // 4, ... < q1 of the cell (cx, cy, cz). NA, NB, NC: extents of the candidate
    // box (2 along the component's own axis, else 3).
    template <int C>
    // (po, pv): the particle q0, loaded by the caller ahead of time.
    static __device__ __forceinline__ void sums(const Args& a, int32_t q0, int32_t q1, float maxd,
                                                float fcx, float fcy, float fcz, float hh,
                                                const float4& po, const float4& pv,
                                                float (&w)[18], float (&u)[18]) {
#pragma unroll
        for (int n = 0; n < 18; ++n) w[n] = u[n] = 0.0f;
        if (q0 >= q1) return;
        const float wcap = fmaf(hh, 1.0e5f, -1.0f);
        float4 on = po;
        float vn = (&pv.x)[C];
        for (int32_t q = q0; q < q1; q += 4) {
            const float4 o = on;
            const float v = vn;
            // the next particle's loads in flight during this one's sums (the last
            // iteration reloads its own particle: no branch in the loop)
            const int32_t qn = min(q + 4, q1 - 1);
            on = a.pos[qn];
            vn = (&a.vel[qn].x)[C];
            const float fx = gmax(0.0f, gmin(maxd, maxd * o.x)) - fcx;
            const float fy = gmax(0.0f, gmin(maxd, maxd * o.y)) - fcy;
            const float fz = gmax(0.0f, gmin(maxd, maxd * o.z)) - fcz;
            float X[3], Y[3], Z[3];
            X[0] = fac(C == 0 ? fx : fx + 0.5f);
            X[1] = fac(C == 0 ? fx - 1.0f : fx - 0.5f);
            X[2] = fac(fx - 1.5f);
            Y[0] = fac(C == 1 ? fy : fy + 0.5f);
            Y[1] = fac(C == 1 ? fy - 1.0f : fy - 0.5f);
            Y[2] = fac(fy - 1.5f);
            Z[0] = fac(C == 2 ? fz : fz + 0.5f);
            Z[1] = fac(C == 2 ? fz - 1.0f : fz - 0.5f);
            Z[2] = fac(fz - 1.5f);
            constexpr int NA = C == 0 ? 2 : 3, NB = C == 1 ? 2 : 3, NC = C == 2 ? 2 : 3;
#pragma unroll
            for (int aa = 0; aa < NA; ++aa)
#pragma unroll
                for (int bb = 0; bb < NB; ++bb) {
                    const float xy = TRI ? X[aa] * Y[bb] : X[aa] + Y[bb];
#pragma unroll
                    for (int cc = 0; cc < NC; ++cc) {
                        float wt;
                        if constexpr (TRI) {
                            wt = xy * Z[cc];
                        } else {
                            // max(hh / max(r2, 1e-5) - 1, 0) as one clamp to
                            // [0, hh / 1e-5 - 1] (r2 = 0 gives +inf, clamped)
                            wt = __builtin_amdgcn_fmed3f(fmaf(hh, __builtin_amdgcn_rcpf(xy + Z[cc]), -1.0f),
                                                         0.0f, wcap);
                        }
                        const int n = (aa * NB + bb) * NC + cc;
                        w[n] += wt;
                        u[n] = fmaf(wt, v, u[n]);
                    }
                }
        }
    }
    static constexpr int kRounds = 4;   // 512 cells, 128 per round (16 per wave)
    static __device__ void run(const Args& a) {
        const Params& p = a.p;
        const int N = p.N;
        const int nt = (N + kP2gTileEdge - 1) / kP2gTileEdge;
        const int64_t nbl = (int64_t)nt * nt * nt;
        const int64_t b = xcd_block(nbl);
        if (b >= nbl) return;
        const int tz = (int)(b % nt), ty = (int)((b / nt) % nt), tx = (int)(b / ((int64_t)nt * nt));
        const int t = threadIdx.x;
        const int wave = t >> 6, lane = t & 63, c16 = lane >> 2, slot = lane & 3;
        // round r: the wave's cells are the 8 cells of columns 2 (wave + 8 r) and
        // 2 (wave + 8 r) + 1 of the tile (column = 8 lx + ly)
        int32_t qa[kRounds], qb[kRounds];
        int cell[kRounds];
        bool any = false;
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const int col = 2 * (wave + 8 * r) + (c16 >> 3);
            const int lx = col >> 3, ly = col & 7, lz = c16 & 7;
            cell[r] = (lx * 8 + ly) * 8 + lz;
            const int cx = tx * 8 + lx, cy = ty * 8 + ly, cz = tz * 8 + lz;
            qa[r] = qb[r] = 0;
            if (cx < N && cy < N && cz < N) {
                const int64_t c = cidx(N, cx, cy, cz);
                qa[r] = a.offsets[c];
                qb[r] = a.offsets[c + 1];
            }
            any = any || qb[r] > qa[r];
        }
        if (!__syncthreads_or(any)) {
            if (t == 0) a.tflag[b] = 0;
            return;
        }
        // the cells' candidate sums of one component: [candidate][cell], weights
        // then weighted velocities
        __shared__ float cs[2][18][512];
        const float maxd = (float)N;
        const float hh = p.sharpen_re * p.sharpen_re;
        float* __restrict__ out = a.part + b * kP2gTilePart;
        // the first particle of the next round, loaded while the current round's
        // sums are reduced and stored
        float4 po = make_float4(0, 0, 0, 0), pv = make_float4(0, 0, 0, 0);
        auto first = [&](int r) {
            const int32_t q = qa[r] + slot;
            if (q < qb[r]) {
                po = a.pos[q];
                pv = a.vel[q];
            }
        };
        first(0);
#pragma unroll 1
        for (int comp = 0; comp < 3; ++comp) {
            if (comp > 0) __syncthreads();   // the previous component's reads are done
#pragma unroll 1
            for (int r = 0; r < kRounds; ++r) {
                const int lc = cell[r];
                const float fcx = (float)(tx * 8 + (lc >> 6)), fcy = (float)(ty * 8 + ((lc >> 3) & 7)),
                            fcz = (float)(tz * 8 + (lc & 7));
                float w[18], u[18];
                const int32_t q0 = qa[r] + slot, q1 = qb[r];
                if (comp == 0) sums<0>(a, q0, q1, maxd, fcx, fcy, fcz, hh, po, pv, w, u);
                else if (comp == 1) sums<1>(a, q0, q1, maxd, fcx, fcy, fcz, hh, po, pv, w, u);
                else sums<2>(a, q0, q1, maxd, fcx, fcy, fcz, hh, po, pv, w, u);
                first(r + 1 < kRounds ? r + 1 : 0);   // next round (the next component's first)
#pragma unroll
                for (int n = 0; n < 18; ++n) {
                    w[n] = quad_sum(w[n]);
                    u[n] = quad_sum(u[n]);
                }
                if (slot == 0) {
#pragma unroll
                    for (int n = 0; n < 18; ++n) {
                        cs[0][n][lc] = w[n];
                        cs[1][n][lc] = u[n];
                    }
                }
            }
            __syncthreads();
            // the candidate box's extents along x, y, z and the region's
            const int NA = comp == 0 ? 2 : 3, NB = comp == 1 ? 2 : 3, NC = comp == 2 ? 2 : 3;
            const int RB = 7 + NB, RC = 7 + NC;
            for (int f = t; f < kP2gTileFaces; f += kThreads) {
                const int fk = f % RC, fj = (f / RC) % RB, fi = f / (RC * RB);
                float sw = 0.0f, su = 0.0f;
                for (int aa = 0; aa < NA; ++aa) {
                    const int x = fi - aa;
                    if (x < 0 || x > 7) continue;
                    for (int bb = 0; bb < NB; ++bb) {
                        const int y = fj - bb;
                        if (y < 0 || y > 7) continue;
                        for (int cc = 0; cc < NC; ++cc) {
                            const int z = fk - cc;
                            if (z < 0 || z > 7) continue;
                            const int n = (aa * NB + bb) * NC + cc, c = (x * 8 + y) * 8 + z;
                            sw += cs[0][n][c];
                            su += cs[1][n][c];
                        }
                    }
                }
                out[comp * 2 * kP2gTileFaces + f] = sw;
                out[comp * 2 * kP2gTileFaces + kP2gTileFaces + f] = su;
            }
        }
        if (t == 0) a.tflag[b] = 1;
    }
};

// Pass 2: the faces at (i, j, k) for 0 <= i, j, k <= N (as KP2gCols), each the sum
// of its window's tiles' partials taken in tile order.
struct KP2gCombine : KP2gTilesBase {
    static constexpr int kThreads = 256;
    struct Args {
        Params p;
        const float* part;
        const uint8_t* tflag;
        const uint8_t* A;
        float* ux;
        float* uy;
        float* uz;
        float* dux;
        float* duy;
        float* duz;
    };
    // sum over the tiles holding window cells [lo, hi] per axis (cells clamped to
    // the grid); (oi, oj, ok) = 1 on tangential axes (the region starts one cell
    // before the tile there)
    static __device__ __forceinline__ void gather(const Args& a, int nt, int comp, int i, int j, int k,
                                                  int xlo, int xhi, int ylo, int yhi, int zlo, int zhi,
                                                  int oi, int oj, int ok, float& sw, float& su) {
        sw = 0.0f;
        su = 0.0f;
        for (int tx = xlo >> 3; tx <= xhi >> 3; ++tx)
            for (int ty = ylo >> 3; ty <= yhi >> 3; ++ty)
                for (int tz = zlo >> 3; tz <= zhi >> 3; ++tz) {
                    const int64_t tb = ((int64_t)tx * nt + ty) * nt + tz;
                    if (!a.tflag[tb]) continue;
                    const int li = i - 8 * tx + oi, lj = j - 8 * ty + oj, lk = k - 8 * tz + ok;
                    const int f = comp == 0 ? xi(li, lj, lk) : comp == 1 ? yi(li, lj, lk) : zi(li, lj, lk);
                    const float* t = a.part + tb * kP2gTilePart + comp * 2 * kP2gTileFaces;
                    sw += t[f];
                    su += t[kP2gTileFaces + f];
                }
    }
    static __device__ void run(const Args& a) {
        const int N = a.p.N;
        const int nt = (N + kP2gTileEdge - 1) / kP2gTileEdge;
        const int64_t n1 = N + 1;
        const int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x;
        if (idx >= n1 * n1 * n1) return;
        const int k = (int)(idx % n1), j = (int)((idx / n1) % n1), i = (int)(idx / (n1 * n1));
        const bool vx = j < N && k < N, vy = i < N && k < N, vz = i < N && j < N;
        float swx = 0, sux = 0, swy = 0, suy = 0, swz = 0, suz = 0;
        const int m = N - 1;
        if (vx)
            gather(a, nt, 0, i, j, k, max(i - 1, 0), min(i, m), max(j - 1, 0), min(j + 1, m),
                   max(k - 1, 0), min(k + 1, m), 0, 1, 1, swx, sux);
        if (vy)
            gather(a, nt, 1, i, j, k, max(i - 1, 0), min(i + 1, m), max(j - 1, 0), min(j, m),
                   max(k - 1, 0), min(k + 1, m), 1, 0, 1, swy, suy);
        if (vz)
            gather(a, nt, 2, i, j, k, max(i - 1, 0), min(i + 1, m), max(j - 1, 0), min(j + 1, m),
                   max(k - 1, 0), min(k, m), 1, 1, 0, swz, suz);
        p2g_store(N, a.A, i, j, k, vx, vy, vz, sux, swx, suy, swy, suz, swz, a.ux, a.uy, a.uz,
                  a.dux, a.duy, a.duz);
    }
};

// ------------------------------------------- SolvePicFlip + AdvectParticles
// flip.cpp:287-322 and 229-285: FLIP = u + I(du), PIC = I(u), blend
// (1-r)*PIC + r*FLIP; p += dt * I(u)(p) (the PIC velocity at the same position);
// wall clamp to [1/maxd, 1-1/maxd]. The re-sort inside AdvectParticles only
// serves the solid-particle push, which is inert without SOLID particles.
struct KG2pAdvect {
    static constexpr int kThreads = kBlock;
    struct Args {
        Params p;
        float4* pos;
        float4* vel;
        int64_t np;
        float* ux;
        float* uy;
        float* uz;
        float* dux;
        float* duy;
        float* duz;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        float4* __restrict__ pos = kargs.pos;
        float4* __restrict__ vel = kargs.vel;
        int64_t np = kargs.np;
        float* ux = kargs.ux;
        float* uy = kargs.uy;
        float* uz = kargs.uz;
        float* dux = kargs.dux;
        float* duy = kargs.duy;
        float* duz = kargs.duz;

        const int64_t s = xcd_block((np + blockDim.x - 1) / blockDim.x) * blockDim.x + threadIdx.x;
        if (s >= np) return;
        const int N = p.N;
        float* g[3] = {ux, uy, uz};
        float* d[3] = {dux, duy, duz};
        float4 me = pos[s];
        float4 v = vel[s];
        float dx, dy, dz, px, py, pz;
        interpolate_velocity(d, N, me.x, me.y, me.z, dx, dy, dz);
        float tx = dx + v.x, ty = dy + v.y, tz = dz + v.z;
        interpolate_velocity(g, N, me.x, me.y, me.z, px, py, pz);
        const float r = p.picflip;
        v.x = (1.0f - r) * px + r * tx;
        v.y = (1.0f - r) * py + r * ty;
        v.z = (1.0f - r) * pz + r * tz;
        me.x = me.x + p.dt * px;
        me.y = me.y + p.dt * py;
        me.z = me.z + p.dt * pz;
        const float rr = 1.0f / p.maxd;
        me.x = gmax(rr, gmin(1.0f - rr, me.x));
        me.y = gmax(rr, gmin(1.0f - rr, me.y));
        me.z = gmax(rr, gmin(1.0f - rr, me.z));
        pos[s] = me;
        vel[s] = v;
    }
};

// ------------------------------------------------------- ResampleParticles
// particleresampler.inl:35-148. Neighbour cells come from GetCellNeighbors with
// a float index: loop variables start at (int)(x - 1) and run while
// (float)s <= x + 1, skipping cells outside [0, N-1].
__device__ __forceinline__ void cell_range(float x, int N, int& lo, int& hi) {
    lo = (int)(x - 1.0f);
    hi = lo;
    while ((float)(hi + 1) <= x + 1.0f) ++hi;
    if (lo < 0) lo = 0;
    if (hi > N - 1) hi = N - 1;
}

// Distance from the grid coordinate x to the cell interval [c, c+1], less a
// margin of 1e-3 cells plus a few ulps of the coordinate that covers the
// rounding of x and of the particles' own grid coordinates.
__device__ __forceinline__ float cell_gap(float x, int c, float margin) {
    const float g = fmaxf(fmaxf((float)c - x, x - (float)(c + 1)), 0.0f);
    return fmaxf(g - margin, 0.0f);
}

// Narrows the z cells [zl, zh] of column (gx2 + gy2) to those whose squared
// distance bound does not exceed gbound. The bound grows away from the point,
// so the cells dropped are at the ends of the contiguous range.
__device__ __forceinline__ void narrow_z(float z, float gxy2, float gbound, float margin, int& zl,
                                        int& zh) {
    while (zl <= zh) {
        const float g = cell_gap(z, zl, margin);
        if (gxy2 + g * g <= gbound) break;
        ++zl;
    }
    while (zh >= zl) {
        const float g = cell_gap(z, zh, margin);
        if (gxy2 + g * g <= gbound) break;
        --zh;
    }
}

constexpr int kSpringQueue = 8;   // queued in-radius neighbours per thread

struct KResample {
    static constexpr int kThreads = kStageBlock;
    struct Args {
        Params p;
        const float4* pos;
        const float4* vel;
        int64_t np;
        const int32_t* offsets;
        float4* pout;
        float4* vout;
        uint8_t* accept;
    };
    static __device__ void run(const Args& kargs) {
        const Params& p = kargs.p;
        const float4* __restrict__ pos = kargs.pos;
        const float4* __restrict__ vel = kargs.vel;
        int64_t np = kargs.np;
        const int32_t* __restrict__ offsets = kargs.offsets;
        float4* __restrict__ pout = kargs.pout;
        float4* __restrict__ vout = kargs.vout;
        uint8_t* __restrict__ accept = kargs.accept;

        __shared__ StageLds st;
        __shared__ int s_queue[kSpringQueue][kStageBlock];
        const int64_t s = xcd_block((np + blockDim.x - 1) / blockDim.x) * blockDim.x + threadIdx.x;
        const bool valid = s < np;
        const int N = p.N;
        const float maxd = p.maxd;
        const float re = p.density / maxd;
        const float dt = p.dt;
        const float springforce = p.spring;
        const float4 me = valid ? pos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
        const uint32_t my_id = (uint32_t)__float_as_int(me.w);
        const float re2 = re * re;
        const float yre2 = 1.0f / re2;
        const float re2_skip = re2 * 1.0001f;
        // Cells are visited only when the distance bound to them (grid units, see
        // cell_gap) is within re*1.0005: every pair in a skipped cell has
        // dist^2 > re^2*1.0001 and would be skipped by the pair tests below.
        const float gbound = re2 * (maxd * maxd) * 1.001f;
        const float margin = 1.0e-3f + 4.0e-7f * maxd;
        const uint32_t fr = (uint32_t)p.frame;
        float sx_ = 0.0f, sy_ = 0.0f, sz_ = 0.0f;
        int grp = -1;   // the particle's staging group
        // spring term of one neighbour within re (first loop of Resample); dx = me - o
        // q is the neighbour's particle index, or its LDS slot when `slot` is set
        // (resolved only on the rare coincident-pair path)
        auto spring = [&](float dx, float dy, float dz, float d2, int32_t q, bool slot) {
            float dist = sqrtf(d2);
            float w = springforce * 1.0f * smooth_fixed(dist * dist, re2, yre2);
            if (dist > 0.1f * re) {
                sx_ += w * dx / dist * re;
                sy_ += w * dy / dist * re;
                sz_ += w * dz / dist * re;
            } else {
                const int32_t qq = slot ? staged_particle(st, q) : q;
                const uint32_t qid = (uint32_t)__float_as_int(pos[qq].w);
                uint32_t hx = hash4(p.seed, 3u * fr + 0u, my_id, qid);
                uint32_t hy = hash4(p.seed, 3u * fr + 1u, my_id, qid);
                uint32_t hz = hash4(p.seed, 3u * fr + 2u, my_id, qid);
                sx_ += 0.01f * re / dt * (float)(int)(hx % 101u) / 100.0f;
                sy_ += 0.01f * re / dt * (float)(int)(hy % 101u) / 100.0f;
                sz_ += 0.01f * re / dt * (float)(int)(hz % 101u) / 100.0f;
            }
        };
        {
            float x = gmax(0.0f, gmin(maxd, maxd * me.x));
            float y = gmax(0.0f, gmin(maxd, maxd * me.y));
            float z = gmax(0.0f, gmin(maxd, maxd * me.z));
            const float xg = maxd * me.x, yg = maxd * me.y, zg = maxd * me.z;
            int xl, xh, yl, yh, zl, zh;
            cell_range(x, N, xl, xh);
            cell_range(y, N, yl, yh);
            cell_range(z, N, zl, zh);
            grp = stage_neighbourhood(st, valid && zl <= zh, xl, xh, yl, yh, zl, zh, N, offsets, pos);
            if (!valid) return;
            // With the neighbourhood staged, in-radius neighbours are queued (LDS
            // index, in visiting order) and their spring terms evaluated in batches:
            // evaluated inline, nearly every iteration of the wave would run the
            // square root and divisions for the few lanes with a neighbour in range.
            int queued = 0;
            // An entry is the neighbour's LDS slot when the block is staged and its
            // particle index otherwise.
            const bool staged = st.ok && grp >= 0;
            auto flush = [&]() {
                for (int t = 0; t < queued; ++t) {
                    const int e = s_queue[t][threadIdx.x];
                    float ox, oy, oz;
                    if (staged) {
                        ox = st.x[e]; oy = st.y[e]; oz = st.z[e];
                    } else {
                        const float4 o = pos[e];
                        ox = o.x; oy = o.y; oz = o.z;
                    }
                    const float dx = me.x - ox, dy = me.y - oy, dz = me.z - oz;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    spring(dx, dy, dz, d2, e, staged);
                }
                queued = 0;
            };
            // Per axis, the cells of the range whose distance bound is within the
            // spring radius, ascending (at most 3, normally 2). Columns are taken in
            // the reference's order (cx, then cy, ascending) from these lists, so a
            // wave iterates about 2 x 2 columns instead of 3 x 3 of which each lane
            // needs some.
            int cxs[3], cys[3], nx = 0, ny = 0;
            float gxs[3], gys[3], gzs[3];
            bool zin[3];
    #pragma unroll
            for (int m = 0; m < 3; ++m) {
                const int c = xl + m;
                const float g = cell_gap(xg, c, margin);
                if (c <= xh && g * g <= gbound) { cxs[nx] = c; gxs[nx] = g; ++nx; }
            }
    #pragma unroll
            for (int m = 0; m < 3; ++m) {
                const int c = yl + m;
                const float g = cell_gap(yg, c, margin);
                if (c <= yh && g * g <= gbound) { cys[ny] = c; gys[ny] = g; ++ny; }
            }
    #pragma unroll
            for (int m = 0; m < 3; ++m) {
                gzs[m] = cell_gap(zg, zl + m, margin);
                zin[m] = zl + m <= zh;
            }
            // The candidate ranges of the columns, (cx, cy) in the reference's order,
            // each narrowed to the z cells within the bound (a contiguous run: the
            // bound grows away from the point), are fetched together before the scan.
            int32_t cq0[9], cq1[9];
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
    #pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const int slot = 3 * a + b;
                    cq0[slot] = cq1[slot] = 0;
                    if (a >= nx || b >= ny) continue;
                    const float gxy2 = gxs[a] * gxs[a] + gys[b] * gys[b];
                    int za = zh + 1, zb = zl - 1;
    #pragma unroll
                    for (int m = 2; m >= 0; --m)
                        if (zin[m] && gxy2 + gzs[m] * gzs[m] <= gbound) za = zl + m;
    #pragma unroll
                    for (int m = 0; m < 3; ++m)
                        if (zin[m] && gxy2 + gzs[m] * gzs[m] <= gbound) zb = zl + m;
                    if (za > zb) continue;
                    const int64_t col = cidx(N, cxs[a], cys[b], 0);
                    cq0[slot] = offsets[col + za];
                    cq1[slot] = offsets[col + zb + 1];
                }
            }
            // Candidates within re (bit-exact skip, below) are queued by LDS slot
            // (staged) or particle index; the queue write is unconditional and only
            // a taken candidate advances it, so the scan has one rare branch.
    #pragma unroll
            for (int slot = 0; slot < 9; ++slot) {
                const int32_t q0 = cq0[slot], q1 = cq1[slot];
                if (q0 >= q1) continue;
                if (staged) {
                    const int l0 = staged_index(st, grp, cxs[slot / 3], cys[slot % 3], q0);
                    for (int32_t q = q0, l = l0; q < q1; ++q, ++l) {
                        const float dx = me.x - st.x[l], dy = me.y - st.y[l], dz = me.z - st.z[l];
                        const float d2 = dx * dx + dy * dy + dz * dz;
                        // Smooth(dist^2, re) is exactly 0 once dist^2 >= re^2, and
                        // then dist > 0.1 re: the term adds +-0 to the spring. d2
                        // above re^2 by a margin far beyond the rounding of sqrt and
                        // the square guarantees that, so the pair is skipped without
                        // changing a bit.
                        const bool take = q != (int32_t)s && d2 <= re2_skip;
                        s_queue[queued][threadIdx.x] = l;
                        queued += take ? 1 : 0;
                        if (queued == kSpringQueue) flush();
                    }
                } else {
                    for (int32_t q = q0; q < q1; ++q) {
                        const float4 o = pos[q];
                        const float dx = me.x - o.x, dy = me.y - o.y, dz = me.z - o.z;
                        const float d2 = dx * dx + dy * dy + dz * dz;
                        const bool take = q != (int32_t)s && d2 <= re2_skip;
                        s_queue[queued][threadIdx.x] = q;
                        queued += take ? 1 : 0;
                        if (queued == kSpringQueue) flush();
                    }
                }
            }
            flush();
        }
        const float4 mv = vel[s];
        const float tx = me.x + dt * sx_, ty = me.y + dt * sy_, tz = me.z + dt * sz_;
        float rx = 0.0f, ry = 0.0f, rz = 0.0f, wsum = 0.0f;
        // With re^2 <= 1e-5 every Sharpen weight max(re^2/max(d2,1e-5) - 1, 0) is 0:
        // the velocity is kept (Resample's "else ru = u"), without visiting neighbours.
        if (re2 > 1.0e-5f) {
            float x = gmax(0.0f, gmin(maxd - 1, maxd * tx));
            float y = gmax(0.0f, gmin(maxd - 1, maxd * ty));
            float z = gmax(0.0f, gmin(maxd - 1, maxd * tz));
            const float xg = maxd * tx, yg = maxd * ty, zg = maxd * tz;
            int xl, xh, yl, yh, zl, zh;
            cell_range(x, N, xl, xh);
            cell_range(y, N, yl, yh);
            cell_range(z, N, zl, zh);
            for (int cx = xl; cx <= xh; ++cx) {
                const float gx = cell_gap(xg, cx, margin);
                for (int cy = yl; cy <= yh; ++cy) {
                    const float gy = cell_gap(yg, cy, margin);
                    const float gxy2 = gx * gx + gy * gy;
                    if (gxy2 > gbound) continue;
                    int za = zl, zb = zh;
                    narrow_z(zg, gxy2, gbound, margin, za, zb);
                    if (za > zb) continue;
                    const int64_t col = cidx(N, cx, cy, 0);
                    const int32_t q0 = offsets[col + za], q1 = offsets[col + zb + 1];
                    // Sharpen(dist2, re) = max(re^2/max(dist2,1e-5) - 1, 0) is exactly 0
                    // when max(dist2, 1e-5) >= re^2; the +0 term is skipped.
                    auto term = [&](int32_t q, float ox, float oy, float oz) {
                        float dist2 = sqrlength(tx, ty, tz, ox, oy, oz);
                        if (gmax(dist2, 1.0e-5f) >= re2) return;
                        float w = 1.0f * sharpen_nr(dist2, re2);
                        const float4 ov = vel[q];
                        rx = rx + w * ov.x;
                        ry = ry + w * ov.y;
                        rz = rz + w * ov.z;
                        wsum += w;
                    };
                    // t lies near the particle, so its cells are normally in its
                    // group's staged box: their positions are read from LDS (the
                    // same bits)
                    if (st.ok && grp >= 0 && staged_has(st, grp, cx, cy, za, zb)) {
                        const int l0 = staged_index(st, grp, cx, cy, q0);
                        for (int32_t q = q0, l = l0; q < q1; ++q, ++l) term(q, st.x[l], st.y[l], st.z[l]);
                    } else {
                        for (int32_t q = q0; q < q1; ++q) {
                            const float4 o = pos[q];
                            term(q, o.x, o.y, o.z);
                        }
                    }
                }
            }
        }
        if (wsum != 0.0f) {
            rx = rx / wsum; ry = ry / wsum; rz = rz / wsum;
        } else {
            rx = mv.x; ry = mv.y; rz = mv.z;
        }
        // accept unless t lies inside a solid (the static walls)
        bool ok = true;
        if (p.wall > 0) {
            const float lo = (float)p.wall, hi = (float)(N - p.wall);
            float gx = tx * maxd, gy = ty * maxd, gz = tz * maxd;
            ok = !(gx < lo || gy < lo || gz < lo || gx > hi || gy > hi || gz > hi);
        }
        accept[s] = ok ? 1 : 0;
        if (ok) {
            pout[s] = make_float4(tx, ty, tz, me.w);
            vout[s] = make_float4(rx, ry, rz, mv.w);
        } else {
            pout[s] = me;
            vout[s] = mv;
        }
    }
};

// Scatter sorted particles back to particle-vector (id) order for the accessors.
struct KUnsort {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float4* pos;
        const float4* vel;
        const uint8_t* accept;
        int64_t np;
        float4* pout;
        float4* vout;
        uint8_t* aout;
        int32_t* ids_out;
    };
    static __device__ void run(const Args& kargs) {
        const float4* __restrict__ pos = kargs.pos;
        const float4* __restrict__ vel = kargs.vel;
        const uint8_t* __restrict__ accept = kargs.accept;
        int64_t np = kargs.np;
        float4* __restrict__ pout = kargs.pout;
        float4* __restrict__ vout = kargs.vout;
        uint8_t* __restrict__ aout = kargs.aout;
        int32_t* __restrict__ ids_out = kargs.ids_out;

        int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (s >= np) return;
        float4 pp = pos[s];
        int32_t id = __float_as_int(pp.w);
        if (pout) pout[id] = pp;
        if (vout) vout[id] = vel[s];
        if (aout && accept) aout[id] = accept[s];
        if (ids_out) ids_out[s] = id;
    }
};

struct KFillI32 {
    static constexpr int kThreads = kBlock;
    struct Args {
        int32_t* p;
        int64_t n;
        int32_t v;
    };
    static __device__ void run(const Args& kargs) {
        int32_t* p = kargs.p;
        int64_t n = kargs.n;
        int32_t v = kargs.v;

        int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i < n) p[i] = v;
    }
};
struct KFillF32 {
    static constexpr int kThreads = kBlock;
    struct Args {
        float* p;
        int64_t n;
        float v;
    };
    static __device__ void run(const Args& kargs) {
        float* p = kargs.p;
        int64_t n = kargs.n;
        float v = kargs.v;

        int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i < n) p[i] = v;
    }
};

}  // namespace

void launch_bin(hipStream_t s, const float4* pos, int64_t np, int N, float maxd, int32_t* cell_of,
                int32_t* counts) {
    if (np == 0) return;
    launch<KBin>(s, dim3(grid_for(np, kBlock)), dim3(kBlock), {pos, np, N, maxd, cell_of, counts});
}

void launch_scan(hipStream_t s, const int32_t* in, int32_t* out, int64_t n, int32_t* block_sums) {
    int nb = (int)((n + kScanTile - 1) / kScanTile);
    if (nb < 1) nb = 1;
    launch<KScanReduce>(s, dim3(nb), dim3(kScanThreads), {in, n, block_sums});
    launch<KScanBsums>(s, dim3(1), dim3(kScanThreads), {block_sums, nb});
    launch<KScanApply>(s, dim3(nb), dim3(kScanThreads), {in, n, block_sums, out, nb});
}

void launch_scatter(hipStream_t s, const int32_t* cell_of, int64_t np, const int32_t* offsets,
                    int32_t* counts, int32_t* slots, const float4* pos, int32_t* sid) {
    if (np == 0) return;
    launch<KScatter>(s, dim3(grid_for(np, kBlock)), dim3(kBlock), {cell_of, np, offsets, counts, slots, pos, sid});
}

void launch_cellsort(hipStream_t s, const int32_t* offsets, int64_t nc, int32_t* sid,
                     int32_t* slots) {
    launch<KCellsort>(s, dim3(grid_for(nc, KCellsort::kThreads)), dim3(KCellsort::kThreads), {offsets, nc, sid, slots});
}

void launch_gather(hipStream_t s, const int32_t* slots, int64_t np, const float4* pos_in,
                   const float4* vel_in, float4* pos_out, float4* vel_out) {
    if (np == 0) return;
    launch<KGather>(s, dim3(grid_for(np, kBlock)), dim3(kBlock), {slots, np, pos_in, vel_in, pos_out, vel_out});
}

void launch_density(hipStream_t s, const Params& p, const float4* pos, float4* vel, int64_t np,
                    const int32_t* offsets, int apply_forces) {
    if (np == 0) return;
    launch<KDensity>(s, dim3((unsigned)xcd_grid(grid_for(np, kStageBlock))), dim3(kStageBlock), {p, pos, vel, np, offsets, apply_forces});
}

void launch_mark(hipStream_t s, const Params& p, const float4* vel, const int32_t* offsets,
                 uint8_t* A, float* L) {
    int64_t nc = (int64_t)p.N * p.N * p.N;
    launch<KMark>(s, dim3(grid_for(nc, kBlock)), dim3(kBlock), {p, vel, offsets, A, L});
}

void launch_p2g(hipStream_t s, const Params& p, const float4* pos, const float4* vel,
                const int32_t* offsets, const uint8_t* A, float* const u[3], float* const du[3],
                float* part, uint8_t* tflag) {
    if (p.p2g == ARIEL_P2G_TILED) {
        const int64_t nb = p2g_tile_count(p.N);
        const dim3 g((unsigned)xcd_grid(nb));
        if (p.splat == 1)
            launch<KP2gTiles<true>>(s, g, dim3(KP2gTiles<true>::kThreads), {p, pos, vel, offsets, part, tflag});
        else
            launch<KP2gTiles<false>>(s, g, dim3(KP2gTiles<false>::kThreads), {p, pos, vel, offsets, part, tflag});
        const int64_t n1 = p.N + 1;
        launch<KP2gCombine>(s, dim3(grid_for(n1 * n1 * n1, KP2gCombine::kThreads)), dim3(KP2gCombine::kThreads),
                            {p, part, tflag, A, u[0], u[1], u[2], du[0], du[1], du[2]});
        return;
    }
    const int64_t nkc = (p.N + 1 + 63) / 64;
    const int64_t nti = (p.N + 1 + p.p2g_tile - 1) / p.p2g_tile;
    const int64_t nb = nti * nti * p.p2g_tile * p.p2g_tile * nkc;
    if (p.splat == 1)
        launch<KP2gCols<true>>(s, dim3((unsigned)xcd_grid(nb)), dim3(64), {p, pos, vel, offsets, A, u[0], u[1], u[2], du[0], du[1], du[2]});
    else
        launch<KP2gCols<false>>(s, dim3((unsigned)xcd_grid(nb)), dim3(64), {p, pos, vel, offsets, A, u[0], u[1], u[2], du[0], du[1], du[2]});
}

void launch_g2p_advect(hipStream_t s, const Params& p, float4* pos, float4* vel, int64_t np,
                       float* const u[3], float* const du[3]) {
    if (np == 0) return;
    launch<KG2pAdvect>(s, dim3((unsigned)xcd_grid(grid_for(np, kBlock))), dim3(kBlock), {p, pos, vel, np, u[0], u[1], u[2], du[0], du[1], du[2]});
}

void launch_resample(hipStream_t s, const Params& p, const float4* pos, const float4* vel,
                     int64_t np, const int32_t* offsets, float4* pos_out, float4* vel_out,
                     uint8_t* accept) {
    if (np == 0) return;
    launch<KResample>(s, dim3((unsigned)xcd_grid(grid_for(np, kStageBlock))), dim3(kStageBlock), {p, pos, vel, np, offsets, pos_out, vel_out, accept});
}

void launch_unsort(hipStream_t s, const float4* pos, const float4* vel, const uint8_t* accept,
                   int64_t np, float4* pos_out, float4* vel_out, uint8_t* accept_out,
                   int32_t* ids_out) {
    if (np == 0) return;
    launch<KUnsort>(s, dim3(grid_for(np, kBlock)), dim3(kBlock), {pos, vel, accept, np, pos_out, vel_out, accept_out, ids_out});
}

void launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v) {
    if (n == 0) return;
    launch<KFillI32>(s, dim3(grid_for(n, kBlock)), dim3(kBlock), {p, n, v});
}
void launch_fill_f32(hipStream_t s, float* p, int64_t n, float v) {
    if (n == 0) return;
    launch<KFillF32>(s, dim3(grid_for(n, kBlock)), dim3(kBlock), {p, n, v});
}

}  // namespace ariel
