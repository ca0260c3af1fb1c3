// learned_kernels.h — the learned preconditioner's device kernels (BF16 MFMA
// convolutions, staging, output layer), shared by inference (learned.hip) and
// training (learned_train.hip). Each translation unit gets its own copies
// (anonymous namespace).
#pragma once
#include "ariel_device.h"
#include "ariel_internal.h"

// A 4-level U-Net over the grid, BF16 weights and activations, FP32 accumulate:
//   e0 = relu(conv3(x;  2->16))              x = (w r/s, w), full grid
//   e1 = relu(conv3s2(e0; 16->32))           1/2 resolution
//   e2 = relu(conv3s2(e1; 32->48))           1/4
//   e3 = relu(conv3s2(e2; 48->64))           1/8
//   d2 = relu(conv3(e2 + up(proj(e3; 64->48)); 48->48))
//   d1 = relu(conv3(e1 + up(proj(d2; 48->32)); 32->32))
//   d0 = relu(conv3(e0 + up(proj(d1; 32->16)); 16->16))
//   z  = s * w * conv3(d0; 16->1)
// w = sqrt(6 / diag) on FLUID cells (0 elsewhere) is a symmetric diagonal
// scaling: the network sees the matrix with its ghost-fluid diagonal (up to
// ~1e3 next to the free surface) scaled to about one, and w also marks the
// FLUID cells. No biases, so the map is positively homogeneous: the scale s
// (RMS of r over the active cells) only keeps the activations in BF16 range.
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x16bf16_1k: one wave owns
// 16 consecutive output voxels (rows) and every output channel (16 per MFMA
// column block); K runs over taps x input channels in chunks of 16. Operand
// layout of the instruction: A row = lane%16 with k = 4*(lane/16)+j, B column =
// lane%16 with the same k, D column = lane%16 with rows 4*(lane/16)+j.
// Activations are channels-last BF16 (voxel-major, channel contiguous);
// weights are [tap][cout][cin] so that 4 consecutive input channels load as 8
// bytes. Zero padding outside the grid.
namespace ariel {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x32_bf16: D(16x16) += A(16x32) B(32x16). Lane l holds A row
// l%16 and B column l%16 at 8 k-slots; D as for the 16x16x16 form (column l%16,
// rows 4*(l/16)+j). The instruction sums over the slots, so any labelling of
// the slots that A and B share is valid: here lane group g = l/16 takes the 8
// consecutive k = 8g..8g+7 of each 32-wide chunk, which with CIN a multiple of
// 8 are 8 consecutive channels of one tap (one 16-byte load per operand).
__device__ __forceinline__ f32x4 mfma32(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 zero8() {
    return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u));
}
__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bf16x2 as_bf2(uint32_t u) { return __builtin_bit_cast(bf16x2, u); }
// Transpose of the 4x4 block held by the 4 lanes of a quad (lane a of the quad
// holds v[0..3] = M[a][0..3]; afterwards it holds M[0..3][a]): two exchanges
// across the quad with DPP quad permutations, first of lane bit 0 with
// register bit 0, then of bit 1 with bit 1. b0, b1: bits 0 and 1 of the lane
// index within the quad.
__device__ __forceinline__ float quad_swap(float x, int ctrl_is_pairs) {
    const int v = __float_as_int(x);
    return __int_as_float(ctrl_is_pairs ? __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false)    // [2,3,0,1]
                                        : __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false));  // [1,0,3,2]
}
__device__ __forceinline__ void quad_transpose4(float (&v)[4], bool b0, bool b1) {
    const float x0 = quad_swap(v[0], 0), x1 = quad_swap(v[1], 0);
    const float x2 = quad_swap(v[2], 0), x3 = quad_swap(v[3], 0);
    const float n0 = b0 ? x1 : v[0], n1 = b0 ? v[1] : x0;
    const float n2 = b0 ? x3 : v[2], n3 = b0 ? v[3] : x2;
    const float y0 = quad_swap(n0, 1), y1 = quad_swap(n1, 1);
    const float y2 = quad_swap(n2, 1), y3 = quad_swap(n3, 1);
    v[0] = b1 ? y2 : n0;
    v[1] = b1 ? y3 : n1;
    v[2] = b1 ? n2 : y0;
    v[3] = b1 ? n3 : y1;
}

// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
// Two floats rounded to BF16 (nearest even) and packed, lo in the low half:
// one v_cvt_pk_bf16_f32, equal to f2bf of each for finite inputs.
typedef float f32x2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
    const f32x2v f = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, bf16x2));
}

// symmetric diagonal scaling of a FLUID cell: sqrt(6 / diag), IEEE divide and root
__device__ __forceinline__ float diag_weight(float dg) {
    return dg > 0.0f ? sqrtf(6.0f / dg) : 0.0f;
}

__device__ __forceinline__ void decode3(int64_t v, int D, int& i, int& j, int& k) {
    k = (int)(v % D);
    const int64_t q = v / D;
    j = (int)(q % D);
    i = (int)(q / D);
}

// Generic conv: TAPS = 27 (3x3x3, padding 1) or 1 (1x1x1); STRIDE 1 or 2.
// out[v][co] = act( sum_{tap,ci} w[tap][co][ci] * in[src(v,tap)][ci] ) as BF16.
// Every kernel of the network returns at once when `gate` says the solve is
// done or has fallen back to MIC(0).
__device__ __forceinline__ bool gated(const SolverScalars* g) { return g && (g->done || g->use_mic); }

template <int TAPS, int STRIDE, int CIN, int COUT, bool RELU, int SPLIT = 1>
struct KConv {
    static constexpr int kThreads = 256;
    struct Args {
        const uint16_t* in;    // Din^3 x CIN
        const uint16_t* w;     // TAPS x COUT x CIN
        uint16_t* out;         // Dout^3 x COUT
        int Din, Dout;
        const SolverScalars* gate;
    };
    static __device__ void run(const Args& a) {
        if (gated(a.gate)) return;
        const int lane = threadIdx.x & 63;
        // SPLIT waves share 16 output voxels, each taking NB output-channel blocks
        // (more waves for the coarsest level; elsewhere the repeated A loads cost
        // more than the extra parallelism gains)
        const int64_t gw = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
        const int64_t wave = gw / SPLIT;
        const int nsplit = (int)(gw % SPLIT);
        const int64_t nvox = (int64_t)a.Dout * a.Dout * a.Dout;
        const int64_t v0 = wave * 16;
        if (v0 >= nvox) return;
        const int row = lane & 15, kq = lane >> 4;   // A row / k quad
        const int64_t v = v0 + row;
        const bool vin = v < nvox;
        int oi = 0, oj = 0, ok = 0;
        if (vin) decode3(v, a.Dout, oi, oj, ok);
        constexpr int NB = COUT / 16 / SPLIT;
        const int nb0 = nsplit * NB;
        f32x4 acc[NB];
#pragma unroll
        for (int n = 0; n < NB; ++n) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
        // K = taps x input channels, in chunks of 32 (see mfma32); unrolled
        // so that the loads of several chunks are in flight together
        constexpr int K = TAPS * CIN, NCH = (K + 31) / 32;
#pragma unroll 7
        for (int ch = 0; ch < NCH; ++ch) {
            const int k0 = 32 * ch + 8 * kq;
            bf16x8 av = zero8();
            bf16x8 bv[NB];
#pragma unroll
            for (int n = 0; n < NB; ++n) bv[n] = zero8();
            if (k0 < K) {
                const int t = k0 / CIN, c0 = k0 - t * CIN;
                const int di = TAPS == 27 ? t / 9 - 1 : 0, dj = TAPS == 27 ? (t / 3) % 3 - 1 : 0,
                          dk = TAPS == 27 ? t % 3 - 1 : 0;
                const int ii = oi * STRIDE + di, jj = oj * STRIDE + dj, kk = ok * STRIDE + dk;
                if (vin && ii >= 0 && ii < a.Din && jj >= 0 && jj < a.Din && kk >= 0 && kk < a.Din)
                    av = load8(a.in + (((int64_t)ii * a.Din + jj) * a.Din + kk) * CIN + c0);
#pragma unroll
                for (int n = 0; n < NB; ++n)
                    bv[n] = load8(a.w + ((int64_t)t * COUT + (nb0 + n) * 16 + row) * CIN + c0);
            }
#pragma unroll
            for (int n = 0; n < NB; ++n) acc[n] = mfma32(av, bv[n], acc[n]);
        }
        // D: column (output channel) = lane%16, rows (voxels) 4*(lane/16)+j;
        // transposed within quads of lanes (see KConvTile) for 8-byte stores
        const bool b0 = (row & 1) != 0, b1 = (row & 2) != 0;
        const int64_t vo = v0 + kq * 4 + (row & 3);
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            float x[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) x[j] = RELU ? fmaxf(acc[n][j], 0.0f) : acc[n][j];
            quad_transpose4(x, b0, b1);
            if (vo < nvox)
                *reinterpret_cast<uint2*>(a.out + vo * COUT + (nb0 + n) * 16 + (row & ~3)) =
                    make_uint2(f2bf2(x[0], x[1]), f2bf2(x[2], x[3]));
        }
    }
};

// t = e + up(p): nearest-neighbour upsampling of the coarse p (half resolution).
template <int C>
struct KUpAdd {
    static constexpr int kThreads = 256;
    struct Args {
        const uint16_t* e;     // D^3 x C
        const uint16_t* p;     // (D/2)^3 x C
        uint16_t* t;
        int D;
        const SolverScalars* gate;
    };
    static __device__ void run(const Args& a) {
        if (gated(a.gate)) return;
        const int64_t n = (int64_t)a.D * a.D * a.D * C;
        const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (g >= n) return;
        const int c = (int)(g % C);
        const int64_t v = g / C;
        int i, j, k;
        decode3(v, a.D, i, j, k);
        const int H = a.D / 2;
        const int64_t pv = ((int64_t)(i >> 1) * H + (j >> 1)) * H + (k >> 1);
        a.t[g] = f2bf(bf2f(a.e[g]) + bf2f(a.p[pv * C + c]));
    }
};

// ---------------------------------------------------------------- tiled layers
// Full- and half-resolution stride-1 layers stage their input once per
// workgroup in LDS: a workgroup computes a 4x4x16 brick of output voxels (16
// z-rows of 16 voxels, 4 rows per wave) from a 6x6x18 halo, and every tap's A
// operand is an LDS read instead of a global one. Weights (the B operand) come
// from global memory, each load shared by the wave's 4 rows. Modes of the
// staged input:
//   kPlain  in[v]
//   kUpAdd  bf16(e[v] + p[parent(v)])   (the decoder's skip + upsampling)
//   kInput  (w r/s, w) of the network's first layer, K packed as 2*tap + c
namespace tile {
constexpr int TX = 4, TY = 4, TZ = 16;
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
constexpr int HV = HX * HY * HZ;
constexpr int kPlain = 0, kUpAdd = 1, kInput = 2;

__device__ __forceinline__ int64_t bricks(int D) {
    return (int64_t)((D + TX - 1) / TX) * ((D + TY - 1) / TY) * ((D + TZ - 1) / TZ);
}

// brick index -> origin (z fastest, so consecutive bricks of one XCD share halos)
__device__ __forceinline__ void brick_origin(int64_t b, int D, int& x0, int& y0, int& z0) {
    const int nz = (D + TZ - 1) / TZ, ny = (D + TY - 1) / TY;
    z0 = (int)(b % nz) * TZ;
    const int64_t q = b / nz;
    y0 = (int)(q % ny) * TY;
    x0 = (int)(q / ny) * TX;
}

__device__ __forceinline__ float input_scale(const SolverScalars* sc, int64_t nseg) {
    const int64_t ns = nseg >= 0 ? nseg : sc->nseg;
    const double cells = (double)(ns > 0 ? ns : 1) * 8.0;
    return (float)sqrt(sc->rnorm2 / cells);
}

// Stage the (6x6x18) x CIN halo of the brick at (x0, y0, z0) into LDS with
// channel stride CP (bf16 units); zero outside the grid.
template <int CIN, int CP, int MODE>
__device__ __forceinline__ void stage(uint16_t* s_in, const uint16_t* in, const uint16_t* up,
                                      int D, int x0, int y0, int z0) {
    // Each thread first issues the loads of all its entries (a fixed count for
    // the callers' 256 threads) and only then waits, rounds and stores: a loop
    // that waited on each entry's load before the next kept one load in flight.
    constexpr int C8 = CIN / 8;
    constexpr int kTotal = HV * C8;
    constexpr int kNT = 256;
    constexpr int kPer = (kTotal + kNT - 1) / kNT;
    const int H = D >> 1;
    uint4 val[kPer];
    uint4 upv[MODE == kUpAdd ? kPer : 1];
    int dst[kPer];
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
        const int u = threadIdx.x + m * kNT;
        val[m] = make_uint4(0u, 0u, 0u, 0u);
        if constexpr (MODE == kUpAdd) upv[m] = make_uint4(0u, 0u, 0u, 0u);
        dst[m] = -1;
        if (u < kTotal) {
            const int hv = u / C8, c8 = u - hv * C8;
            dst[m] = hv * CP + c8 * 8;
            const int hz = hv % HZ, hy = (hv / HZ) % HY, hx = hv / (HZ * HY);
            const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
            if (gx >= 0 && gx < D && gy >= 0 && gy < D && gz >= 0 && gz < D) {
                const int64_t v = ((int64_t)gx * D + gy) * D + gz;
                val[m] = *reinterpret_cast<const uint4*>(in + v * CIN + c8 * 8);
                if constexpr (MODE == kUpAdd) {
                    const int64_t pv = ((int64_t)(gx >> 1) * H + (gy >> 1)) * H + (gz >> 1);
                    upv[m] = *reinterpret_cast<const uint4*>(up + pv * CIN + c8 * 8);
                }
            }
        }
    }
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
        if (dst[m] < 0) continue;
        if constexpr (MODE == kUpAdd) {
            // e + up(p) rounded to BF16 (outside the grid both are +0 and so is the sum)
            uint32_t* e = reinterpret_cast<uint32_t*>(&val[m]);
            const uint32_t* q = reinterpret_cast<const uint32_t*>(&upv[m]);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                e[c] = f2bf2(bf2f((uint16_t)(e[c] & 0xFFFFu)) + bf2f((uint16_t)(q[c] & 0xFFFFu)),
                             bf2f((uint16_t)(e[c] >> 16)) + bf2f((uint16_t)(q[c] >> 16)));
            }
        }
        uint2* d2 = reinterpret_cast<uint2*>(s_in + dst[m]);
        d2[0] = make_uint2(val[m].x, val[m].y);
        d2[1] = make_uint2(val[m].z, val[m].w);
    }
}
}  // namespace tile

template <int CIN, int COUT, int MODE>
struct KConvTile {
    static constexpr int kThreads = 256;
    // LDS channel stride: 16-channel layers use 16x16x16 MFMAs over one tap each
    // (8-byte rows; measured faster than pairing taps into 16x16x32), wider ones
    // 16x16x32 over 8-channel slices (16-byte rows)
    static constexpr bool K16 = MODE != tile::kInput && CIN == 16;
    static constexpr int CP = MODE == tile::kInput ? 2 : (K16 ? CIN + 4 : CIN + 8);
    static constexpr int NB = COUT / 16;
    // the input layer's weights (2 KB) in LDS; the others' from global memory
    // (LDS-resident 16x16 weights lowered occupancy and ran slower)
    static constexpr int WN = MODE == tile::kInput ? 16 * 64 : 27 * COUT * CIN;
    static constexpr bool WLDS = MODE == tile::kInput;
    struct Args {
        const uint16_t* in;    // kPlain: D^3 x CIN; kUpAdd: the skip e
        const uint16_t* up;    // kUpAdd: coarse p, (D/2)^3 x CIN
        const float* r;        // kInput
        const float* wd;       // kInput: sqrt(6/diag) on FLUID cells, else 0
        const uint16_t* w;     // 27 x COUT x CIN (kInput: 16 x 64)
        uint16_t* out;         // D^3 x COUT, ReLU, BF16
        int D;
        const SolverScalars* sc;
        int64_t nseg;
    };
    static __device__ void run(const Args& a) {
        using namespace tile;
        if (gated(a.sc)) return;
        __shared__ __align__(16) uint16_t s_in[HV * CP];
        __shared__ __align__(16) uint16_t s_w[WLDS ? WN : 8];
        const int D = a.D;
        if constexpr (WLDS) {
            for (int i = threadIdx.x; i < WN / 8; i += blockDim.x)
                reinterpret_cast<uint4*>(s_w)[i] = reinterpret_cast<const uint4*>(a.w)[i];
        }
        const uint16_t* W = WLDS ? s_w : a.w;
        const int64_t nb = bricks(D);
        const int64_t b = xcd_block(nb);
        if (b >= nb) return;
        int x0, y0, z0;
        brick_origin(b, D, x0, y0, z0);
        if constexpr (MODE == kInput) {
            const float s = input_scale(a.sc, a.nseg);
            const float inv = s > 0.0f ? 1.0f / s : 0.0f;
            uint32_t* s32 = reinterpret_cast<uint32_t*>(s_in);
            // all loads first (see stage()), then the pairs
            constexpr int kPer = (HV + kThreads - 1) / kThreads;
            float wv[kPer], rv[kPer];
#pragma unroll
            for (int m = 0; m < kPer; ++m) {
                const int hv = threadIdx.x + m * kThreads;
                wv[m] = 0.0f;
                rv[m] = 0.0f;
                if (hv < HV) {
                    const int hz = hv % HZ, hy = (hv / HZ) % HY, hx = hv / (HZ * HY);
                    const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
                    if (gx >= 0 && gx < D && gy >= 0 && gy < D && gz >= 0 && gz < D) {
                        const int64_t q = ((int64_t)gx * D + gy) * D + gz;
                        wv[m] = a.wd[q];
                        rv[m] = a.r[q];
                    }
                }
            }
#pragma unroll
            for (int m = 0; m < kPer; ++m) {
                const int hv = threadIdx.x + m * kThreads;
                if (hv >= HV) continue;
                const float w = wv[m];
                s32[hv] = w != 0.0f ? f2bf2(rv[m] * inv * w, w) : 0u;
            }
        } else {
            stage<CIN, CP, MODE>(s_in, a.in, a.up, D, x0, y0, z0);
        }
        __syncthreads();

        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int row = lane & 15, kq = lane >> 4;
        f32x4 acc[4][NB];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int n = 0; n < NB; ++n) acc[j][n] = f32x4{0.f, 0.f, 0.f, 0.f};
        if constexpr (MODE == kInput) {
            // K = 2 channels x 27 taps packed as k = 2*tap + c (64 with padding):
            // a lane's 8 slots are 4 consecutive taps, both channels
            const uint32_t* s32 = reinterpret_cast<const uint32_t*>(s_in);
#pragma unroll
            for (int ch = 0; ch < 2; ++ch) {
                const int k0 = 32 * ch + 8 * kq;
                const bf16x8 bv = load8(W + row * 64 + k0);
                const int t0 = k0 >> 1;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int col = wave * 4 + j, cx = col / TY, cy = col % TY;
                    uint32_t p[4];
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const int t = t0 + m;
                        p[m] = 0u;
                        if (t < 27)
                            p[m] = s32[((cx + t / 9) * HY + (cy + (t / 3) % 3)) * HZ + row + t % 3];
                    }
                    const bf16x8 av = __builtin_bit_cast(bf16x8, make_uint4(p[0], p[1], p[2], p[3]));
                    acc[j][0] = mfma32(av, bv, acc[j][0]);
                }
            }
        } else if constexpr (K16) {
            for (int t = 0; t < 27; ++t) {
                const int di = t / 9, dj = (t / 3) % 3, dk = t % 3;
                bf16x4 bv[NB];
#pragma unroll
                for (int n = 0; n < NB; ++n)
                    bv[n] = *reinterpret_cast<const bf16x4*>(W + (t * COUT + n * 16 + row) * CIN + kq * 4);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int col = wave * 4 + j, cx = col / TY, cy = col % TY;
                    const int hv = ((cx + di) * HY + (cy + dj)) * HZ + row + dk;
                    const bf16x4 av = *reinterpret_cast<const bf16x4*>(s_in + hv * CP + kq * 4);
#pragma unroll
                    for (int n = 0; n < NB; ++n)
                        acc[j][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv[n], acc[j][n], 0, 0, 0);
                }
            }
        } else {
            // K = 27 taps x CIN in chunks of 32 (see mfma32)
            constexpr int K = 27 * CIN, NCH = (K + 31) / 32;
#pragma unroll 2
            for (int ch = 0; ch < NCH; ++ch) {
                const int k0 = 32 * ch + 8 * kq;
                const bool live = k0 < K;
                const int t = live ? k0 / CIN : 0, c0 = live ? k0 - t * CIN : 0;
                const int di = t / 9, dj = (t / 3) % 3, dk = t % 3;
                bf16x8 bv[NB];
#pragma unroll
                for (int n = 0; n < NB; ++n)
                    bv[n] = live ? load8(W + (t * COUT + n * 16 + row) * CIN + c0) : zero8();
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int col = wave * 4 + j, cx = col / TY, cy = col % TY;
                    const int hv = ((cx + di) * HY + (cy + dj)) * HZ + row + dk;
                    const bf16x8 av = live ? load8(s_in + hv * CP + c0) : zero8();
#pragma unroll
                    for (int n = 0; n < NB; ++n) acc[j][n] = mfma32(av, bv[n], acc[j][n]);
                }
            }
        }
        // D: column (output channel) = lane%16, rows (z offsets) 4*(lane/16)+i.
        // Each 4x4 block of (lanes 4m..4m+3) x (registers 0..3) is transposed
        // within its quad, so lane 4m+a then holds channels 4m..4m+3 of voxel
        // 4*kq+a and writes them as one 8-byte store (a wave writes 16 voxels x
        // 16 channels as one contiguous 512-byte run per block of channels).
        const bool b0 = (row & 1) != 0, b1 = (row & 2) != 0;
        const int gz = z0 + kq * 4 + (row & 3);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = wave * 4 + j;
            const int gx = x0 + col / TY, gy = y0 + col % TY;
            if (gx >= D || gy >= D) continue;   // uniform over the wave
#pragma unroll
            for (int n = 0; n < NB; ++n) {
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) v[i] = fmaxf(acc[j][n][i], 0.0f);
                quad_transpose4(v, b0, b1);
                if (gz < D) {
                    const int64_t vo = ((int64_t)gx * D + gy) * D + gz;
                    const uint2 w = {f2bf2(v[0], v[1]), f2bf2(v[2], v[3])};
                    *reinterpret_cast<uint2*>(a.out + vo * COUT + n * 16 + (row & ~3)) = w;
                }
            }
        }
    }
};

// Output layer on the VALU: z = s * w * sum_{tap, c} W[tap][c] d0[v + tap][c],
// one thread per voxel of a 4x4x16 brick, d0's halo staged in LDS, FP32. The
// weights are the same for every lane: they are read with scalar loads (kept
// out of LDS, whose bandwidth the halo reads need).
struct KConvOutTile {
    static constexpr int kThreads = 256;
    static constexpr int CP = 20;
    struct Args {
        const uint16_t* in;    // N^3 x 16
        const uint16_t* w;     // 27 x 16 x 16, column 0 used
        const float* wd;       // sqrt(6/diag) on FLUID cells, else 0
        float* z;
        int N;
        const SolverScalars* sc;
        int64_t nseg;
    };
    static __device__ void run(const Args& a) {
        using namespace tile;
        if (gated(a.sc)) return;
        __shared__ __align__(16) uint16_t s_in[HV * CP];
        const uint32_t* __restrict__ w32 = reinterpret_cast<const uint32_t*>(a.w);   // row 0 of tap t at 128 t
        const int N = a.N;
        const int64_t nb = bricks(N);
        const int64_t b = xcd_block(nb);
        if (b >= nb) return;
        int x0, y0, z0;
        brick_origin(b, N, x0, y0, z0);
        stage<16, CP, kPlain>(s_in, a.in, nullptr, N, x0, y0, z0);
        __syncthreads();
        const int cz = threadIdx.x & 15, cy = (threadIdx.x >> 4) & 3, cx = threadIdx.x >> 6;
        const int gx = x0 + cx, gy = y0 + cy, gz = z0 + cz;
        if (gx >= N || gy >= N || gz >= N) return;
        float acc = 0.0f;
#pragma unroll 3
        for (int t = 0; t < 27; ++t) {
            const int hv = ((cx + t / 9) * HY + (cy + (t / 3) % 3)) * HZ + cz + t % 3;
            const uint2* src = reinterpret_cast<const uint2*>(s_in + hv * CP);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint2 v = src[q];
                acc = __builtin_amdgcn_fdot2_f32_bf16(as_bf2(v.x), as_bf2(w32[t * 128 + 2 * q]), acc, false);
                acc = __builtin_amdgcn_fdot2_f32_bf16(as_bf2(v.y), as_bf2(w32[t * 128 + 2 * q + 1]), acc, false);
            }
        }
        const int64_t vo = ((int64_t)gx * N + gy) * N + gz;
        const float s = input_scale(a.sc, a.nseg);
        const float w = a.wd[vo];
        a.z[vo] = w != 0.0f ? acc * s * w : 0.0f;
    }
};

struct KDiagWeight {
    static constexpr int kThreads = 256;
    struct Args {
        const uint8_t* A;
        const float* diag;
        float* wd;
        int64_t nc;
    };
    static __device__ void run(const Args& a) {
        const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= a.nc) return;
        a.wd[c] = a.A[c] == kFluid ? diag_weight(a.diag[c]) : 0.0f;
    }
};

template <class K>
void conv_tiled(hipStream_t s, int D, const typename K::Args& a) {
    const int64_t nb = (int64_t)((D + tile::TX - 1) / tile::TX) * ((D + tile::TY - 1) / tile::TY) *
                       ((D + tile::TZ - 1) / tile::TZ);
    launch<K>(s, dim3((unsigned)xcd_grid(nb)), dim3(256), a);
}

template <class K, int SPLIT = 1>
void conv(hipStream_t s, int Dout, const typename K::Args& a) {
    const int64_t nvox = (int64_t)Dout * Dout * Dout;
    const int64_t waves = (nvox + 15) / 16 * SPLIT;
    launch<K>(s, dim3((unsigned)((waves + 3) / 4)), dim3(256), a);
}

}  // namespace
}  // namespace ariel
