// learned.hip — the learned preconditioner's network on MFMA (gfx950).
//
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
#include "ariel_device.h"
#include "ariel_internal.h"

namespace ariel {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
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

template <int TAPS, int STRIDE, int CIN, int COUT, bool RELU>
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
        const int64_t wave = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
        const int64_t nvox = (int64_t)a.Dout * a.Dout * a.Dout;
        const int64_t v0 = wave * 16;
        if (v0 >= nvox) return;
        const int row = lane & 15, kq = lane >> 4;   // A row / k quad
        const int64_t v = v0 + row;
        const bool vin = v < nvox;
        int oi = 0, oj = 0, ok = 0;
        if (vin) decode3(v, a.Dout, oi, oj, ok);
        constexpr int NB = COUT / 16;
        f32x4 acc[NB];
#pragma unroll
        for (int n = 0; n < NB; ++n) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int t = 0; t < TAPS; ++t) {
            const int di = TAPS == 27 ? t / 9 - 1 : 0, dj = TAPS == 27 ? (t / 3) % 3 - 1 : 0,
                      dk = TAPS == 27 ? t % 3 - 1 : 0;
            const int ii = oi * STRIDE + di, jj = oj * STRIDE + dj, kk = ok * STRIDE + dk;
            const bool inb = vin && ii >= 0 && ii < a.Din && jj >= 0 && jj < a.Din && kk >= 0 &&
                             kk < a.Din;
            const int64_t src = ((int64_t)ii * a.Din + jj) * a.Din + kk;
#pragma unroll
            for (int kc = 0; kc < CIN / 16; ++kc) {
                bf16x4 av = {0, 0, 0, 0};
                if (inb) av = *reinterpret_cast<const bf16x4*>(a.in + src * CIN + kc * 16 + kq * 4);
#pragma unroll
                for (int n = 0; n < NB; ++n) {
                    const bf16x4 bv = *reinterpret_cast<const bf16x4*>(
                        a.w + ((int64_t)t * COUT + n * 16 + row) * CIN + kc * 16 + kq * 4);
                    acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[n], 0, 0, 0);
                }
            }
        }
        // D: column (output channel) = lane%16, rows (voxels) 4*(lane/16)+j
#pragma unroll
        for (int n = 0; n < NB; ++n)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t vo = v0 + kq * 4 + j;
                if (vo < nvox) {
                    float x = acc[n][j];
                    if (RELU) x = fmaxf(x, 0.0f);
                    a.out[vo * COUT + n * 16 + row] = f2bf(x);
                }
            }
    }
};

// First layer: input channels (w r/s, w) built on the fly from the residual,
// the cell classes and the diagonal; K = 27 taps x 2 channels packed as
// k = 2*tap+c, padded to 64 (weights [16][64], zero beyond 54).
struct KConvIn {
    static constexpr int kThreads = 256;
    struct Args {
        const float* r;
        const uint8_t* A;
        const float* diag;
        const uint16_t* w;     // 16 x 64
        uint16_t* out;         // N^3 x 16
        int N;
        const SolverScalars* sc;
        int64_t nseg;          // active segments (-1: from sc)
    };
    static __device__ void run(const Args& a) {
        if (gated(a.sc)) return;
        const int lane = threadIdx.x & 63;
        const int64_t wave = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
        const int N = a.N;
        const int64_t nvox = (int64_t)N * N * N;
        const int64_t v0 = wave * 16;
        if (v0 >= nvox) return;
        const int64_t ns = a.nseg >= 0 ? a.nseg : a.sc->nseg;
        const double cells = (double)(ns > 0 ? ns : 1) * 8.0;
        const float s = (float)sqrt(a.sc->rnorm2 / cells);
        const float inv = s > 0.0f ? 1.0f / s : 0.0f;
        const int row = lane & 15, kq = lane >> 4;
        const int64_t v = v0 + row;
        const bool vin = v < nvox;
        int oi = 0, oj = 0, ok = 0;
        if (vin) decode3(v, N, oi, oj, ok);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) {
            bf16x4 av;
#pragma unroll
            for (int jx = 0; jx < 4; ++jx) {
                const int k = kc * 16 + kq * 4 + jx;
                const int t = k >> 1, c = k & 1;
                float x = 0.0f;
                if (t < 27 && vin) {
                    const int ii = oi + t / 9 - 1, jj = oj + (t / 3) % 3 - 1, kk = ok + t % 3 - 1;
                    if (ii >= 0 && ii < N && jj >= 0 && jj < N && kk >= 0 && kk < N) {
                        const int64_t q = ((int64_t)ii * N + jj) * N + kk;
                        if (a.A[q] == kFluid) {
                            const float w = diag_weight(a.diag[q]);
                            x = c == 0 ? a.r[q] * inv * w : w;
                        }
                    }
                }
                av[jx] = (short)f2bf(x);
            }
            const bf16x4 bv = *reinterpret_cast<const bf16x4*>(a.w + row * 64 + kc * 16 + kq * 4);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t vo = v0 + kq * 4 + j;
            if (vo < nvox) a.out[vo * 16 + row] = f2bf(fmaxf(acc[j], 0.0f));
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

// Output layer: z = s * w * conv3(d0; 16->1), channel 0 of a 16-wide MFMA
// block (weights of channels 1..15 are zero), written as FP32 (0 off FLUID).
struct KConvOut {
    static constexpr int kThreads = 256;
    struct Args {
        const uint16_t* in;    // N^3 x 16
        const uint16_t* w;     // 27 x 16 x 16
        const uint8_t* A;
        const float* diag;
        float* z;
        int N;
        const SolverScalars* sc;
        int64_t nseg;
    };
    static __device__ void run(const Args& a) {
        if (gated(a.sc)) return;
        const int lane = threadIdx.x & 63;
        const int64_t wave = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
        const int N = a.N;
        const int64_t nvox = (int64_t)N * N * N;
        const int64_t v0 = wave * 16;
        if (v0 >= nvox) return;
        const int64_t ns = a.nseg >= 0 ? a.nseg : a.sc->nseg;
        const double cells = (double)(ns > 0 ? ns : 1) * 8.0;
        const float s = (float)sqrt(a.sc->rnorm2 / cells);
        const int row = lane & 15, kq = lane >> 4;
        const int64_t v = v0 + row;
        const bool vin = v < nvox;
        int oi = 0, oj = 0, ok = 0;
        if (vin) decode3(v, N, oi, oj, ok);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int t = 0; t < 27; ++t) {
            const int ii = oi + t / 9 - 1, jj = oj + (t / 3) % 3 - 1, kk = ok + t % 3 - 1;
            const bool inb = vin && ii >= 0 && ii < N && jj >= 0 && jj < N && kk >= 0 && kk < N;
            bf16x4 av = {0, 0, 0, 0};
            if (inb) av = *reinterpret_cast<const bf16x4*>(a.in + (((int64_t)ii * N + jj) * N + kk) * 16 + kq * 4);
            const bf16x4 bv = *reinterpret_cast<const bf16x4*>(a.w + ((int64_t)t * 16 + row) * 16 + kq * 4);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
        if (row == 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t vo = v0 + kq * 4 + j;
                if (vo < nvox) a.z[vo] = a.A[vo] == kFluid ? acc[j] * s * diag_weight(a.diag[vo]) : 0.0f;
            }
        }
    }
};

template <class K>
void conv(hipStream_t s, int Dout, const typename K::Args& a) {
    const int64_t nvox = (int64_t)Dout * Dout * Dout;
    const int64_t waves = (nvox + 15) / 16;
    launch<K>(s, dim3((unsigned)((waves + 3) / 4)), dim3(256), a);
}

}  // namespace

// Offsets (in BF16 elements) of each layer's weights inside the packed blob, in
// the order listed in the file header; returns the total count.
int64_t learned_layout(int64_t* off) {
    const int64_t sizes[] = {
        16 * 64,            // in   2->16 (packed taps)
        27 * 32 * 16,       // e1  16->32 s2
        27 * 48 * 32,       // e2  32->48 s2
        27 * 64 * 48,       // e3  48->64 s2
        64 * 48,            // p2  proj 64->48 (1x1, [cout][cin] = 48 x 64)
        27 * 48 * 48,       // d2
        48 * 32,            // p1  proj 48->32
        27 * 32 * 32,       // d1
        32 * 16,            // p0  proj 32->16
        27 * 16 * 16,       // d0
        27 * 16 * 16,       // out 16->1 (columns 1..15 zero)
    };
    int64_t acc = 0;
    for (int i = 0; i < kLearnedLayers; ++i) {
        if (off) off[i] = acc;
        acc += sizes[i];
    }
    return acc;
}

void launch_learned_apply(hipStream_t s, const LearnedNet& net, int N, const float* r,
                          const uint8_t* A, const float* diag, float* z, const SolverScalars* sc,
                          int64_t nseg) {
    const uint16_t* w = net.weights;
    const int64_t* o = net.off;
    const int D0 = N, D1 = N / 2, D2 = N / 4, D3 = N / 8;
    conv<KConvIn>(s, D0, {r, A, diag, w + o[0], net.e0, N, sc, nseg});
    conv<KConv<27, 2, 16, 32, true>>(s, D1, {net.e0, w + o[1], net.e1, D0, D1, sc});
    conv<KConv<27, 2, 32, 48, true>>(s, D2, {net.e1, w + o[2], net.e2, D1, D2, sc});
    conv<KConv<27, 2, 48, 64, true>>(s, D3, {net.e2, w + o[3], net.e3, D2, D3, sc});
    // decoder, level 2
    conv<KConv<1, 1, 64, 48, false>>(s, D3, {net.e3, w + o[4], net.p, D3, D3, sc});
    {
        const int64_t n = (int64_t)D2 * D2 * D2 * 48;
        launch<KUpAdd<48>>(s, dim3((unsigned)((n + 255) / 256)), dim3(256), {net.e2, net.p, net.t, D2, sc});
    }
    conv<KConv<27, 1, 48, 48, true>>(s, D2, {net.t, w + o[5], net.d2, D2, D2, sc});
    // level 1
    conv<KConv<1, 1, 48, 32, false>>(s, D2, {net.d2, w + o[6], net.p, D2, D2, sc});
    {
        const int64_t n = (int64_t)D1 * D1 * D1 * 32;
        launch<KUpAdd<32>>(s, dim3((unsigned)((n + 255) / 256)), dim3(256), {net.e1, net.p, net.t, D1, sc});
    }
    conv<KConv<27, 1, 32, 32, true>>(s, D1, {net.t, w + o[7], net.d1, D1, D1, sc});
    // level 0
    conv<KConv<1, 1, 32, 16, false>>(s, D1, {net.d1, w + o[8], net.p, D1, D1, sc});
    {
        const int64_t n = (int64_t)D0 * D0 * D0 * 16;
        launch<KUpAdd<16>>(s, dim3((unsigned)((n + 255) / 256)), dim3(256), {net.e0, net.p, net.t, D0, sc});
    }
    conv<KConv<27, 1, 16, 16, true>>(s, D0, {net.t, w + o[9], net.d0, D0, D0, sc});
    conv<KConvOut>(s, D0, {net.d0, w + o[10], A, diag, z, N, sc, nseg});
}

}  // namespace ariel
