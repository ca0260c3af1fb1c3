// learned_train.hip — training the learned preconditioner on the GPU with this
// library's own kernels. The forward pass is the inference network
// (learned_kernels.h: BF16 MFMA convolutions) run with one buffer per layer so
// that every activation is kept; the backward pass below propagates FP32
// gradients through it (the BF16 roundings of activations and weights pass the
// gradient through unchanged), reduces weight gradients per block in a fixed
// order, and Adam updates FP32 master weights, from which the BF16 weights of
// the next forward pass are rounded.
//
// Loss per pair (r, e) with A e = r: the signed A-angle 1 - (r.z)|r.z| / ((r.e)(z.Az))
// between the network output z and e, averaged over the batch; A is the
// pressure matrix given by its diagonal and FLUID mask (off-diagonals -1
// between FLUID neighbours, scaled by N^2).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ariel.h"
#include "learned_kernels.h"

namespace ariel {
namespace {

constexpr int kTB = 256;
constexpr int kMaxRedBlocks = 1024;
constexpr int kWChunks = 16384;   // weight-gradient partials per layer: chunks x taps

__device__ __forceinline__ void block_sum3(double v0, double v1, double v2, double* out) {
    __shared__ double sm[3][kTB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    v2 = wave_sum(v2);
    if (lane == 0) {
        sm[0][w] = v0;
        sm[1][w] = v1;
        sm[2][w] = v2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0, c = 0;
        for (int i = 0; i < kTB / 64; ++i) {
            a += sm[0][i];
            b += sm[1][i];
            c += sm[2][i];
        }
        out[0] = a;
        out[1] = b;
        out[2] = c;
    }
}

// Sums of the first nparts triples of part, in a fixed order, by one whole
// wave (threads 0-63 all call it); lane 0 receives the sums.
__device__ __forceinline__ void wave_sum_parts(const double* part, int nparts, double& s0, double& s1,
                                               double& s2) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = threadIdx.x & 63; i < nparts; i += 64) {
        a += part[3 * i];
        b += part[3 * i + 1];
        c += part[3 * i + 2];
    }
    s0 = wave_sum(a);
    s1 = wave_sum(b);
    s2 = wave_sum(c);
}

// Per-sample set-up: diagonal weights wd = sqrt(6/diag) on FLUID cells, and
// block partials of sum_FLUID r^2 and of the number of 8-cell z-segments that
// hold a FLUID cell (the network's input scale).
struct KTrainPrep {
    static constexpr int kThreads = kTB;
    struct Args {
        const float* r;
        const float* diag;
        const float* fluid;
        float* wd;
        double* part;   // 3 per block
        int64_t nc;
    };
    static __device__ void run(const Args& a) {
        const int64_t per = (a.nc + (int64_t)gridDim.x * kTB - 1) / ((int64_t)gridDim.x * kTB);
        double r2 = 0.0, segs = 0.0;
        for (int64_t k = 0; k < per; ++k) {
            const int64_t c = ((int64_t)blockIdx.x * per + k) * kTB + threadIdx.x;
            if (c >= a.nc) break;
            const bool fl = a.fluid[c] > 0.0f;
            a.wd[c] = fl ? diag_weight(a.diag[c]) : 0.0f;
            if (fl) r2 += (double)a.r[c] * (double)a.r[c];
            if ((c & 7) == 0) {
                bool any = false;
                for (int m = 0; m < 8; ++m) any = any || a.fluid[c + m] > 0.0f;
                segs += any ? 1.0 : 0.0;
            }
        }
        block_sum3(r2, segs, 0.0, a.part + 3 * blockIdx.x);
    }
};

struct KTrainPrepFinal {
    static constexpr int kThreads = 64;
    struct Args {
        const double* part;
        int nparts;
        SolverScalars* sc;
    };
    static __device__ void run(const Args& a) {
        double r2, segs, unused;
        wave_sum_parts(a.part, a.nparts, r2, segs, unused);
        if (threadIdx.x != 0) return;
        a.sc->rnorm2 = r2;
        a.sc->nseg = (int64_t)segs;
        a.sc->done = 0;
        a.sc->use_mic = 0;
    }
};

// The network's input as the first layer saw it: (bf16(w r / s), bf16(w)) as floats.
struct KTrainInput {
    static constexpr int kThreads = kTB;
    struct Args {
        const float* r;
        const float* wd;
        const SolverScalars* sc;
        float* x;   // nc x 2
        int64_t nc;
    };
    static __device__ void run(const Args& a) {
        const int64_t c = (int64_t)blockIdx.x * kTB + threadIdx.x;
        if (c >= a.nc) return;
        const float s = tile::input_scale(a.sc, -1);
        const float inv = s > 0.0f ? 1.0f / s : 0.0f;
        const float w = a.wd[c];
        float x0 = 0.0f, x1 = 0.0f;
        if (w != 0.0f) {
            x0 = bf2f(f2bf(a.r[c] * inv * w));
            x1 = bf2f(f2bf(w));
        }
        a.x[2 * c] = x0;
        a.x[2 * c + 1] = x1;
    }
};

// Az with A given by diag and the FLUID mask, scaled by N^2.
struct KApplyA {
    static constexpr int kThreads = kTB;
    struct Args {
        const float* z;
        const float* diag;
        const float* fluid;
        float* Az;
        int N;
    };
    static __device__ void run(const Args& a) {
        const int N = a.N;
        const int64_t c = (int64_t)blockIdx.x * kTB + threadIdx.x;
        if (c >= (int64_t)N * N * N) return;
        if (!(a.fluid[c] > 0.0f)) {
            a.Az[c] = 0.0f;
            return;
        }
        int i, j, k;
        decode3(c, N, i, j, k);
        float nb = 0.0f;
        const int q[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                             {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        for (int m = 0; m < 6; ++m) {
            const int x = q[m][0], y = q[m][1], w = q[m][2];
            if (x < 0 || x >= N || y < 0 || y >= N || w < 0 || w >= N) continue;
            const int64_t o = ((int64_t)x * N + y) * N + w;
            if (a.fluid[o] > 0.0f) nb += a.z[o];
        }
        a.Az[c] = (a.diag[c] * a.z[c] - nb) * (float)N * (float)N;
    }
};

struct KDots3 {
    static constexpr int kThreads = kTB;
    struct Args {
        const float* r;
        const float* e;
        const float* z;
        const float* Az;
        double* part;
        int64_t nc;
    };
    static __device__ void run(const Args& a) {
        const int64_t per = (a.nc + (int64_t)gridDim.x * kTB - 1) / ((int64_t)gridDim.x * kTB);
        double rz = 0.0, re = 0.0, zaz = 0.0;
        for (int64_t k = 0; k < per; ++k) {
            const int64_t c = ((int64_t)blockIdx.x * per + k) * kTB + threadIdx.x;
            if (c >= a.nc) break;
            const double r = a.r[c], z = a.z[c];
            rz += r * z;
            re += r * (double)a.e[c];
            zaz += z * (double)a.Az[c];
        }
        block_sum3(rz, re, zaz, a.part + 3 * blockIdx.x);
    }
};

// loss_i = 1 - rz |rz| / (re zAz) (the A-angle, signed so that a network
// pointing away from the solution is not a minimum: flexible PCG needs
// z.r > 0); dL/dz = a r + b Az with a = -2 |rz| / (re zAz) and
// b = 2 rz |rz| / (re zAz^2), both divided by the batch size.
struct KLossFinal {
    static constexpr int kThreads = 64;
    struct Args {
        const double* part;
        int nparts;
        double* loss_acc;
        float* coef;
        float inv_batch;
    };
    static __device__ void run(const Args& a) {
        double rz, re, zaz;
        wave_sum_parts(a.part, a.nparts, rz, re, zaz);
        if (threadIdx.x != 0) return;
        re = re > 1e-30 ? re : 1e-30;
        zaz = zaz > 1e-30 ? zaz : 1e-30;
        const double arz = rz < 0.0 ? -rz : rz;
        *a.loss_acc += 1.0 - rz * arz / (re * zaz);
        a.coef[0] = (float)(-2.0 * arz / (re * zaz) * a.inv_batch);
        a.coef[1] = (float)(2.0 * rz * arz / (re * zaz * zaz) * a.inv_batch);
    }
};

// Gradient at the output layer's pre-scale output o (z = s w o): s w (a r + b Az).
struct KOutGrad {
    static constexpr int kThreads = kTB;
    struct Args {
        const float* r;
        const float* Az;
        const float* wd;
        const float* coef;
        const SolverScalars* sc;
        float* g;
        int64_t nc;
    };
    static __device__ void run(const Args& a) {
        const int64_t c = (int64_t)blockIdx.x * kTB + threadIdx.x;
        if (c >= a.nc) return;
        const float s = tile::input_scale(a.sc, -1);
        a.g[c] = s * a.wd[c] * (a.coef[0] * a.r[c] + a.coef[1] * a.Az[c]);
    }
};

// G *= (Y > 0): the ReLU's gradient mask from the layer's stored output.
struct KReluMask {
    static constexpr int kThreads = kTB;
    struct Args {
        float* g;
        const uint16_t* y;
        int64_t n;
    };
    static __device__ void run(const Args& a) {
        const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
        if (i >= a.n) return;
        if (!(bf2f(a.y[i]) > 0.0f)) a.g[i] = 0.0f;
    }
};

// t = e + up(p): dP[pv][c] = sum over the 8 children of dT (dE = dT is the same buffer).
template <int C>
struct KUpBwd {
    static constexpr int kThreads = kTB;
    struct Args {
        const float* gt;   // D^3 x C
        float* gp;         // (D/2)^3 x C
        int D;
    };
    static __device__ void run(const Args& a) {
        const int H = a.D / 2;
        const int64_t n = (int64_t)H * H * H * C;
        const int64_t g = (int64_t)blockIdx.x * kTB + threadIdx.x;
        if (g >= n) return;
        const int c = (int)(g % C);
        int i, j, k;
        decode3(g / C, H, i, j, k);
        float s = 0.0f;
        for (int m = 0; m < 8; ++m) {
            const int64_t v = ((int64_t)(2 * i + (m >> 2)) * a.D + 2 * j + ((m >> 1) & 1)) * a.D + 2 * k + (m & 1);
            s += a.gt[v * C + c];
        }
        a.gp[g] = s;
    }
};

// dX[u][ci] (+)= sum_t sum_co W[t][co][ci] G[v][co] over the v with
// u = S v + off(t); W is BF16 with (t, co, ci) at w[(t*WCO + co)*CIN + ci].
template <int TAPS, int S, int CIN, int COUT, int WCO>
struct KBwdData {
    static constexpr int kThreads = kTB;
    struct Args {
        const float* g;
        const uint16_t* w;
        float* dx;
        int Din, Dout;
        int acc;
    };
    static constexpr int NCI = CIN / 4;   // each thread: one voxel, 4 input channels
    static __device__ void run(const Args& a) {
        static_assert(CIN % 4 == 0, "input channels in fours");
        const int64_t total = (int64_t)a.Din * a.Din * a.Din * NCI;
        const int64_t gid = (int64_t)blockIdx.x * kTB + threadIdx.x;
        if (gid >= total) return;
        const int cb = (int)(gid % NCI) * 4;
        const int64_t u = gid / NCI;
        int ui, uj, uk;
        decode3(u, a.Din, ui, uj, uk);
        float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int t = 0; t < TAPS; ++t) {
            const int di = TAPS == 27 ? t / 9 - 1 : 0, dj = TAPS == 27 ? (t / 3) % 3 - 1 : 0,
                      dk = TAPS == 27 ? t % 3 - 1 : 0;
            const int xi = ui - di, xj = uj - dj, xk = uk - dk;
            if (S == 2 && ((xi | xj | xk) & 1)) continue;
            const int vi = xi / S, vj = xj / S, vk = xk / S;
            if (xi < 0 || xj < 0 || xk < 0 || vi >= a.Dout || vj >= a.Dout || vk >= a.Dout) continue;
            const float* gv = a.g + (((int64_t)vi * a.Dout + vj) * a.Dout + vk) * COUT;
            const uint16_t* wt = a.w + (int64_t)t * WCO * CIN + cb;
#pragma unroll 4
            for (int co = 0; co < COUT; ++co) {
                const float gc = gv[co];
                const uint2 w2 = *reinterpret_cast<const uint2*>(wt + co * CIN);
                sum[0] += __uint_as_float(w2.x << 16) * gc;
                sum[1] += __uint_as_float(w2.x & 0xFFFF0000u) * gc;
                sum[2] += __uint_as_float(w2.y << 16) * gc;
                sum[3] += __uint_as_float(w2.y & 0xFFFF0000u) * gc;
            }
        }
        float4* d = reinterpret_cast<float4*>(a.dx + u * CIN + cb);
        if (a.acc) {
            const float4 o = *d;
            sum[0] = o.x + sum[0];
            sum[1] = o.y + sum[1];
            sum[2] = o.z + sum[2];
            sum[3] = o.w + sum[3];
        }
        *d = make_float4(sum[0], sum[1], sum[2], sum[3]);
    }
};

// part[chunk][t][co][ci] = sum over the chunk's output voxels v of
// X[S v + off(t)][ci] G[v][co]; blockIdx.y = tap. A chunk is a run of output
// rows (vi, vj); a row whose input row falls outside the grid is skipped whole
// and the k range is clipped once per row, so the voxel loop carries no index
// decode. Each thread owns an RCO x RCI block of (co, ci) pairs (one vector load
// of X and one of G feed RCO*RCI products); the TPV threads covering all pairs
// form a voxel lane, VL lanes split each row's voxels, and the lanes' sums are
// combined through LDS in a fixed order.
template <int TAPS, int S, int CIN, int COUT, bool XF32>
struct KBwdWeight {
    static constexpr int kThreads = kTB;
    static constexpr int PAIRS = CIN * COUT;
    static constexpr int RCI = CIN % 4 == 0 ? 4 : CIN;
    static constexpr int RCO = COUT % 4 == 0 ? 4 : COUT;
    static constexpr int NCI = CIN / RCI;
    static constexpr int TPV = (COUT / RCO) * NCI;
    static constexpr int VL = kTB / TPV;
    struct Args {
        const void* x;
        const float* g;
        float* part;
        int Din, Dout;
        int rows_per_chunk;
    };
    static __device__ void run(const Args& a) {
        static_assert(TPV <= kTB && CIN % RCI == 0 && COUT % RCO == 0, "pair tiling");
        static_assert(RCI == 4 || RCI == 2 || RCI == 1, "x vector width");
        static_assert(RCO == 4 || RCO == 1, "g vector width");
        static_assert(!XF32 || RCI == 2, "the f32 input is read in pairs");
        __shared__ float s_red[VL * PAIRS];
        const int t = blockIdx.y;
        const int di = TAPS == 27 ? t / 9 - 1 : 0, dj = TAPS == 27 ? (t / 3) % 3 - 1 : 0,
                  dk = TAPS == 27 ? t % 3 - 1 : 0;
        const int rows = a.Dout * a.Dout;
        const int r0 = blockIdx.x * a.rows_per_chunk;
        const int r1 = r0 + a.rows_per_chunk < rows ? r0 + a.rows_per_chunk : rows;
        // output k range whose input k = S k + dk lies in [0, Din)
        const int k0 = dk < 0 ? 1 : 0;
        const int kmax = (a.Din - 1 - dk) / S + 1;
        const int k1 = kmax < a.Dout ? kmax : a.Dout;
        const int vl = threadIdx.x / TPV, q = threadIdx.x % TPV;
        const bool live = vl < VL;
        const int cb = (q % NCI) * RCI, ob = (q / NCI) * RCO;
        float acc[RCO][RCI];
#pragma unroll
        for (int m = 0; m < RCO; ++m)
#pragma unroll
            for (int n = 0; n < RCI; ++n) acc[m][n] = 0.0f;
        for (int r = r0; r < r1 && live; ++r) {
            const int vi = r / a.Dout, vj = r - vi * a.Dout;
            const int xi = S * vi + di, xj = S * vj + dj;
            if (xi < 0 || xi >= a.Din || xj < 0 || xj >= a.Din) continue;
            const int64_t xrow = ((int64_t)xi * a.Din + xj) * a.Din + dk;
            const float* grow = a.g + (int64_t)r * a.Dout * COUT + ob;
            for (int vk = k0 + vl; vk < k1; vk += VL) {
                const int64_t xe = (xrow + (int64_t)S * vk) * CIN + cb;
                float xv[RCI], gv[RCO];
                if constexpr (XF32) {
                    const float2 w = *reinterpret_cast<const float2*>(static_cast<const float*>(a.x) + xe);
                    xv[0] = w.x;
                    xv[1] = w.y;
                } else if constexpr (RCI == 4) {
                    const uint2 w = *reinterpret_cast<const uint2*>(static_cast<const uint16_t*>(a.x) + xe);
                    xv[0] = __uint_as_float(w.x << 16);
                    xv[1] = __uint_as_float(w.x & 0xFFFF0000u);
                    xv[2] = __uint_as_float(w.y << 16);
                    xv[3] = __uint_as_float(w.y & 0xFFFF0000u);
                } else {
#pragma unroll
                    for (int n = 0; n < RCI; ++n) xv[n] = bf2f(static_cast<const uint16_t*>(a.x)[xe + n]);
                }
                if constexpr (RCO == 4) {
                    const float4 w = *reinterpret_cast<const float4*>(grow + (int64_t)vk * COUT);
                    gv[0] = w.x;
                    gv[1] = w.y;
                    gv[2] = w.z;
                    gv[3] = w.w;
                } else {
                    gv[0] = grow[(int64_t)vk * COUT];
                }
#pragma unroll
                for (int m = 0; m < RCO; ++m)
#pragma unroll
                    for (int n = 0; n < RCI; ++n) acc[m][n] += xv[n] * gv[m];
            }
        }
        if (live) {
#pragma unroll
            for (int m = 0; m < RCO; ++m)
#pragma unroll
                for (int n = 0; n < RCI; ++n) s_red[vl * PAIRS + (ob + m) * CIN + cb + n] = acc[m][n];
        }
        __syncthreads();
        for (int p = threadIdx.x; p < PAIRS; p += kTB) {
            float sum = 0.0f;
            for (int m = 0; m < VL; ++m) sum += s_red[m * PAIRS + p];
            a.part[((int64_t)blockIdx.x * TAPS + t) * PAIRS + p] = sum;
        }
    }
};

// grad[packed(t, co, ci)] += scale * sum over chunks (fixed order). Layouts:
// 0 [t][co][ci]; 1 the input layer's [co][2t + c] (rows of 64); 2 the output
// layer's [t][16][16], column 0.
struct KReduceW {
    static constexpr int kThreads = kTB;
    static constexpr int kLanes = 16;             // chunk lanes per element
    static constexpr int kPerBlock = kTB / kLanes;
    struct Args {
        const float* part;
        float* grad;
        int nchunks, taps, cout, cin, layout;
    };
    static __device__ void run(const Args& a) {
        __shared__ float s_sum[kTB];
        const int n = a.taps * a.cout * a.cin;
        const int lane = threadIdx.x / kPerBlock;
        const int e = blockIdx.x * kPerBlock + threadIdx.x % kPerBlock;
        float s = 0.0f;
        if (e < n)
            for (int c = lane; c < a.nchunks; c += kLanes) s += a.part[(int64_t)c * n + e];
        s_sum[threadIdx.x] = s;
        __syncthreads();
        if (lane != 0 || e >= n) return;
        s = 0.0f;
        for (int m = 0; m < kLanes; ++m) s += s_sum[m * kPerBlock + threadIdx.x];
        const int ci = e % a.cin, co = (e / a.cin) % a.cout, t = e / (a.cin * a.cout);
        int idx = e;
        if (a.layout == 1) idx = co * 64 + 2 * t + ci;
        else if (a.layout == 2) idx = (t * 16 + co) * 16 + ci;
        a.grad[idx] += s;
    }
};

struct KGradNorm {
    static constexpr int kThreads = kTB;
    struct Args {
        const float* g;
        double* part;
        int64_t n;
    };
    static __device__ void run(const Args& a) {
        const int64_t per = (a.n + (int64_t)gridDim.x * kTB - 1) / ((int64_t)gridDim.x * kTB);
        double s = 0.0;
        for (int64_t k = 0; k < per; ++k) {
            const int64_t i = ((int64_t)blockIdx.x * per + k) * kTB + threadIdx.x;
            if (i >= a.n) break;
            s += (double)a.g[i] * (double)a.g[i];
        }
        block_sum3(s, 0.0, 0.0, a.part + 3 * blockIdx.x);
    }
};

// Adam (Kingma & Ba 2015) with the gradient scaled to norm <= clip; writes the
// BF16 weights of the next forward pass. A batch whose gradient is not finite
// changes nothing.
struct KAdam {
    static constexpr int kThreads = kTB;
    struct Args {
        float* w;
        float* m;
        float* v;
        const float* g;
        uint16_t* wbf;
        const double* part;
        int nparts;
        float lr, clip, b1, b2, eps, bc1, bc2;
        int64_t n;
    };
    static __device__ void run(const Args& a) {
        __shared__ double s_nrm2;
        if (threadIdx.x < 64) {
            double n2, u1, u2;
            wave_sum_parts(a.part, a.nparts, n2, u1, u2);
            if (threadIdx.x == 0) s_nrm2 = n2;
        }
        __syncthreads();
        const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
        if (i >= a.n) return;
        const double nrm2 = s_nrm2;
        if (!(nrm2 < 1e300)) return;   // non-finite gradient: the batch is skipped
        const float nrm = (float)sqrt(nrm2);
        const float sc = a.clip > 0.0f && nrm > a.clip ? a.clip / (nrm + 1e-6f) : 1.0f;
        const float g = a.g[i] * sc;
        const float m = a.b1 * a.m[i] + (1.0f - a.b1) * g;
        const float v = a.b2 * a.v[i] + (1.0f - a.b2) * g * g;
        a.m[i] = m;
        a.v[i] = v;
        const float w = a.w[i] - a.lr * (m / a.bc1) / (sqrtf(v / a.bc2) + a.eps);
        a.w[i] = w;
        a.wbf[i] = f2bf(w);
    }
};

struct KToBf16 {
    static constexpr int kThreads = kTB;
    struct Args {
        const float* w;
        uint16_t* wbf;
        int64_t n;
    };
    static __device__ void run(const Args& a) {
        const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
        if (i < a.n) a.wbf[i] = f2bf(a.w[i]);
    }
};

unsigned blocks(int64_t n) { return (unsigned)((n + kTB - 1) / kTB); }

}  // namespace
}  // namespace ariel

using namespace ariel;

// Buffers of one grid size: every layer's BF16 output and FP32 gradients.
struct TrainSpace {
    int N = 0;
    uint16_t *e0, *e1, *e2, *e3, *p2, *t2, *d2, *p1, *t1, *d1, *p0, *t0, *d0;
    float *wd, *x, *z, *Az, *go;
    float *ge0, *gd0, *gp0, *ge1, *gd1, *gp1, *ge2, *gd2, *gp2, *ge3;
    float* part;
    int64_t part_n = 0;
    std::vector<void*> mem;
};

struct ariel_trainer {
    hipStream_t stream = nullptr;
    int64_t count = 0;
    int64_t off[kLearnedLayers] = {};
    float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t* wbf = nullptr;
    SolverScalars* sc = nullptr;
    double* red = nullptr;      // 3 * kMaxRedBlocks
    double* loss = nullptr;
    float* coef = nullptr;
    int64_t t = 0;              // Adam steps taken
    std::map<int, TrainSpace> spaces;
    std::string err;
};

namespace {

int tfail(ariel_trainer* T, int code, const std::string& msg) {
    if (T) T->err = msg;
    set_last_error(msg);
    return code;
}

template <typename P>
hipError_t talloc(std::vector<void*>& mem, P** p, int64_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, (size_t)(n > 0 ? n : 1) * sizeof(P));
    if (e == hipSuccess) mem.push_back(q);
    *p = static_cast<P*>(q);
    return e;
}

int space_for(ariel_trainer* T, int N, TrainSpace** out) {
    auto it = T->spaces.find(N);
    if (it != T->spaces.end()) {
        *out = &it->second;
        return ARIEL_OK;
    }
    TrainSpace S;
    S.N = N;
    const int64_t v0 = (int64_t)N * N * N, v1 = v0 / 8, v2 = v0 / 64, v3 = v0 / 512;
    // largest weight-gradient partial: kWChunks x the most (co, ci) pairs (64 x 48)
    const int64_t part = (int64_t)kWChunks * 64 * 48;
    bool ok = true;
    auto A = [&](auto** p, int64_t n) { ok = ok && talloc(S.mem, p, n) == hipSuccess; };
    A(&S.e0, 16 * v0); A(&S.e1, 32 * v1); A(&S.e2, 48 * v2); A(&S.e3, 64 * v3);
    A(&S.p2, 48 * v3); A(&S.t2, 48 * v2); A(&S.d2, 48 * v2); A(&S.p1, 32 * v2);
    A(&S.t1, 32 * v1); A(&S.d1, 32 * v1); A(&S.p0, 16 * v1); A(&S.t0, 16 * v0);
    A(&S.d0, 16 * v0);
    A(&S.wd, v0); A(&S.x, 2 * v0); A(&S.z, v0); A(&S.Az, v0); A(&S.go, v0);
    A(&S.ge0, 16 * v0); A(&S.gd0, 16 * v0); A(&S.gp0, 16 * v1); A(&S.ge1, 32 * v1);
    A(&S.gd1, 32 * v1); A(&S.gp1, 32 * v2); A(&S.ge2, 48 * v2); A(&S.gd2, 48 * v2);
    A(&S.gp2, 48 * v3); A(&S.ge3, 64 * v3);
    A(&S.part, part);
    S.part_n = part;
    if (!ok) {
        for (void* p : S.mem) hipFree(p);
        return tfail(T, ARIEL_ERR_NOMEM, "trainer: device allocation failed");
    }
    auto res = T->spaces.emplace(N, std::move(S));
    *out = &res.first->second;
    return ARIEL_OK;
}

// Weight gradient of one layer: partial sums per voxel chunk, then the fixed-order
// reduction into the packed gradient.
template <int TAPS, int S, int CIN, int COUT, bool XF32>
void weight_grad(ariel_trainer* T, TrainSpace& W, const void* x, const float* g, int Din, int Dout,
                 int layer, int layout) {
    // chunks of whole output rows, about 256 voxels each, fewer where that
    // leaves under 2048 blocks (enough to hide the loads' latency), and
    // kWChunks x TAPS partials at most
    const int rows = Dout * Dout;
    int per = (256 + Dout - 1) / Dout;
    const int fill = rows * TAPS / 2048;
    if (fill < per) per = fill > 1 ? fill : 1;
    const int cap = kWChunks / TAPS > 0 ? kWChunks / TAPS : 1;
    if ((rows + per - 1) / per > cap) per = (rows + cap - 1) / cap;
    const int64_t chunks = (rows + per - 1) / per;
    launch<KBwdWeight<TAPS, S, CIN, COUT, XF32>>(T->stream, dim3((unsigned)chunks, TAPS), dim3(kTB),
                                                {x, g, W.part, Din, Dout, per});
    const int n = TAPS * COUT * CIN;
    launch<KReduceW>(T->stream, dim3((unsigned)((n + KReduceW::kPerBlock - 1) / KReduceW::kPerBlock)), dim3(kTB),
                     {W.part, T->grad + T->off[layer], (int)chunks, TAPS, COUT, CIN, layout});
}

template <int TAPS, int S, int CIN, int COUT, int WCO>
void data_grad(ariel_trainer* T, const float* g, int layer, float* dx, int Din, int Dout, int acc) {
    const int64_t n = (int64_t)Din * Din * Din * (CIN / 4);
    launch<KBwdData<TAPS, S, CIN, COUT, WCO>>(T->stream, dim3(blocks(n)), dim3(kTB),
                                             {g, T->wbf + T->off[layer], dx, Din, Dout, acc});
}

void relu_mask(ariel_trainer* T, float* g, const uint16_t* y, int64_t n) {
    launch<KReluMask>(T->stream, dim3(blocks(n)), dim3(kTB), {g, y, n});
}

// Forward pass of one sample with every layer kept; z = the network's output.
void forward(ariel_trainer* T, TrainSpace& W, const float* r, const float* diag, const float* fluid) {
    using namespace tile;
    hipStream_t s = T->stream;
    const int N = W.N, D0 = N, D1 = N / 2, D2 = N / 4, D3 = N / 8;
    const int64_t v0 = (int64_t)N * N * N, v1 = v0 / 8, v2 = v0 / 64, v3 = v0 / 512;
    const uint16_t* w = T->wbf;
    const int64_t* o = T->off;
    SolverScalars* sc = T->sc;
    const int nred = (int)std::min<int64_t>(kMaxRedBlocks, (v0 + kTB - 1) / kTB);

    launch<KTrainPrep>(s, dim3(nred), dim3(kTB), {r, diag, fluid, W.wd, T->red, v0});
    launch<KTrainPrepFinal>(s, dim3(1), dim3(64), {T->red, nred, sc});
    launch<KTrainInput>(s, dim3(blocks(v0)), dim3(kTB), {r, W.wd, sc, W.x, v0});
    // forward, every layer kept
    conv_tiled<KConvTile<2, 16, kInput>>(s, D0, {nullptr, nullptr, r, W.wd, w + o[0], W.e0, D0, sc, -1});
    conv<KConv<27, 2, 16, 32, true>>(s, D1, {W.e0, w + o[1], W.e1, D0, D1, sc});
    conv<KConv<27, 2, 32, 48, true>>(s, D2, {W.e1, w + o[2], W.e2, D1, D2, sc});
    conv<KConv<27, 2, 48, 64, true, 4>, 4>(s, D3, {W.e2, w + o[3], W.e3, D2, D3, sc});
    conv<KConv<1, 1, 64, 48, false>>(s, D3, {W.e3, w + o[4], W.p2, D3, D3, sc});
    launch<KUpAdd<48>>(s, dim3(blocks(48 * v2)), dim3(kTB), {W.e2, W.p2, W.t2, D2, sc});
    conv<KConv<27, 1, 48, 48, true>>(s, D2, {W.t2, w + o[5], W.d2, D2, D2, sc});
    conv<KConv<1, 1, 48, 32, false>>(s, D2, {W.d2, w + o[6], W.p1, D2, D2, sc});
    launch<KUpAdd<32>>(s, dim3(blocks(32 * v1)), dim3(kTB), {W.e1, W.p1, W.t1, D1, sc});
    conv_tiled<KConvTile<32, 32, kPlain>>(s, D1, {W.t1, nullptr, nullptr, nullptr, w + o[7], W.d1, D1, sc, -1});
    conv<KConv<1, 1, 32, 16, false>>(s, D1, {W.d1, w + o[8], W.p0, D1, D1, sc});
    launch<KUpAdd<16>>(s, dim3(blocks(16 * v0)), dim3(kTB), {W.e0, W.p0, W.t0, D0, sc});
    conv_tiled<KConvTile<16, 16, kPlain>>(s, D0, {W.t0, nullptr, nullptr, nullptr, w + o[9], W.d0, D0, sc, -1});
    conv_tiled<KConvOutTile>(s, N, {W.d0, w + o[10], W.wd, W.z, N, sc, -1});
}

// Forward and backward of one sample; the weight gradients accumulate in T->grad.
void sample(ariel_trainer* T, TrainSpace& W, const float* r, const float* e, const float* diag,
            const float* fluid, float inv_batch) {
    forward(T, W, r, diag, fluid);
    hipStream_t s = T->stream;
    const int N = W.N, D0 = N, D1 = N / 2, D2 = N / 4, D3 = N / 8;
    const int64_t v0 = (int64_t)N * N * N, v1 = v0 / 8, v2 = v0 / 64, v3 = v0 / 512;
    SolverScalars* sc = T->sc;
    const int nred = (int)std::min<int64_t>(kMaxRedBlocks, (v0 + kTB - 1) / kTB);
    // loss and dL/do
    launch<KApplyA>(s, dim3(blocks(v0)), dim3(kTB), {W.z, diag, fluid, W.Az, N});
    launch<KDots3>(s, dim3(nred), dim3(kTB), {r, e, W.z, W.Az, T->red, v0});
    launch<KLossFinal>(s, dim3(1), dim3(64), {T->red, nred, T->loss, T->coef, inv_batch});
    launch<KOutGrad>(s, dim3(blocks(v0)), dim3(kTB), {r, W.Az, W.wd, T->coef, sc, W.go, v0});
    // output layer (16 -> 1, no ReLU)
    weight_grad<27, 1, 16, 1, false>(T, W, W.d0, W.go, D0, D0, 10, 2);
    data_grad<27, 1, 16, 1, 16>(T, W.go, 10, W.gd0, D0, D0, 0);
    // d0 = relu(conv(t0)); t0 = e0 + up(p0): dT0 lands in ge0
    relu_mask(T, W.gd0, W.d0, 16 * v0);
    weight_grad<27, 1, 16, 16, false>(T, W, W.t0, W.gd0, D0, D0, 9, 0);
    data_grad<27, 1, 16, 16, 16>(T, W.gd0, 9, W.ge0, D0, D0, 0);
    launch<KUpBwd<16>>(s, dim3(blocks(16 * v1)), dim3(kTB), {W.ge0, W.gp0, D0});
    // p0 = proj(d1)
    weight_grad<1, 1, 32, 16, false>(T, W, W.d1, W.gp0, D1, D1, 8, 0);
    data_grad<1, 1, 32, 16, 16>(T, W.gp0, 8, W.gd1, D1, D1, 0);
    // d1 = relu(conv(t1)); t1 = e1 + up(p1)
    relu_mask(T, W.gd1, W.d1, 32 * v1);
    weight_grad<27, 1, 32, 32, false>(T, W, W.t1, W.gd1, D1, D1, 7, 0);
    data_grad<27, 1, 32, 32, 32>(T, W.gd1, 7, W.ge1, D1, D1, 0);
    launch<KUpBwd<32>>(s, dim3(blocks(32 * v2)), dim3(kTB), {W.ge1, W.gp1, D1});
    // p1 = proj(d2)
    weight_grad<1, 1, 48, 32, false>(T, W, W.d2, W.gp1, D2, D2, 6, 0);
    data_grad<1, 1, 48, 32, 32>(T, W.gp1, 6, W.gd2, D2, D2, 0);
    // d2 = relu(conv(t2)); t2 = e2 + up(p2)
    relu_mask(T, W.gd2, W.d2, 48 * v2);
    weight_grad<27, 1, 48, 48, false>(T, W, W.t2, W.gd2, D2, D2, 5, 0);
    data_grad<27, 1, 48, 48, 48>(T, W.gd2, 5, W.ge2, D2, D2, 0);
    launch<KUpBwd<48>>(s, dim3(blocks(48 * v3)), dim3(kTB), {W.ge2, W.gp2, D2});
    // p2 = proj(e3)
    weight_grad<1, 1, 64, 48, false>(T, W, W.e3, W.gp2, D3, D3, 4, 0);
    data_grad<1, 1, 64, 48, 48>(T, W.gp2, 4, W.ge3, D3, D3, 0);
    // encoder, coarse to fine: e3, e2, e1 (stride 2), then the input layer
    relu_mask(T, W.ge3, W.e3, 64 * v3);
    weight_grad<27, 2, 48, 64, false>(T, W, W.e2, W.ge3, D2, D3, 3, 0);
    data_grad<27, 2, 48, 64, 64>(T, W.ge3, 3, W.ge2, D2, D3, 1);
    relu_mask(T, W.ge2, W.e2, 48 * v2);
    weight_grad<27, 2, 32, 48, false>(T, W, W.e1, W.ge2, D1, D2, 2, 0);
    data_grad<27, 2, 32, 48, 48>(T, W.ge2, 2, W.ge1, D1, D2, 1);
    relu_mask(T, W.ge1, W.e1, 32 * v1);
    weight_grad<27, 2, 16, 32, false>(T, W, W.e0, W.ge1, D0, D1, 1, 0);
    data_grad<27, 2, 16, 32, 32>(T, W.ge1, 1, W.ge0, D0, D1, 1);
    relu_mask(T, W.ge0, W.e0, 16 * v0);
    weight_grad<27, 1, 2, 16, true>(T, W, W.x, W.ge0, D0, D0, 0, 1);
}

}  // namespace

extern "C" {

int ariel_trainer_create(ariel_trainer** out) {
    if (!out) return tfail(nullptr, ARIEL_ERR_INVALID, "trainer_create: null out");
    *out = nullptr;
    ariel_trainer* T = new ariel_trainer();
    T->count = learned_layout(T->off);
    std::vector<void*> mem;
    bool ok = hipStreamCreateWithFlags(&T->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && talloc(mem, &T->w, T->count) == hipSuccess && talloc(mem, &T->m, T->count) == hipSuccess &&
         talloc(mem, &T->v, T->count) == hipSuccess && talloc(mem, &T->grad, T->count) == hipSuccess &&
         talloc(mem, &T->wbf, T->count) == hipSuccess && talloc(mem, &T->sc, 1) == hipSuccess &&
         talloc(mem, &T->red, 3 * kMaxRedBlocks) == hipSuccess && talloc(mem, &T->loss, 1) == hipSuccess &&
         talloc(mem, &T->coef, 2) == hipSuccess;
    if (ok) {
        ok = hipMemsetAsync(T->w, 0, T->count * sizeof(float), T->stream) == hipSuccess &&
             hipMemsetAsync(T->m, 0, T->count * sizeof(float), T->stream) == hipSuccess &&
             hipMemsetAsync(T->v, 0, T->count * sizeof(float), T->stream) == hipSuccess &&
             hipMemsetAsync(T->wbf, 0, T->count * sizeof(uint16_t), T->stream) == hipSuccess &&
             hipMemsetAsync(T->sc, 0, sizeof(SolverScalars), T->stream) == hipSuccess &&
             hipStreamSynchronize(T->stream) == hipSuccess;
    }
    if (!ok) {
        for (void* p : mem) hipFree(p);
        if (T->stream) hipStreamDestroy(T->stream);
        delete T;
        return tfail(nullptr, ARIEL_ERR_HIP, "trainer_create: no device or out of memory");
    }
    *out = T;
    return ARIEL_OK;
}

int ariel_trainer_destroy(ariel_trainer* T) {
    if (!T) return ARIEL_OK;
    hipStreamSynchronize(T->stream);
    for (auto& kv : T->spaces)
        for (void* p : kv.second.mem) hipFree(p);
    for (void* p : {(void*)T->w, (void*)T->m, (void*)T->v, (void*)T->grad, (void*)T->wbf,
                    (void*)T->sc, (void*)T->red, (void*)T->loss, (void*)T->coef})
        hipFree(p);
    hipStreamDestroy(T->stream);
    delete T;
    return ARIEL_OK;
}

int ariel_trainer_set_weights(ariel_trainer* T, const float* w, int64_t n) {
    if (!T || !w || n != T->count) return tfail(T, ARIEL_ERR_INVALID, "trainer_set_weights: bad arguments");
    if (hipMemcpyAsync(T->w, w, n * sizeof(float), hipMemcpyHostToDevice, T->stream) != hipSuccess ||
        hipMemsetAsync(T->m, 0, n * sizeof(float), T->stream) != hipSuccess ||
        hipMemsetAsync(T->v, 0, n * sizeof(float), T->stream) != hipSuccess)
        return tfail(T, ARIEL_ERR_HIP, "trainer_set_weights: copy failed");
    launch<KToBf16>(T->stream, dim3(blocks(n)), dim3(kTB), {T->w, T->wbf, n});
    T->t = 0;
    if (hipStreamSynchronize(T->stream) != hipSuccess) return tfail(T, ARIEL_ERR_HIP, "trainer_set_weights: sync");
    return ARIEL_OK;
}

int ariel_trainer_get_weights(ariel_trainer* T, float* w, int64_t n) {
    if (!T || !w || n != T->count) return tfail(T, ARIEL_ERR_INVALID, "trainer_get_weights: bad arguments");
    if (hipMemcpyAsync(w, T->w, n * sizeof(float), hipMemcpyDeviceToHost, T->stream) != hipSuccess ||
        hipStreamSynchronize(T->stream) != hipSuccess)
        return tfail(T, ARIEL_ERR_HIP, "trainer_get_weights: copy failed");
    return ARIEL_OK;
}

int ariel_trainer_step(ariel_trainer* T, int32_t N, int32_t batch, const float* r, const float* e,
                       const float* diag, const float* fluid, float lr, float clip, int32_t update,
                       double* loss_out, float* grad_out) {
    if (!T) return tfail(T, ARIEL_ERR_INVALID, "trainer_step: null trainer");
    if (N < 16 || N % 16 != 0 || batch < 1 || !r || !e || !diag || !fluid)
        return tfail(T, ARIEL_ERR_INVALID, "trainer_step: N must be a positive multiple of 16, batch >= 1");
    TrainSpace* W = nullptr;
    int rc = space_for(T, N, &W);
    if (rc != ARIEL_OK) return rc;
    const int64_t v0 = (int64_t)N * N * N;
    hipStream_t s = T->stream;
    if (hipMemsetAsync(T->grad, 0, T->count * sizeof(float), s) != hipSuccess ||
        hipMemsetAsync(T->loss, 0, sizeof(double), s) != hipSuccess)
        return tfail(T, ARIEL_ERR_HIP, "trainer_step: memset failed");
    for (int b = 0; b < batch; ++b)
        sample(T, *W, r + b * v0, e + b * v0, diag + b * v0, fluid + b * v0, 1.0f / (float)batch);
    if (update) {
        T->t += 1;
        const int nred = (int)std::min<int64_t>(kMaxRedBlocks, (T->count + kTB - 1) / kTB);
        launch<KGradNorm>(s, dim3(nred), dim3(kTB), {T->grad, T->red, T->count});
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.0f - std::pow(b1, (float)T->t), bc2 = 1.0f - std::pow(b2, (float)T->t);
        launch<KAdam>(s, dim3(blocks(T->count)), dim3(kTB),
                      {T->w, T->m, T->v, T->grad, T->wbf, T->red, nred, lr, clip, b1, b2, 1e-8f, bc1, bc2,
                       T->count});
    }
    if (hipGetLastError() != hipSuccess) return tfail(T, ARIEL_ERR_HIP, "trainer_step: launch failed");
    double loss = 0.0;
    if (hipMemcpyAsync(&loss, T->loss, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
        return tfail(T, ARIEL_ERR_HIP, "trainer_step: readback failed");
    if (grad_out &&
        hipMemcpyAsync(grad_out, T->grad, T->count * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess)
        return tfail(T, ARIEL_ERR_HIP, "trainer_step: gradient readback failed");
    if (hipStreamSynchronize(s) != hipSuccess) return tfail(T, ARIEL_ERR_HIP, "trainer_step: sync failed");
    if (loss_out) *loss_out = loss / batch;
    return ARIEL_OK;
}

int ariel_trainer_forward(ariel_trainer* T, int32_t N, int32_t batch, const float* r, const float* diag,
                          const float* fluid, float* z) {
    if (!T) return tfail(T, ARIEL_ERR_INVALID, "trainer_forward: null trainer");
    if (N < 16 || N % 16 != 0 || batch < 1 || !r || !diag || !fluid || !z)
        return tfail(T, ARIEL_ERR_INVALID, "trainer_forward: N must be a positive multiple of 16, batch >= 1");
    TrainSpace* W = nullptr;
    int rc = space_for(T, N, &W);
    if (rc != ARIEL_OK) return rc;
    const int64_t v0 = (int64_t)N * N * N;
    for (int b = 0; b < batch; ++b) {
        forward(T, *W, r + b * v0, diag + b * v0, fluid + b * v0);
        if (hipMemcpyAsync(z + b * v0, W->z, v0 * sizeof(float), hipMemcpyDeviceToDevice, T->stream) != hipSuccess)
            return tfail(T, ARIEL_ERR_HIP, "trainer_forward: copy failed");
    }
    if (hipStreamSynchronize(T->stream) != hipSuccess) return tfail(T, ARIEL_ERR_HIP, "trainer_forward: sync failed");
    return ARIEL_OK;
}

const char* ariel_trainer_last_error(const ariel_trainer* T) { return T ? T->err.c_str() : ""; }

}  // extern "C"
