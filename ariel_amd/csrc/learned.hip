// learned.hip — the learned preconditioner's network on MFMA (gfx950):
// weight layout and the inference sequence (kernels in learned_kernels.h).
#include "learned_kernels.h"

namespace ariel {

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
    using namespace tile;
    const uint16_t* w = net.weights;
    const int64_t* o = net.off;
    const int D0 = N, D1 = N / 2, D2 = N / 4, D3 = N / 8;
    conv_tiled<KConvTile<2, 16, kInput>>(s, D0, {nullptr, nullptr, r, net.wd, w + o[0], net.e0, D0, sc, nseg});
    conv<KConv<27, 2, 16, 32, true>>(s, D1, {net.e0, w + o[1], net.e1, D0, D1, sc});
    conv<KConv<27, 2, 32, 48, true>>(s, D2, {net.e1, w + o[2], net.e2, D1, D2, sc});
    conv<KConv<27, 2, 48, 64, true, 4>, 4>(s, D3, {net.e2, w + o[3], net.e3, D2, D3, sc});
    // decoder, level 2 (small: skip-add as its own pass)
    conv<KConv<1, 1, 64, 48, false>>(s, D3, {net.e3, w + o[4], net.p, D3, D3, sc});
    {
        const int64_t n = (int64_t)D2 * D2 * D2 * 48;
        launch<KUpAdd<48>>(s, dim3((unsigned)((n + 255) / 256)), dim3(256), {net.e2, net.p, net.t, D2, sc});
    }
    conv<KConv<27, 1, 48, 48, true>>(s, D2, {net.t, w + o[5], net.d2, D2, D2, sc});
    // levels 1 and 0: skip-add fused into the tiled layer's staging
    conv<KConv<1, 1, 48, 32, false>>(s, D2, {net.d2, w + o[6], net.p, D2, D2, sc});
    conv_tiled<KConvTile<32, 32, kUpAdd>>(s, D1, {net.e1, net.p, nullptr, nullptr, w + o[7], net.d1, D1, sc, nseg});
    conv<KConv<1, 1, 32, 16, false>>(s, D1, {net.d1, w + o[8], net.p, D1, D1, sc});
    conv_tiled<KConvTile<16, 16, kUpAdd>>(s, D0, {net.e0, net.p, nullptr, nullptr, w + o[9], net.d0, D0, sc, nseg});
    conv_tiled<KConvOutTile>(s, N, {net.d0, w + o[10], net.wd, z, N, sc, nseg});
}

void launch_diag_weight(hipStream_t s, int N, const uint8_t* A, const float* diag, float* wd) {
    const int64_t nc = (int64_t)N * N * N;
    launch<KDiagWeight>(s, dim3((unsigned)((nc + 255) / 256)), dim3(256), {A, diag, wd, nc});
}

}  // namespace ariel
