// noalloc.hip — the particle level set of ariel_export_levelset for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc). The level-set kernels need per-call
// workspace (per-cell counts and offsets, the scan's block sums, and the
// particles binned by cell); here it lives in one file-scope device array of
// fixed size instead of buffers allocated per call. The caller passes the
// output, a device buffer of (N+1)^3 floats.
//
// Each device holds its own copy of the array, at its own address, from module
// load until the process ends. Host code never names the array in a launch: the
// address comes from hipGetSymbolAddress, taken once per device with that device
// current and kept in a per-device table. Every call on a device shares that
// device's array and it keeps the last call's contents; the per-cell counts are
// cleared by a fill kernel at the start of each call, every other part is
// written before it is read. Calls that overlap on one device (on different
// streams, or from several host threads) race on the array and must be
// serialised by the caller.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/ariel.h"
#include "ariel_internal.h"

namespace ariel {

namespace {

// Limits of the allocation-free level set: grids up to 64^3 cells and up to
// 2^20 particles.
constexpr int64_t kNoAllocMaxN = ARIEL_NOALLOC_MAX_N;
constexpr int64_t kNoAllocMaxParticles = ARIEL_NOALLOC_MAX_PARTICLES;

constexpr int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// Workspace layout for N^3 cells and np particles (byte offsets, each section
// 256-byte aligned): counts[N^3], offsets[N^3 + 1], scan block sums, cell_of[np],
// rank[np], the binned positions (float4)[np].
struct Layout {
    int64_t counts, offsets, sums, cell_of, rank, sorted, total;
};
constexpr Layout layout(int64_t nc, int64_t np) {
    Layout l{};
    l.counts = 0;
    l.offsets = l.counts + align256(4 * nc);
    l.sums = l.offsets + align256(4 * (nc + 1));
    l.cell_of = l.sums + align256(4 * (nc / 4096 + 64));
    l.rank = l.cell_of + align256(4 * np);
    l.sorted = l.rank + align256(4 * np);
    l.total = l.sorted + align256(16 * np);
    return l;
}
constexpr int64_t kScratchBytes =
    layout(kNoAllocMaxN * kNoAllocMaxN * kNoAllocMaxN, kNoAllocMaxParticles).total;

// The workspace (about 26 MiB). Host code must not pass `scratch` itself to a
// kernel: on the host it is a stand-in symbol, not a device address.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the first
// call on that device.
hipError_t scratch_address(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

hipError_t noalloc_levelset(hipStream_t s, int device, const float4* pos, int64_t np, int N,
                            float maxd, float band, float rad, float* phi, std::string& err) {
    if (N > kNoAllocMaxN) {
        err = "allocation-free level set: grid edge " + std::to_string(N) + " above the limit " +
              std::to_string(kNoAllocMaxN);
        return hipErrorInvalidValue;
    }
    if (np > kNoAllocMaxParticles) {
        err = "allocation-free level set: " + std::to_string(np) + " particles above the limit " +
              std::to_string(kNoAllocMaxParticles);
        return hipErrorInvalidValue;
    }
    void* base = nullptr;
    const hipError_t e = scratch_address(device, &base);
    if (e != hipSuccess) {
        err = std::string("allocation-free level set: hipGetSymbolAddress: ") + hipGetErrorString(e);
        return e;
    }
    const int64_t nc = (int64_t)N * N * N;
    const Layout l = layout(nc, np);
    unsigned char* b = static_cast<unsigned char*>(base);
    SdfScratch w;
    w.counts = reinterpret_cast<int32_t*>(b + l.counts);
    w.offsets = reinterpret_cast<int32_t*>(b + l.offsets);
    w.scan_sums = reinterpret_cast<int32_t*>(b + l.sums);
    w.cell_of = reinterpret_cast<int32_t*>(b + l.cell_of);
    w.rank = reinterpret_cast<int32_t*>(b + l.rank);
    w.sorted = reinterpret_cast<float4*>(b + l.sorted);
    launch_particle_sdf(s, pos, np, N, maxd, band, rad, w, phi);
    return hipGetLastError();
}

}  // namespace ariel
