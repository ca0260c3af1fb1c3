// selftest.hip — exhaustive comparison of the kernels' division routines with
// IEEE division over every float in a range (bit patterns lo..hi), and of the
// network's hardware BF16 rounding with the integer round-to-nearest-even.
#include "ariel_device.h"
#include "ariel_internal.h"
#include "learned_kernels.h"

namespace ariel {
namespace {

__global__ void k_div_check(int which, float a, uint32_t lo, uint32_t hi,
                            unsigned long long* mismatches) {
    const float ya = 1.0f / a;
    unsigned long long bad = 0;
    const uint64_t n = (uint64_t)hi - lo + 1;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const float x = __uint_as_float(lo + (uint32_t)t);
        bool same;
        if (which == 0) {
            const float ref = x / a;
            const float got = div_fixed(x, a, ya);
            if (x >= 0x1p-100f) {
                same = __float_as_uint(ref) == __float_as_uint(got);
            } else {
                same = __float_as_uint(gmax(1.0f - ref, 0.0f)) ==
                       __float_as_uint(gmax(1.0f - got, 0.0f));
            }
        } else if (which == 1) {
            const float ref = a / x;
            const float got = div_nr(a, x);
            same = __float_as_uint(ref) == __float_as_uint(got);
        } else {
            // the numerator is fixed (kSharpenHH); a is unused
            const float ref = kSharpenHH / x;
            const float got = div_sharpen_hh(x);
            same = __float_as_uint(ref) == __float_as_uint(got);
        }
        bad += same ? 0 : 1;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) bad += __shfl_down(bad, off, 64);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(mismatches, bad);
}

// Every pair of consecutive float bit patterns (2t, 2t+1): f2bf2 (one
// v_cvt_pk_bf16_f32) against f2bf on each finite input.
__global__ void k_bf16_check(unsigned long long* mismatches) {
    unsigned long long bad = 0;
    const uint64_t n = 1ull << 31;   // pairs
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const float a0 = __uint_as_float((uint32_t)(2 * t)), a1 = __uint_as_float((uint32_t)(2 * t + 1));
        const uint32_t hw = f2bf2(a0, a1);
        const uint32_t sw = (uint32_t)f2bf(a0) | ((uint32_t)f2bf(a1) << 16);
        const uint32_t m = (isfinite(a0) ? 0xFFFFu : 0u) | (isfinite(a1) ? 0xFFFF0000u : 0u);
        bad += ((hw ^ sw) & m) != 0 ? 1 : 0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) bad += __shfl_down(bad, off, 64);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(mismatches, bad);
}

}  // namespace

hipError_t run_bf16_selftest(int64_t* mismatches) {
    unsigned long long* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    e = hipMemset(d, 0, sizeof(unsigned long long));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bf16_check, dim3(8192), dim3(256), 0, 0, d);
        e = hipGetLastError();
    }
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
    hipFree(d);
    *mismatches = (int64_t)h;
    return e;
}

hipError_t run_division_selftest(int which, float a, float lo, float hi, int64_t* mismatches) {
    unsigned long long* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    e = hipMemset(d, 0, sizeof(unsigned long long));
    if (e == hipSuccess) {
        const uint32_t blo = __builtin_bit_cast(uint32_t, lo);
        const uint32_t bhi = __builtin_bit_cast(uint32_t, hi);
        hipLaunchKernelGGL(k_div_check, dim3(8192), dim3(256), 0, 0, which, a, blo, bhi, d);
        e = hipGetLastError();
    }
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
    hipFree(d);
    *mismatches = (int64_t)h;
    return e;
}

}  // namespace ariel
