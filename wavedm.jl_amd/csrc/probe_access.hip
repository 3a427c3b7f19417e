// Bandwidth ceiling of the column passes' access patterns at 384^3 complex
// FP32 (453 MB), measured without any FFT arithmetic.  Each kernel reads
// every element of the field once and writes it back in place, the traffic
// of one pass (2 x 453 MB), and the program prints the sustained rate.
//
//   rows-y   8 lanes across rows of stride 384 elements, 8 consecutive x
//            columns per wave, 48 values per lane held in registers: the lane
//            layout of the y passes
//   rows-z   the same layout with row stride 384^2 elements: the z passes
//   planes-z 64 consecutive x columns per wave, one z plane per load
//            instruction, planes walked in chunks of 48
//   stream   64 consecutive elements per load instruction, 48 per lane in
//            flight: the contiguous in-place ceiling
//
// Standalone executable (no library code); run it through bench.py
// --access-probe.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr unsigned kN = 384, kCols = kN * kN, kCells = kN * kN * kN, kHold = 48;

__global__ __launch_bounds__(256) void k_rows(float2* __restrict__ d, unsigned sline, int zaxis) {
    const unsigned lane = threadIdx.x & 63u, wave = blockIdx.x * 4u + threadIdx.x / 64u;
    const unsigned col = wave * 8u + lane / 8u, j = lane % 8u;
    if (col >= kCols) return;
    const unsigned x = col % kN, o = col / kN;
    const unsigned base = zaxis ? (o * kN + x) : (o * kCols + x);
    float2 v[kHold];
#pragma unroll
    for (unsigned i = 0; i < kHold; ++i) v[i] = d[base + (j + 8u * i) * sline];
#pragma unroll
    for (unsigned i = 0; i < kHold; ++i) {
        v[i].x *= 1.0000001f;
        d[base + (j + 8u * i) * sline] = v[i];
    }
}

__global__ __launch_bounds__(256) void k_planes(float2* __restrict__ d) {
    const unsigned lane = threadIdx.x & 63u, wave = blockIdx.x * 4u + threadIdx.x / 64u;
    const unsigned col = wave * 64u + lane;
    if (col >= kCols) return;
    for (unsigned p0 = 0; p0 < kN; p0 += kHold) {
        float2 v[kHold];
#pragma unroll
        for (unsigned i = 0; i < kHold; ++i) v[i] = d[col + (p0 + i) * kCols];
#pragma unroll
        for (unsigned i = 0; i < kHold; ++i) {
            v[i].x *= 1.0000001f;
            d[col + (p0 + i) * kCols] = v[i];
        }
    }
}

// wave w owns the block [w * 64 * kHold, (w + 1) * 64 * kHold)
__global__ __launch_bounds__(256) void k_stream(float2* __restrict__ d) {
    const unsigned lane = threadIdx.x & 63u, wave = blockIdx.x * 4u + threadIdx.x / 64u;
    const unsigned base = wave * 64u * kHold + lane;
    if (base >= kCells) return;
    float2 v[kHold];
#pragma unroll
    for (unsigned i = 0; i < kHold; ++i) v[i] = d[base + 64u * i];
#pragma unroll
    for (unsigned i = 0; i < kHold; ++i) {
        v[i].x *= 1.0000001f;
        d[base + 64u * i] = v[i];
    }
}

int check(hipError_t e, const char* what) {
    if (e != hipSuccess) std::fprintf(stderr, "probe_access: %s: %s\n", what, hipGetErrorString(e));
    return e != hipSuccess;
}

}  // namespace

int main() {
    static_assert(kCols % (8u * 4u) == 0 && kCols % (64u * 4u) == 0 && kN % kHold == 0, "exact grids");
    static_assert(kCells % (64u * kHold * 4u) == 0, "exact stream grid");
    float2* d = nullptr;
    if (check(hipMalloc(&d, sizeof(float2) * size_t(kCells)), "hipMalloc")) return 1;
    if (check(hipMemset(d, 0, sizeof(float2) * size_t(kCells)), "hipMemset")) return 1;
    hipEvent_t a, b;
    if (check(hipEventCreate(&a), "event") || check(hipEventCreate(&b), "event")) return 1;
    const double bytes = 2.0 * sizeof(float2) * double(kCells);
    const char* names[4] = {"rows-y", "rows-z", "planes-z", "stream"};
    const unsigned grids[4] = {kCols / 8u / 4u, kCols / 8u / 4u, kCols / 64u / 4u, kCells / (64u * kHold) / 4u};
    const int iters = 20, warm = 3;
    for (int k = 0; k < 4; ++k) {
        auto launch = [&] {
            if (k == 0) k_rows<<<grids[k], 256>>>(d, kN, 0);
            if (k == 1) k_rows<<<grids[k], 256>>>(d, kCols, 1);
            if (k == 2) k_planes<<<grids[k], 256>>>(d);
            if (k == 3) k_stream<<<grids[k], 256>>>(d);
        };
        for (int w = 0; w < warm; ++w) launch();
        if (check(hipGetLastError(), "launch") || check(hipEventRecord(a), "record")) return 1;
        for (int it = 0; it < iters; ++it) launch();
        if (check(hipEventRecord(b), "record") || check(hipEventSynchronize(b), "sync")) return 1;
        float ms = 0.0f;
        if (check(hipEventElapsedTime(&ms, a, b), "elapsed")) return 1;
        const double us = 1e3 * double(ms) / iters;
        std::printf("{\"pattern\": \"%s\", \"us_per_pass\": %.2f, \"GB_per_s\": %.1f, \"bytes\": %.0f}\n",
                    names[k], us, bytes / (us * 1e-6) / 1e9, bytes);
        std::fflush(stdout);
    }
    check(hipFree(d), "hipFree");
    return 0;
}
