// Build-mode switch: the HIP build for gfx950 (the product) or the CPU fiber
// emulation build used by the host-side bounds-checking tests.
#pragma once

#include <type_traits>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#define WDM_GPU 1
#define WDM_FI __device__ __forceinline__
#define WDM_HDFI __host__ __device__ __forceinline__
#define WDM_KERNEL(nthreads) __global__ __launch_bounds__(nthreads, WDM_MIN_WAVES)
#ifndef WDM_MIN_WAVES
#define WDM_MIN_WAVES 1
#endif
// Orders LDS traffic between the lanes of one wavefront (program order inside a
// wave is kept by the LDS unit; this stops the compiler reordering across it).
#define WDM_WAVE_SYNC()                                   \
    do {                                                  \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); \
        __builtin_amdgcn_wave_barrier();                  \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); \
    } while (0)
#define WDM_SHARED __shared__
// Keeps the scheduler from hoisting the next chunk's table loads above this
// point (bounds register pressure in the unrolled line FFTs).
#define WDM_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
// Makes a register value opaque to the optimiser at this point, so work that
// depends on it cannot be hoisted above (used to keep FP64 conversions at
// their point of use).
#define WDM_OPAQUE(x) asm volatile("" : "+v"(x))
#else
#include "emu/emu_runtime.hpp"
#define WDM_EMU 1
#define WDM_FI inline
#define WDM_HDFI inline
#define WDM_KERNEL(nthreads)
#define WDM_WAVE_SYNC() emu::barrier()
#define WDM_SHARED static
#define WDM_SCHED_FENCE() ((void)0)
#define WDM_OPAQUE(x) ((void)0)
#endif

#define WDM_UNROLL _Pragma("unroll")

namespace wdm {
constexpr int kWave = 64;  // CDNA wavefront width
}
