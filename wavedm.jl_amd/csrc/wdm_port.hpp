// Build-mode switch: the HIP build for gfx950 (the product) or the CPU fiber
// emulation build used by the host-side bounds-checking tests.
#pragma once

#include <cstring>
#include <type_traits>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#define WDM_GPU 1
#define WDM_FI __device__ __forceinline__
#define WDM_HDFI __host__ __device__ __forceinline__
#define WDM_KERNEL(nthreads) __global__ __launch_bounds__(nthreads, WDM_MIN_WAVES)
// with a per-kernel minimum of waves per SIMD (caps the registers per lane)
#define WDM_KERNEL_W(nthreads, waves) __global__ __launch_bounds__(nthreads, waves)
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
// Waits for this wave's outstanding vector memory operations (including the
// asynchronous global->LDS copies, which the compiler does not track across
// loop iterations) / its outstanding LDS operations.
#define WDM_WAIT_VMEM() asm volatile("s_waitcnt vmcnt(0)" ::: "memory")
#define WDM_WAIT_LDS() asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory")
#else
#include "emu/emu_runtime.hpp"
#define WDM_EMU 1
#define WDM_FI inline
#define WDM_HDFI inline
#define WDM_KERNEL(nthreads)
#define WDM_KERNEL_W(nthreads, waves)
#define WDM_WAVE_SYNC() emu::barrier()
#define WDM_SHARED static
#define WDM_SCHED_FENCE() ((void)0)
#define WDM_OPAQUE(x) ((void)0)
#define WDM_WAIT_VMEM() ((void)0)
#define WDM_WAIT_LDS() ((void)0)
#endif

#define WDM_UNROLL _Pragma("unroll")

namespace wdm {
constexpr int kWave = 64;  // CDNA wavefront width

// Asynchronous 16-byte-per-lane copy global -> LDS (global_load_lds_dwordx4):
// lane l's 16 bytes at gsrc land at lds_base + 16 l; lds_base must be the same
// for every lane of the wave.  Completion is tracked like a vector load; read
// the LDS only after a wait (the compiler inserts it) and a WDM_WAVE_SYNC.
#ifdef WDM_GPU
__device__ __forceinline__ void lds_dma16(const void* gsrc, void* lds_base) {
    __builtin_amdgcn_global_load_lds(gsrc, (__attribute__((address_space(3))) void*)lds_base, 16, 0, 0);
}
#else
inline void lds_dma16(const void* gsrc, void* lds_base) {
    std::memcpy(static_cast<char*>(lds_base) + 16 * (threadIdx.x & (kWave - 1)), gsrc, 16);
}
#endif
}
