// CPU emulation of the small HIP subset the wavedm kernels use.
//
// DEBUG/TEST BUILD ONLY (libwavedm_cpuemu.so).  It lets the exact kernel
// source run on the host under AddressSanitizer so every global and LDS access
// is bounds-checked before a kernel ever reaches a GPU.  One block runs at a
// time; each GPU thread is a ucontext fiber and __syncthreads()/wave barriers
// switch fibers.  The shipped library (libwavedm_hip.so) never contains this.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>

struct emu_dim3 {
    unsigned x, y, z;
    emu_dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
typedef emu_dim3 dim3;

namespace emu {
struct Ctx {
    emu_dim3 tid, bid, bdim, gdim;
};
extern Ctx g_ctx;
void barrier();            // block-wide barrier (also used for wave barriers)
void launch(dim3 grid, dim3 block, const std::function<void()>& body);
[[noreturn]] void fail(const char* msg);
}  // namespace emu

#define threadIdx (emu::g_ctx.tid)
#define blockIdx (emu::g_ctx.bid)
#define blockDim (emu::g_ctx.bdim)
#define gridDim (emu::g_ctx.gdim)
#define __syncthreads() emu::barrier()

// Fibers switch only at barriers, so a read-modify-write is atomic here.
inline unsigned long long atomicAdd(unsigned long long* a, unsigned long long v) {
    const unsigned long long o = *a;
    *a = o + v;
    return o;
}

// ---- fake runtime API (host memory stands in for device memory) ----------
typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                     hipMemcpyDefault };
inline hipError_t hipMalloc(void** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    *p = std::malloc(n);
    if (!*p) return hipErrorOutOfMemory;
    // poison so uninitialised reads show up as NaNs in float data
    std::memset(*p, 0xff, n);
    return hipSuccess;
}
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    if (n) std::memmove(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
    return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
    std::memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipStreamCreate(hipStream_t* s) { *s = nullptr; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipPeekAtLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "emu error"; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = nullptr; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
// streams are synchronous here: waits are no-ops, graphs are never captured
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
constexpr unsigned hipEventDisableTiming = 2;
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = nullptr; return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
