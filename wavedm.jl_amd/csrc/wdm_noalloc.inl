// Allocation-free Poisson solve (included at the end of wdm_plan.cpp; shares
// its helpers).  wdm_poisson_noalloc is the variant of wdm_poisson for
// callers that cannot allocate device memory: code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc.  The caller passes
// the density and the potential as device buffers; the work memory the plan
// would own (the R2C half spectrum, the twiddle and k^2 tables) lives in the
// file-scope device array `scratch`, of fixed size, that every GPU the library
// runs on holds until the process ends.  Each call on a GPU reuses that GPU's
// array, so calls that overlap on one GPU race on it.

#include <mutex>

namespace {

constexpr size_t kNoAllocScratchBytes = size_t(WDM_NOALLOC_SCRATCH_BYTES);
constexpr int kNoAllocMaxDevices = 64;

}  // namespace

#ifdef WDM_GPU
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
#else
// emulation build: host memory stands in for the device array
alignas(256) static unsigned char scratch[kNoAllocScratchBytes];
#define HIP_SYMBOL(s) (&(s))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
    *p = const_cast<void*>(sym);
    return hipSuccess;
}
#endif

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device, resolved once per device
// with that device current (each GPU holds its own copy at its own address);
// host code uses this address, never the array's name.
void* scratch_address() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        fail(WDM_ERR_UNSUPPORTED, "wdm_poisson_noalloc: device index " + std::to_string(dev) + " out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) fail(WDM_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Layout of one call's work memory in `scratch`: the half spectrum, then the
// twiddle and k^2 tables of the three axes (each 256-byte aligned).
template <class T>
struct NoAllocLayout {
    size_t hat, tw[3], k2[3], total;
    NoAllocLayout(const int n[3], int hpitch) {
        typedef typename TwEntry<T>::type TwE;
        size_t off = 0;
        hat = off;
        off += align256(sizeof(cx<T>) * size_t(hpitch) * size_t(n[1]) * size_t(n[2]));
        for (int a = 0; a < 3; ++a) {
            tw[a] = off;
            off += align256(sizeof(TwE) * size_t(n[a]));
        }
        for (int a = 0; a < 3; ++a) {
            k2[a] = off;
            off += align256(sizeof(T) * size_t(k2_count(a, n[a])));
        }
        total = off;
    }
};

// The twiddle and k^2 tables of one axis, built on the device by the same
// generators (contraction off) the plan's host code uses: bit-identical to a
// plan's tables, with no host staging, so the call is capturable.
template <class T>
WDM_KERNEL(256)
void k_noalloc_axis_tables(typename TwEntry<T>::type* __restrict__ tw, T* __restrict__ k2, int a, int n, double X) {
    const int j = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (j < n) tw[j] = twiddle_entry<T>(j, n);
    if (j < k2_count(a, n)) k2[j] = k2_entry<T>(a, n, j, X);
}

template <class T>
void poisson_noalloc(const int n[3], const double X[3], const double* rho, double* phi, double scale, hipStream_t s) {
    typedef typename TwEntry<T>::type TwE;
    const AxisOps<T>* ops[3];
    for (int a = 0; a < 3; ++a) {
        ops[a] = axis_ops(n[a], (T*)nullptr);
        if (!ops[a])
            fail(WDM_ERR_UNSUPPORTED, "wdm_poisson_noalloc: axis length " + std::to_string(n[a]) + " is not supported");
    }
    const int hpitch = ((n[0] / 2 + 1) + WDM_HPITCH_ALIGN - 1) / WDM_HPITCH_ALIGN * WDM_HPITCH_ALIGN;
    const NoAllocLayout<T> L(n, hpitch);
    if (L.total > kNoAllocScratchBytes)
        fail(WDM_ERR_UNSUPPORTED, "wdm_poisson_noalloc: the grid needs " + std::to_string(L.total) +
                                      " bytes of work memory, above the " + std::to_string(kNoAllocScratchBytes) +
                                      "-byte scratch array");
    unsigned char* base = static_cast<unsigned char*>(scratch_address());
    for (int a = 0; a < 3; ++a)
        WDM_LAUNCH((k_noalloc_axis_tables<T>), dim3(unsigned((n[a] + 255) / 256)), dim3(256), s,
                   reinterpret_cast<TwE*>(base + L.tw[a]), reinterpret_cast<T*>(base + L.k2[a]), a, n[a], X[a]);
    cx<T>* hat = reinterpret_cast<cx<T>*>(base + L.hat);
    const TwE* tw[3] = {reinterpret_cast<const TwE*>(base + L.tw[0]), reinterpret_cast<const TwE*>(base + L.tw[1]),
                        reinterpret_cast<const TwE*>(base + L.tw[2])};
    const int nlines = n[1] * n[2];
    ops[0]->x_r2c_real(rho, hat, nlines, hpitch, tw[0], s);
    poisson_passes<T>(ops[1], ops[2], n[0], n[1], n[2], hpitch, hat, tw[1], tw[2], base + L.k2[0], base + L.k2[1],
                      base + L.k2[2], scale, s, [](int, bool) {});
    ops[0]->x_c2r_real(hat, phi, nlines, hpitch, tw[0], s);
    hip_check(hipGetLastError(), "wdm_poisson_noalloc launch");
}

}  // namespace

extern "C" {

int wdm_poisson_noalloc(int nx, int ny, int nz, double xmax, double ymax, double zmax, int precision,
                        const double* rho_dev, double* phi_dev, double scale, void* stream) {
    if (!rho_dev || !phi_dev) return fail_code(nullptr, WDM_ERR_ARG, "wdm_poisson_noalloc: NULL buffer");
    if (nx < 1 || ny < 1 || nz < 1) return fail_code(nullptr, WDM_ERR_ARG, "wdm_poisson_noalloc: bad grid size");
    if (!(xmax > 0) || !(ymax > 0) || !(zmax > 0))
        return fail_code(nullptr, WDM_ERR_ARG, "wdm_poisson_noalloc: box half-widths must be positive");
    if (precision != WDM_F64 && precision != WDM_F32)
        return fail_code(nullptr, WDM_ERR_ARG, "wdm_poisson_noalloc: precision must be WDM_F64 or WDM_F32");
    if (!std::isfinite(scale) || scale == 0.0)
        return fail_code(nullptr, WDM_ERR_ARG, "wdm_poisson_noalloc: scale must be finite and nonzero");
    const int n[3] = {nx, ny, nz};
    const double X[3] = {xmax, ymax, zmax};
    hipStream_t s = static_cast<hipStream_t>(stream);
    return guarded(nullptr, [&] {
        if (precision == WDM_F64) poisson_noalloc<double>(n, X, rho_dev, phi_dev, scale, s);
        else poisson_noalloc<float>(n, X, rho_dev, phi_dev, scale, s);
    });
}

}  // extern "C"
