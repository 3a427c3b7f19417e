// Per-axis-length launcher tables.  Each supported line length N has its own
// translation unit (wdm_inst_<N>.cpp) that instantiates every pass kernel for
// FP32 and FP64 and fills one AxisOps table per precision; the plan picks the
// tables for Nx, Ny and Nz at creation time.
#pragma once

#include "wdm_kernels.hpp"
#include "wdm_adj_kernels.hpp"

#ifdef WDM_GPU
namespace wdm {
// Optional start/stop events attached to the next pass launch (per-kernel
// timing; null events launch normally).
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
inline LaunchEvents& launch_events() {
    static thread_local LaunchEvents ev;
    return ev;
}
}  // namespace wdm
#define WDM_LAUNCH(KERNEL, GRID, BLOCK, STREAM, ...)                                                      \
    hipExtLaunchKernelGGL(KERNEL, GRID, BLOCK, 0, STREAM, wdm::launch_events().start, wdm::launch_events().stop, \
                          0, __VA_ARGS__)
#else
#define WDM_LAUNCH(KERNEL, GRID, BLOCK, STREAM, ...) \
    emu::launch(GRID, BLOCK, [&]() { KERNEL(__VA_ARGS__); })
#endif

namespace wdm {

// Waves per workgroup of the pass kernels, per precision.  Measured at 384^3
// on MI355X (autotuner): FP32 kernels (130-200 registers, two or three waves
// per SIMD) do best with 4-wave groups, which leave no wave slots unused at
// three waves per SIMD; FP64 kernels (256 VGPRs + AGPRs, one wave per SIMD)
// with 2.
#ifndef WDM_NW_F32
#define WDM_NW_F32 4
#endif
#ifndef WDM_NW_F64
#define WDM_NW_F64 2
#endif
template <class T>
constexpr int kNW = sizeof(T) == 4 ? WDM_NW_F32 : WDM_NW_F64;
// tiles per wave of the column passes (1: one tile per wave, no prefetch)
#ifndef WDM_COL_TPW
#define WDM_COL_TPW 1
#endif
// workgroup size of the fused z passes when they ask for more than one wave
// per SIMD: 4 for the drift pass, 2 for the Poisson pass (384^3 FP32: drift
// 256 us at 4, 291 at 2 or 8; Poisson 125 us at 2, 131 at 4, 140 at 8, 143 at 1)
#ifndef WDM_NW_FMI
#define WDM_NW_FMI 4
#endif
#ifndef WDM_NW_PFMI
#define WDM_NW_PFMI 2
#endif
template <int MODE, int KF, int MINW, int NW>
constexpr int fmi_group() {
    return (MODE == COL_FMI && MINW > 1) ? (KF == KF_POISSON ? WDM_NW_PFMI : WDM_NW_FMI) : NW;
}
// LDS-staged persistent column passes (k_cols) where the tile shape allows
#ifndef WDM_COL_STAGED
#define WDM_COL_STAGED 0
#endif

// compute units of the current device (cached for the process)
inline int device_cus() {
#ifdef WDM_GPU
    static int cus = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            v = 256;
        return v;
    }();
    return cus;
#else
    return 4;
#endif
}

template <class T>
struct AxisOps {
    typedef typename TwEntry<T>::type TwE;
    int n, n1, n2, lpw;
    int nw;  // waves per workgroup of every pass kernel of this table
    void (*col)(int mode, int kf, cx<T>* data, int ncols, int nother, unsigned sline, unsigned sother,
                const TwE* tw, KTabs kt, hipStream_t s);
    void (*x_kick)(bool fwd, cx<T>* psi, const cx<T>* phihat, int nlines, int hpitch, const TwE* tw,
                   const KickArgs& ka, hipStream_t s);  // kernel picked by the potential kind of ka.vm
    void (*x_inv_dens)(cx<T>* psi, cx<T>* rhohat, int nlines, int hpitch, const TwE* tw, const double* border,
                       double* partials, hipStream_t s);
    // the same with psi streamed past the caches (k_x_inv_dens NTS)
    void (*x_inv_dens_stream)(cx<T>* psi, cx<T>* rhohat, int nlines, int hpitch, const TwE* tw,
                              const double* border, double* partials, hipStream_t s);
    void (*x_dens)(const cx<T>* psi, cx<T>* rhohat, int nlines, int hpitch, const TwE* tw, double* partials,
                   hipStream_t s);
    void (*x_r2c_real)(const double* src, cx<T>* rhohat, int nlines, int hpitch, const TwE* tw, hipStream_t s);
    void (*x_c2r_real)(const cx<T>* h, double* dst, int nlines, int hpitch, const TwE* tw, hipStream_t s);
    void (*x_fft)(int dir, cx<T>* psi, int nlines, const TwE* tw, hipStream_t s);
    // adjoint step x passes (k_adj_xs, k_adj_xg); adj_xs writes kAdjParams
    // partials per workgroup, x_blocks of them
    void (*adj_xs)(cx<T>* g, const cx<T>* psi_n, cx<T>* shat, int nlines, int hpitch, const TwE* tw,
                   const AdjXArgs& a, double* partials, hipStream_t s);
    void (*adj_xg)(bool fwd, cx<T>* g, const cx<T>* psi_n, const cx<T>* uhat, int nlines, int hpitch, const TwE* tw,
                   const AdjXArgs& a, hipStream_t s);
    // read-only reductions (k_x_red: RS_COUNT partials per launched wave;
    // k_col_red: one per launched wave)
    void (*x_red)(const cx<T>* psi, const cx<T>* phihat, int nlines, int hpitch, const TwE* tw, const KickArgs& ka,
                  double kscale, double* partials, hipStream_t s);
    void (*col_red)(const cx<T>* data, int ncols, int nother, unsigned sline, unsigned sother, const TwE* tw,
                    double kscale, double* partials, hipStream_t s);
    // batched passes over nmem members of this shape (the step's passes only:
    // the x kick with FFT, the column modes of the step and the inverse x
    // pass with density); the members' arguments are device arrays
    void (*col_b)(int mode, int kf, const ColMember<T>* mem, int nmem, int ncols, int nother, unsigned sline,
                  unsigned sother, const TwE* tw, hipStream_t s);
    void (*x_kick_b)(int vk, const XMember<T>* mem, int nmem, int nlines, int hpitch, const TwE* tw, hipStream_t s);
    void (*x_inv_dens_b)(const XMember<T>* mem, int nmem, int nlines, int hpitch, const TwE* tw, hipStream_t s);
};

// Grid sizes (in workgroups of NW waves).
template <int NW, class F>
inline unsigned x_blocks(int nlines) {
    const int waves = (nlines + F::LPW - 1) / F::LPW;
    return unsigned((waves + NW - 1) / NW);
}
template <int NW, class F>
inline unsigned col_blocks(int ncols, int nother) {
    const long waves = long((ncols + F::LPW - 1) / F::LPW) * nother;
    return unsigned((waves + NW - 1) / NW);
}

template <class T, class F, int NW = kNW<T>>
struct Launch {
    typedef typename TwEntry<T>::type TwE;
    typedef typename std::conditional<std::is_same<T, float>::value, PsiPolicyF32, Plain>::type PsiP;

    // staged variant: one workgroup per CU walking tpw tiles per wave
    template <class PP, int MODE, int KF>
    static bool col_staged(cx<T>* d, int ncols, int nother, unsigned sl, unsigned so, const TwE* tw, KTabs kt,
                           hipStream_t s) {
        typedef Stage<F, T> S;
        // 4 waves per workgroup: 4 staging buffers + transpose buffers fit
        // the 160 KB of LDS of one CU
        constexpr int SNW = 4;
        if constexpr (!WDM_COL_STAGED || !S::OK || MODE == COL_FMIS || MODE == COL_FMIA ||
                      !std::is_same<T, float>::value) {
            return false;
        } else {
            // a lane copies E adjacent columns: the last one read must stay
            // inside the row (pitch = elements between rows of columns)
            const unsigned pitch = sl < so ? sl : so;
            if (unsigned((ncols + S::E - 1) / S::E * S::E) > pitch) return false;
            const long tiles = long((ncols + F::LPW - 1) / F::LPW) * nother;
            const long waves = long(device_cus()) * SNW;
            const int tpw = int((tiles + waves - 1) / waves);
            const unsigned blocks = unsigned((tiles + long(SNW) * tpw - 1) / (long(SNW) * tpw));
            WDM_LAUNCH((k_cols<T, PP, F, MODE, KF, SNW>), dim3(blocks), dim3(SNW * 64), s, d, ncols, nother, sl, so,
                       tw, kt, tpw);
            return true;
        }
    }
    template <int MODE, int KF>
    static void col_t(cx<T>* d, int ncols, int nother, unsigned sl, unsigned so, const TwE* tw, KTabs kt,
                      hipStream_t s) {
        if (col_staged<PsiP, MODE, KF>(d, ncols, nother, sl, so, tw, kt, s)) return;
        if constexpr (WDM_COL_TPW > 1 && MODE != COL_FMIS && MODE != COL_FMIA) {
            WDM_LAUNCH((k_colp<T, PsiP, F, MODE, KF, NW, WDM_COL_TPW>),
                       dim3(col_blocks<NW * WDM_COL_TPW, F>(ncols, nother)), dim3(NW * 64), s, d, ncols, nother, sl,
                       so, tw, kt);
        } else {
            constexpr int W = fmi_group<MODE, KF, col_min_waves<T, F, MODE, KF>(), NW>();
            WDM_LAUNCH((k_col<T, PsiP, F, MODE, KF, W>), dim3(col_blocks<W, F>(ncols, nother)), dim3(W * 64), s,
                       d, ncols, nother, sl, so, tw, kt);
        }
    }
    typedef PoissonPolicy PoisP;
    template <int MODE, int KF>
    static void colp_t(cx<T>* d, int ncols, int nother, unsigned sl, unsigned so, const TwE* tw, KTabs kt,
                       hipStream_t s) {
        if (col_staged<PoisP, MODE, KF>(d, ncols, nother, sl, so, tw, kt, s)) return;
        if constexpr (WDM_COL_TPW > 1) {
            WDM_LAUNCH((k_colp<T, PoisP, F, MODE, KF, NW, WDM_COL_TPW>),
                       dim3(col_blocks<NW * WDM_COL_TPW, F>(ncols, nother)), dim3(NW * 64), s, d, ncols, nother, sl,
                       so, tw, kt);
        } else {
            constexpr int W = fmi_group<MODE, KF, col_min_waves<T, F, MODE, KF>(), NW>();
            WDM_LAUNCH((k_col<T, PoisP, F, MODE, KF, W>), dim3(col_blocks<W, F>(ncols, nother)), dim3(W * 64), s,
                       d, ncols, nother, sl, so, tw, kt);
        }
    }
    // mode/kf combinations used by the plan; KF_POISSON passes run in Plain.
    static void col(int mode, int kf, cx<T>* d, int ncols, int nother, unsigned sl, unsigned so, const TwE* tw,
                    KTabs kt, hipStream_t s) {
        if (kf == KF_POISSON) {
            colp_t<COL_FMI, KF_POISSON>(d, ncols, nother, sl, so, tw, kt, s);
            return;
        }
        if (kf == KF_PNONE) {  // Poisson-chain y passes
            if (mode == COL_FWD) colp_t<COL_FWD, KF_PNONE>(d, ncols, nother, sl, so, tw, kt, s);
            else colp_t<COL_INV, KF_PNONE>(d, ncols, nother, sl, so, tw, kt, s);
            return;
        }
        switch (mode) {
            case COL_FWD: col_t<COL_FWD, KF_NONE>(d, ncols, nother, sl, so, tw, kt, s); break;
            case COL_INV: col_t<COL_INV, KF_NONE>(d, ncols, nother, sl, so, tw, kt, s); break;
            case COL_FMI: col_t<COL_FMI, KF_DRIFT>(d, ncols, nother, sl, so, tw, kt, s); break;
            case COL_FMIS: col_t<COL_FMIS, KF_DRIFT>(d, ncols, nother, sl, so, tw, kt, s); break;
            case COL_FMIA: col_t<COL_FMIA, KF_DRIFT>(d, ncols, nother, sl, so, tw, kt, s); break;
            case COL_MI: col_t<COL_MI, KF_DRIFT>(d, ncols, nother, sl, so, tw, kt, s); break;
        }
    }
    template <int VK>
    static void x_kick_v(bool fwd, cx<T>* psi, const cx<T>* ph, int nlines, int hp, const TwE* tw, const KickArgs& ka,
                         hipStream_t s) {
        if (fwd)
            WDM_LAUNCH((k_x_kick<T, PsiP, F, true, NW, VK>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, psi, ph,
                       nlines, hp, tw, ka);
        else
            WDM_LAUNCH((k_x_kick<T, PsiP, F, false, NW, VK>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, psi,
                       ph, nlines, hp, tw, ka);
    }
    static void x_kick(bool fwd, cx<T>* psi, const cx<T>* ph, int nlines, int hp, const TwE* tw, const KickArgs& ka,
                       hipStream_t s) {
        switch (vkind_of(ka.vm.flags)) {
            case VK_NONE: x_kick_v<VK_NONE>(fwd, psi, ph, nlines, hp, tw, ka, s); break;
            case VK_QUAD: x_kick_v<VK_QUAD>(fwd, psi, ph, nlines, hp, tw, ka, s); break;
            case VK_HERN: x_kick_v<VK_HERN>(fwd, psi, ph, nlines, hp, tw, ka, s); break;
            default: x_kick_v<VK_ANY>(fwd, psi, ph, nlines, hp, tw, ka, s); break;
        }
    }
    static void x_inv_dens(cx<T>* psi, cx<T>* rh, int nlines, int hp, const TwE* tw, const double* border,
                           double* partials, hipStream_t s) {
        WDM_LAUNCH((k_x_inv_dens<T, PsiP, F, NW>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, psi, rh, nlines,
                   hp, tw, border, partials);
    }
    static void x_inv_dens_stream(cx<T>* psi, cx<T>* rh, int nlines, int hp, const TwE* tw, const double* border,
                                  double* partials, hipStream_t s) {
        WDM_LAUNCH((k_x_inv_dens<T, PsiP, F, NW, true>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, psi, rh,
                   nlines, hp, tw, border, partials);
    }
    static void x_dens(const cx<T>* psi, cx<T>* rh, int nlines, int hp, const TwE* tw, double* partials,
                       hipStream_t s) {
        WDM_LAUNCH((k_x_dens<T, F, NW>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, psi, rh, nlines, hp, tw,
                   partials);
    }
    static void x_r2c_real(const double* src, cx<T>* rh, int nlines, int hp, const TwE* tw, hipStream_t s) {
        WDM_LAUNCH((k_x_r2c_real<T, F, NW>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, src, rh, nlines, hp,
                   tw);
    }
    static void x_c2r_real(const cx<T>* h, double* dst, int nlines, int hp, const TwE* tw, hipStream_t s) {
        WDM_LAUNCH((k_x_c2r_real<T, F, NW>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, h, dst, nlines, hp,
                   tw);
    }
    static void x_fft(int dir, cx<T>* psi, int nlines, const TwE* tw, hipStream_t s) {
        if (dir < 0)
            WDM_LAUNCH((k_x_fft<T, PsiP, F, -1, NW>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, psi, nlines,
                       tw);
        else
            WDM_LAUNCH((k_x_fft<T, PsiP, F, +1, NW>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, psi, nlines,
                       tw);
    }
    static void adj_xs(cx<T>* g, const cx<T>* pn, cx<T>* sh, int nlines, int hp, const TwE* tw, const AdjXArgs& a,
                       double* partials, hipStream_t s) {
        if (a.vm.flags & V_QUADRUPOLE)
            WDM_LAUNCH((k_adj_xs<T, PsiP, F, NW, true>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, g, pn, sh,
                       nlines, hp, tw, a, partials);
        else
            WDM_LAUNCH((k_adj_xs<T, PsiP, F, NW, false>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, g, pn, sh,
                       nlines, hp, tw, a, partials);
    }
    static void adj_xg(bool fwd, cx<T>* g, const cx<T>* pn, const cx<T>* uh, int nlines, int hp, const TwE* tw,
                       const AdjXArgs& a, hipStream_t s) {
        if (fwd)
            WDM_LAUNCH((k_adj_xg<T, PsiP, F, true, NW>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, g, pn, uh,
                       nlines, hp, tw, a);
        else
            WDM_LAUNCH((k_adj_xg<T, PsiP, F, false, NW>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, g, pn, uh,
                       nlines, hp, tw, a);
    }
    static void x_red(const cx<T>* psi, const cx<T>* ph, int nlines, int hp, const TwE* tw, const KickArgs& ka,
                      double kscale, double* partials, hipStream_t s) {
        WDM_LAUNCH((k_x_red<T, PsiP, F, NW>), dim3(x_blocks<NW, F>(nlines)), dim3(NW * 64), s, psi, ph, nlines, hp, tw,
                   ka, kscale, partials);
    }
    static void col_red(const cx<T>* d, int ncols, int nother, unsigned sl, unsigned so, const TwE* tw, double kscale,
                        double* partials, hipStream_t s) {
        WDM_LAUNCH((k_col_red<T, PsiP, F, NW>), dim3(col_blocks<NW, F>(ncols, nother)), dim3(NW * 64), s, d, ncols,
                   nother, sl, so, tw, kscale, partials);
    }
    // batched launches: the same kernel bodies as k_col / k_x_kick /
    // k_x_inv_dens of the default build, member index in the grid's y
    template <class PP, int MODE, int KF>
    static void colb_t(const ColMember<T>* mem, int nmem, int ncols, int nother, unsigned sl, unsigned so,
                       const TwE* tw, hipStream_t s) {
        constexpr int W = fmi_group<MODE, KF, col_min_waves<T, F, MODE, KF>(), NW>();
        WDM_LAUNCH((k_col_b<T, PP, F, MODE, KF, W>), dim3(col_blocks<W, F>(ncols, nother), unsigned(nmem)),
                   dim3(W * 64), s, mem, ncols, nother, sl, so, tw);
    }
    static void col_b(int mode, int kf, const ColMember<T>* mem, int nmem, int ncols, int nother, unsigned sl,
                      unsigned so, const TwE* tw, hipStream_t s) {
        if (kf == KF_POISSON) {
            colb_t<PoisP, COL_FMI, KF_POISSON>(mem, nmem, ncols, nother, sl, so, tw, s);
        } else if (kf == KF_PNONE) {
            if (mode == COL_FWD) colb_t<PoisP, COL_FWD, KF_PNONE>(mem, nmem, ncols, nother, sl, so, tw, s);
            else colb_t<PoisP, COL_INV, KF_PNONE>(mem, nmem, ncols, nother, sl, so, tw, s);
        } else if (mode == COL_FWD) {
            colb_t<PsiP, COL_FWD, KF_NONE>(mem, nmem, ncols, nother, sl, so, tw, s);
        } else if (mode == COL_INV) {
            colb_t<PsiP, COL_INV, KF_NONE>(mem, nmem, ncols, nother, sl, so, tw, s);
        } else {
            colb_t<PsiP, COL_FMI, KF_DRIFT>(mem, nmem, ncols, nother, sl, so, tw, s);
        }
    }
    template <int VK>
    static void x_kick_bv(const XMember<T>* mem, int nmem, int nlines, int hp, const TwE* tw, hipStream_t s) {
        WDM_LAUNCH((k_x_kick_b<T, PsiP, F, NW, VK>), dim3(x_blocks<NW, F>(nlines), unsigned(nmem)), dim3(NW * 64), s,
                   mem, nlines, hp, tw);
    }
    static void x_kick_b(int vk, const XMember<T>* mem, int nmem, int nlines, int hp, const TwE* tw, hipStream_t s) {
        switch (vk) {
            case VK_NONE: x_kick_bv<VK_NONE>(mem, nmem, nlines, hp, tw, s); break;
            case VK_QUAD: x_kick_bv<VK_QUAD>(mem, nmem, nlines, hp, tw, s); break;
            case VK_HERN: x_kick_bv<VK_HERN>(mem, nmem, nlines, hp, tw, s); break;
            default: x_kick_bv<VK_ANY>(mem, nmem, nlines, hp, tw, s); break;
        }
    }
    static void x_inv_dens_b(const XMember<T>* mem, int nmem, int nlines, int hp, const TwE* tw, hipStream_t s) {
        WDM_LAUNCH((k_x_inv_dens_b<T, PsiP, F, NW>), dim3(x_blocks<NW, F>(nlines), unsigned(nmem)), dim3(NW * 64), s,
                   mem, nlines, hp, tw);
    }
    static const AxisOps<T>* table() {
        static const AxisOps<T> ops = {F::N, F::N1, F::N2, F::LPW, NW, &col, &x_kick, &x_inv_dens,
                                       &x_inv_dens_stream, &x_dens, &x_r2c_real, &x_c2r_real, &x_fft,
                                       &adj_xs, &adj_xg, &x_red, &col_red, &col_b, &x_kick_b, &x_inv_dens_b};
        return &ops;
    }
};

// Supported axis lengths and their (N1, N2) factorisations.
#define WDM_FOR_EACH_SIZE(X) \
    X(1, 1, 1)               \
    X(8, 8, 1)               \
    X(12, 12, 1)             \
    X(16, 16, 1)             \
    X(24, 24, 1)             \
    X(32, 8, 4)              \
    X(48, 12, 4)             \
    X(64, 8, 8)              \
    X(96, 24, 4)             \
    X(128, 16, 8)            \
    X(160, 40, 4)            \
    X(192, 24, 8)            \
    X(256, 32, 8)            \
    X(384, 48, 8)

// Tuning variants per axis length (the autotuner's candidates, wdm_autotune):
//   0  the table's factorisation, default waves per workgroup (kNW)
//   1  the table's factorisation, kNWAlt waves per workgroup
//   2  the alternative factorisation below, default workgroup size
// The alternatives use N2 = 4 instead of 8: twice the values per lane for half
// the lanes per line and twice the lines per wave (wider row segments in the
// strided passes).
#define WDM_FOR_EACH_ALT(X) \
    X(64, 16, 4)            \
    X(128, 32, 4)           \
    X(192, 48, 4)           \
    X(256, 64, 4)

// the other workgroup size tried by the autotuner
template <class T>
constexpr int kNWAlt = sizeof(T) == 4 ? 8 : 4;
constexpr int kMaxVariants = 3;

#define WDM_DECLARE_SIZE(N, A, B)                        \
    const AxisOps<float>* axis_ops_f32_##N(int variant); \
    const AxisOps<double>* axis_ops_f64_##N(int variant);
WDM_FOR_EACH_SIZE(WDM_DECLARE_SIZE)

template <int N>
struct AltFac {
    typedef void type;  // no alternative factorisation
};
#define WDM_ALT_SPEC(N, A, B)   \
    template <>                 \
    struct AltFac<N> {          \
        typedef Fac<A, B> type; \
    };
WDM_FOR_EACH_ALT(WDM_ALT_SPEC)

template <class T, class F, int N>
const AxisOps<T>* axis_variant(int variant) {
    if (variant == 0) return Launch<T, F>::table();
    if (variant == 1) return Launch<T, F, kNWAlt<T>>::table();
    if (variant == 2) {
        typedef typename AltFac<N>::type A;
        if constexpr (std::is_same<A, void>::value) return nullptr;
        else return Launch<T, A>::table();
    }
    return nullptr;
}

#define WDM_DEFINE_SIZE(N, A, B)                                                                  \
    const AxisOps<float>* axis_ops_f32_##N(int v) { return axis_variant<float, Fac<A, B>, N>(v); } \
    const AxisOps<double>* axis_ops_f64_##N(int v) { return axis_variant<double, Fac<A, B>, N>(v); }

// The pass table of axis length n in variant v (nullptr: unsupported length
// or no such variant).
inline const AxisOps<float>* axis_ops(int n, float*, int v = 0) {
#define WDM_CASE_F32(N, A, B) \
    case N: return axis_ops_f32_##N(v);
    switch (n) { WDM_FOR_EACH_SIZE(WDM_CASE_F32) default: return nullptr; }
#undef WDM_CASE_F32
}
inline const AxisOps<double>* axis_ops(int n, double*, int v = 0) {
#define WDM_CASE_F64(N, A, B) \
    case N: return axis_ops_f64_##N(v);
    switch (n) { WDM_FOR_EACH_SIZE(WDM_CASE_F64) default: return nullptr; }
#undef WDM_CASE_F64
}

}  // namespace wdm
