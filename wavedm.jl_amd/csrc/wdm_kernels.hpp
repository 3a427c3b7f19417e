// Pass kernels of the fused SSFM step.  Arrays are x-fastest (Julia column
// order): psi[x + Nx*(y + Ny*z)]; the Poisson half spectrum is
// rho_hat[kx + Hp*(y + Ny*z)] with kx in [0, Nx/2] and row pitch Hp.
//
// Steady-state step (reference split, src/simulation_main.jl:536-551):
//   k_x_kick      x: Phi = C2R_x(Phi_hat) ; psi *= exp(-i (m/hbar) dt_kick (Phi+V)) ; FFT_x
//   k_col FWD     y: FFT_y
//   k_col FMI     z: FFT_z * drift(k) / N * IFFT_z
//   k_col INV     y: IFFT_y
//   k_x_inv_dens  x: IFFT_x ; * border ; store psi ; rho=|psi|^2 ; R2C_x -> rho_hat
//   k_col FWD     y, k_col FMI z (* -4 pi G / k^2 / N), k_col INV y  on rho_hat
// so the next step's kick finds Phi_hat ready.
#pragma once

#include "wdm_lines.hpp"
#include "wdm_tidal.hpp"

namespace wdm {

// Loads of lanes whose line is out of range: guarded (predicated, zero) or
// unconditional from the clamped in-range address (never stored).
#ifndef WDM_GUARDED_LOADS
#define WDM_GUARDED_LOADS 1
#endif
#if WDM_GUARDED_LOADS
#define WDM_LDG(OK, BASE, IDX) ((OK) ? ldg(BASE, IDX) : zero)
#else
#define WDM_LDG(OK, BASE, IDX) ldg(BASE, IDX)
#endif

// Arithmetic policy of the Poisson chain in FP32 plans.
#ifdef WDM_POISSON_COMP
typedef Comp PoissonPolicy;
#else
typedef Plain PoissonPolicy;
#endif

// Arithmetic policy of the psi passes in FP32 plans: compensated twiddles
// (hi/lo pairs) by default; WDM_PSI_PLAIN builds single FP32 twiddles.
#ifdef WDM_PSI_PLAIN
typedef Plain PsiPolicyF32;
#else
typedef Comp PsiPolicyF32;
#endif

// COL_FMIS: COL_FMI that also stores the forward spectrum to KTabs::spec_out
// COL_FMIA: COL_FMI that also sums the adjoint's dt-scale spectral term
//           sum_k w_k Im(conj(h_k) Psi_k), h = the factored spectrum,
//           Psi = KTabs::spec_out (read), w_k = KTabs::wc * k^2, one FP64 sum
//           per wave into KTabs::partials
enum ColMode { COL_FWD = 0, COL_INV = 1, COL_FMI = 2, COL_MI = 3, COL_FMIS = 4, COL_FMIA = 5 };
// KF_PNONE: no factor, Plain arithmetic (the y passes of the Poisson chain)
enum KFactor { KF_NONE = 0, KF_DRIFT = 1, KF_POISSON = 2, KF_PNONE = 3 };

// k-space factor tables for the line, column and "other" axes of a pass.
//   KF_DRIFT:   complex FP64 exp(-i (hbar/m) k_a^2 dt / 2) per axis,
//               multiplied together and by `scale` (= 1/N).  The column x other
//               product is formed once per lane; each element's full factor is
//               formed in FP64 and, in FP32 plans, applied to the widened data
//               in FP64 with one rounding per component (mul_dyn).  (Two
//               compensated FP32 multiplies, by the line factor and then by the
//               lane factor, bias |psi|^2 by ~1e-10 per step where cos(k^2 dt)
//               rounds to 1.0f.)
//   KF_POISSON: k_a^2 per axis (FP64, or FP32 in FP32 plans); factor =
//               scale / (kc^2+ko^2+kl^2), 0 at k = 0 (scale = -4 pi G / N).
struct KTabs {
    const void* col;
    const void* oth;
    const void* line;
    double scale;
    void* spec_out;  // COL_FMIS: the forward spectrum (before the factor) is
                     // also stored here, same layout as data; COL_FMIA: the
                     // spectrum Psi it reads, same layout
    const double* norm;  // KF_DRIFT: device scalar multiplying the factor (mass projection) or null
    // COL_FMIA: k^2 of the column, other and line axes (FP64), the weight
    // w_k = wc * (kc^2 + ko^2 + kl^2), and one partial per wave of the pass
    // (wave w = workgroup * waves per group + wave in group, for the
    // ceil(ncols / LPW) * nother waves that own lines)
    const double* k2[3];
    double wc;
    double* partials;
};

// line-axis table entry of a factor kind
template <int KF, class T>
struct LineEntry {
    typedef char type;  // no table
};
template <class T>
struct LineEntry<KF_DRIFT, T> {
    typedef cx<double> type;
};
template <class T>
struct LineEntry<KF_POISSON, T> {
    typedef T type;
};

// the per-lane part of a factor (column and other axes)
template <int KF, class T>
struct KfLane {
    int unused;
};
template <class T>
struct KfLane<KF_DRIFT, T> {
    double ar, ai;  // col * oth * scale
};
template <class T>
struct KfLane<KF_POISSON, T> {
    T kco;  // kc^2 + ko^2
    T scale;
};

template <int KF, class T>
WDM_FI KfLane<KF, T> kf_lane(const KTabs& kt, int c, int o) {
    KfLane<KF, T> l;
    if constexpr (KF == KF_DRIFT) {
        const cx<double> a = static_cast<const cx<double>*>(kt.col)[c];
        const cx<double> b = static_cast<const cx<double>*>(kt.oth)[o];
        const double sc = kt.norm ? kt.scale * kt.norm[0] : kt.scale;
        // explicit fmas (here and in apply_kf): every kernel that applies the
        // factor (COL_FMI, COL_FMIS, COL_MI, batched or not) rounds it the
        // same way, whatever the compiler would contract
        l.ar = fmaT(a.x, b.x, -(a.y * b.y)) * sc;
        l.ai = fmaT(a.x, b.y, a.y * b.x) * sc;
    } else if constexpr (KF == KF_POISSON) {
        l.kco = static_cast<const T*>(kt.col)[c] + static_cast<const T*>(kt.oth)[o];
        l.scale = T(kt.scale);
    } else {
        (void)kt; (void)c; (void)o;
    }
    return l;
}

template <int KF, class P, class T>
WDM_FI cx<T> apply_kf(cx<T> v, const KfLane<KF, T>& l, const KTabs& kt, int k) {
    if constexpr (KF == KF_DRIFT) {
        const cx<double> c = static_cast<const cx<double>*>(kt.line)[k];
        return mul_dyn<P>(v, fmaT(l.ar, c.x, -(l.ai * c.y)), fmaT(l.ar, c.y, l.ai * c.x));
    } else if constexpr (KF == KF_POISSON) {
        const T s = l.kco + static_cast<const T*>(kt.line)[k];
        if constexpr (std::is_same<T, double>::value) {
            const double g = s > 0.0 ? l.scale / s : 0.0;
            return v * g;
        } else {
            const float g = s > 0.0f ? l.scale * rcpf(s) : 0.0f;
            return v * g;
        }
    } else {
        (void)l; (void)kt; (void)k;
        return v;
    }
}

// ------------------------------------------------------------------ potentials
// External potential models evaluated in-kernel (replacing the host V(x,y,z,psi)
// callback of src/solver/KDK.jl:64).  All parameters FP64.
enum VFlags { V_STATIC = 1, V_HERNQUIST = 2, V_QUADRUPOLE = 4, V_KAPPA = 8, V_TIDAL = 16 };
struct VModel {
    int flags;
    const double* vstatic;   // full grid, x-fastest (V_STATIC)
    double hq_GM, hq_a, hq_c[3];  // V = -GM / (|x - c| + a)
    double hq_G;             // G alone (dV/dM of the adjoint)
    double qT[6];            // Txx Tyy Tzz Txy Txz Tyz at the kick time; V = 1/2 (x-c)^T T (x-c)
    double q_c[3];
    double kappa;            // V = kappa |psi|^2 (GNLSE term)
    int gnx, gny;            // grid x and y lengths (cell indices of the tidal ramps)
    TidalModel tidal;        // MW tidal potential of the kick (V_TIDAL)
};

WDM_HDFI double tidal_at(const VModel& vm, double x, double y, double z, unsigned idx) {
    const int i = int(idx % unsigned(vm.gnx)), j = int((idx / unsigned(vm.gnx)) % unsigned(vm.gny));
    const int k = int(idx / (unsigned(vm.gnx) * unsigned(vm.gny)));
    return tidal_V(vm.tidal, x, y, z, i, j, k);
}

template <class T>
WDM_FI double eval_V(const VModel& vm, double x, double y, double z, unsigned idx, cx<T> psi) {
    double V = 0.0;
    if (vm.flags & V_STATIC) V += vm.vstatic[idx];
    if (vm.flags & V_HERNQUIST) {
        const double dx = x - vm.hq_c[0], dy = y - vm.hq_c[1], dz = z - vm.hq_c[2];
        V -= vm.hq_GM / (__builtin_sqrt(dx * dx + dy * dy + dz * dz) + vm.hq_a);
    }
    if (vm.flags & V_QUADRUPOLE) {
        const double dx = x - vm.q_c[0], dy = y - vm.q_c[1], dz = z - vm.q_c[2];
        V += 0.5 * (vm.qT[0] * dx * dx + vm.qT[1] * dy * dy + vm.qT[2] * dz * dz) +
             vm.qT[3] * dx * dy + vm.qT[4] * dx * dz + vm.qT[5] * dy * dz;
    }
    if (vm.flags & V_KAPPA) V += vm.kappa * double(norm2(psi));
    if (vm.flags & V_TIDAL) V += tidal_at(vm, x, y, z, idx);
    return V;
}

// V along one x line (fixed y, z): the quadrupole is a quadratic in
// dx = x - c_x with per-line coefficients, the Hernquist term needs only dx^2
// on top of the line's transverse radius.  Same FP64 formulas as eval_V up to
// the association of the quadrupole sum.
struct LineV {
    double qa, qb, qc;  // quadrupole: (qa*dx + qb)*dx + qc
    double hr2;         // Hernquist: (y - c_y)^2 + (z - c_z)^2
    double y, z;        // the line's coordinates (tidal term)
};
WDM_FI LineV line_V(const VModel& vm, double y, double z) {
    LineV l;
    l.y = y;
    l.z = z;
    const double dy = y - vm.q_c[1], dz = z - vm.q_c[2];
    l.qa = 0.5 * vm.qT[0];
    l.qb = vm.qT[3] * dy + vm.qT[4] * dz;
    l.qc = 0.5 * (vm.qT[1] * dy * dy + vm.qT[2] * dz * dz) + vm.qT[5] * dy * dz;
    const double hy = y - vm.hq_c[1], hz = z - vm.hq_c[2];
    l.hr2 = hy * hy + hz * hz;
    return l;
}
template <class T>
WDM_FI double eval_V_line(const VModel& vm, const LineV& l, double x, unsigned idx, cx<T> psi) {
    double V = 0.0;
    if (vm.flags & V_STATIC) V += vm.vstatic[idx];
    if (vm.flags & V_HERNQUIST) {
        const double dx = x - vm.hq_c[0];
        V -= vm.hq_GM / (__builtin_sqrt(fmaT(dx, dx, l.hr2)) + vm.hq_a);
    }
    if (vm.flags & V_QUADRUPOLE) {
        const double dx = x - vm.q_c[0];
        V += fmaT(fmaT(l.qa, dx, l.qb), dx, l.qc);
    }
    if (vm.flags & V_KAPPA) V += vm.kappa * double(norm2(psi));
    if (vm.flags & V_TIDAL) V += tidal_at(vm, x, l.y, l.z, idx);
    return V;
}

// ------------------------------------------------------------------ phases
// Reduce theta to r in [-pi/4, pi/4] and quadrant q (theta = q*pi/2 + r), FP64:
// Cody and Waite's reduction with pi/2 split into two constants, each
// subtracted by a fused multiply-add so n*hi and n*lo enter unrounded.
WDM_FI double reduce_quadrant(double theta, int& q) {
    constexpr double kTwoOverPi = 0.63661977236758134307553505349006;
    constexpr double kPio2Hi = 1.5707963267948965579989817342721;
    constexpr double kPio2Lo = 6.1232339957367660359076653e-17;
    const double n = __builtin_rint(theta * kTwoOverPi);
    double r = fmaT(-n, kPio2Hi, theta);
    r = fmaT(-n, kPio2Lo, r);
    // n mod 4 in FP64 (exact for any integer-valued n below 2^53), then a
    // 32-bit conversion of 0..3 (no 64-bit integer conversion)
    q = int(fmaT(-4.0, __builtin_floor(n * 0.25), n));
    return r;
}

// sin and cos on |r| <= pi/4 by Taylor polynomials in r^2 (enough terms for the
// working precision of T).
WDM_FI void sincos_small(double r, double& s, double& c) {
    const double r2 = r * r;
    // sin: r * sum_{k=0..8} (-1)^k r^(2k)/(2k+1)!
    double ps = -1.0 / 355687428096000.0;      // -1/17!
    ps = fmaT(ps, r2, 1.0 / 1307674368000.0);  // 1/15!
    ps = fmaT(ps, r2, -1.0 / 6227020800.0);    // -1/13!
    ps = fmaT(ps, r2, 1.0 / 39916800.0);       // 1/11!
    ps = fmaT(ps, r2, -1.0 / 362880.0);        // -1/9!
    ps = fmaT(ps, r2, 1.0 / 5040.0);           // 1/7!
    ps = fmaT(ps, r2, -1.0 / 120.0);           // -1/5!
    ps = fmaT(ps, r2, 1.0 / 6.0);              // 1/3!
    s = fmaT(-ps * r2, r, r);
    // cos: sum_{k=0..9} (-1)^k r^(2k)/(2k)!
    double pc = 1.0 / 6402373705728000.0;      // 1/18!
    pc = fmaT(pc, r2, -1.0 / 20922789888000.0);  // -1/16!
    pc = fmaT(pc, r2, 1.0 / 87178291200.0);    // 1/14!
    pc = fmaT(pc, r2, -1.0 / 479001600.0);     // -1/12!
    pc = fmaT(pc, r2, 1.0 / 3628800.0);        // 1/10!
    pc = fmaT(pc, r2, -1.0 / 40320.0);         // -1/8!
    pc = fmaT(pc, r2, 1.0 / 720.0);            // 1/6!
    pc = fmaT(pc, r2, -1.0 / 24.0);            // -1/4!
    pc = fmaT(pc, r2, 0.5);                    // 1/2!
    c = fmaT(-pc, r2, 1.0);
}
WDM_FI void sincos_small(float r, float& s, float& c) {
    const float r2 = r * r;
    float ps = 2.7557319e-6f;            // 1/9!
    ps = fmaT(ps, r2, -1.98412698e-4f);  // -1/7!
    ps = fmaT(ps, r2, 8.33333333e-3f);   // 1/5!
    ps = fmaT(ps, r2, -1.66666667e-1f);  // -1/3!
    s = fmaT(ps * r2, r, r);
    float pc = -2.75573192e-7f;          // -1/10!
    pc = fmaT(pc, r2, 2.48015873e-5f);   // 1/8!
    pc = fmaT(pc, r2, -1.38888889e-3f);  // -1/6!
    pc = fmaT(pc, r2, 4.16666667e-2f);   // 1/4!
    pc = fmaT(pc, r2, -0.5f);            // -1/2!
    c = fmaT(pc, r2, 1.0f);
}

// multiply by (-i)^q
template <class T>
WDM_FI cx<T> rot_quadrant(cx<T> a, int q) {
    const cx<T> b = (q & 1) ? mk(a.y, -a.x) : a;
    return (q & 2) ? -b : b;
}

// a * exp(-i theta).  FP64: direct product.  FP32: theta reduced in FP64, sin
// and cos in FP32, applied as three shears (x -= t y; y += s x; x -= t y with
// t = tan(phi/2), s = sin(phi)); the shear product has determinant exactly 1
// whatever t and s round to, so the kick adds no systematic bias to |psi|^2.
WDM_FI cx<double> kick_rot(cx<double> a, double theta) {
    int q;
    const double r = reduce_quadrant(theta, q);
    double s, c;
    sincos_small(r, s, c);
    const cx<double> b = mk(fmaT(a.x, c, a.y * s), fmaT(a.y, c, -(a.x * s)));
    return rot_quadrant(b, q);
}
WDM_FI cx<float> kick_rot(cx<float> a, double theta) {
#ifdef WDM_KICK_F64
    {
        int q;
        const double r = reduce_quadrant(theta, q);
        double s, c;
        sincos_small(r, s, c);
        return rot_quadrant(mul_dyn<Comp>(a, c, -s), q);
    }
#endif
    int q;
    const float r = float(reduce_quadrant(theta, q));
    float s, c;
    sincos_small(r, s, c);
    // rotation by phi = -r; t = tan(phi/2) = -s / (1 + c), correctly rounded
    // (with the hardware reciprocal instead, the 384^3 FP32 run's mass drift
    // over 1000 steps grew from 7.6e-8 to 1.0e-7 for a 0.5% faster kick)
    const float sp = -s;
    const float tp = -s / (1.0f + c);
    float x = a.x, y = a.y;
    x = fmaT(-tp, y, x);
    y = fmaT(sp, x, y);
    x = fmaT(-tp, y, x);
    return rot_quadrant(mk(x, y), q);
}

// multiply by a real FP64 factor in data precision (FP32 compensated)
template <class P>
WDM_FI cx<double> scale_dyn(cx<double> a, double f) { return a * f; }
template <class P>
WDM_FI cx<float> scale_dyn(cx<float> a, double f) {
    if constexpr (std::is_same<P, Plain>::value) {
        return a * float(f);
    } else {
        const float h = float(f), l = float(f - double(h));
        return mk(fmaT(a.x, h, a.x * l), fmaT(a.y, h, a.y * l));
    }
}

// ------------------------------------------------------------------ wave helpers
// Copies a small table (twiddles, k-space factors) into LDS once per
// workgroup and returns the LDS copy.  The line FFTs index these tables per
// lane; from LDS the lookups cost a short LDS round trip instead of an L2 one
// each, which the compiler would otherwise wait for one by one.
template <int N, class E>
WDM_FI const E* stage_table(E* dst, const E* __restrict__ src) {
    for (int i = int(threadIdx.x); i < N; i += int(blockDim.x)) dst[i] = src[i];
    __syncthreads();
    return dst;
}

// The twiddles in the conflict-free layout (TL = 1 of fwd2/inv2): 2N
// entries gathered from the plain table.
template <class F, class E>
WDM_FI const E* stage_tw2(E* dst, const E* __restrict__ src) {
    for (int i = int(threadIdx.x); i < 2 * F::N; i += int(blockDim.x)) {
        const int f = i < F::N ? i : i - F::N;
        const int e = i < F::N ? (f % F::N2) * (f / F::N2) : (f % F::N1) * (f / F::N1);
        dst[i] = src[e];
    }
    __syncthreads();
    return dst;
}

struct LaneMap {
    int lane, wv, j, lw, w;
};

// Workgroup id with the XCD round robin undone: the hardware hands workgroup b
// to XCD b mod 8, so consecutive ids land on different L2s.  Renumbering
// b -> (b mod 8) q + b / 8 (q = gridDim / 8) gives each XCD a contiguous
// range of logical ids, so neighbouring column groups share one L2.  The tail
// beyond 8q keeps its id.
#ifndef WDM_XCD_REMAP
#define WDM_XCD_REMAP 1
#endif
WDM_FI unsigned xcd_block() {
#if WDM_XCD_REMAP
    const unsigned b = blockIdx.x, q = gridDim.x / 8u;
    if (b >= q * 8u) return b;
    return (b % 8u) * q + b / 8u;
#else
    return blockIdx.x;
#endif
}
// b: the workgroup's index within its pass (blockIdx.x, or the member-local
// index of a batched launch)
template <class F, int NW>
WDM_FI LaneMap lane_map(unsigned b) {
    LaneMap m;
    m.lane = int(threadIdx.x) & (kWave - 1);
    m.wv = int(threadIdx.x) / kWave;
    m.j = m.lane % F::N2;
    m.lw = m.lane / F::N2;
    m.w = int(b) * NW + m.wv;
    return m;
}
template <class F, int NW>
WDM_FI LaneMap lane_map() {
    return lane_map<F, NW>(blockIdx.x);
}

// Batched launches (an ensemble's members of one shape): grid (g, nmembers),
// the workgroup of linear id L = y g + x.  The XCD renumbering of xcd_block()
// is applied to L over the whole grid, then split into (member, index within
// the member's pass of g workgroups).
WDM_FI void batch_block(unsigned& member, unsigned& b) {
    const unsigned g = gridDim.x, total = gridDim.x * gridDim.y;
    unsigned L = blockIdx.y * g + blockIdx.x;
#if WDM_XCD_REMAP
    const unsigned q = total / 8u;
    if (L < q * 8u) L = (L % 8u) * q + L / 8u;
#endif
    member = L / g;
    b = L % g;
}

// Deterministic wave sum/max of per-lane FP64 values through the wave's LDS
// buffer; lane 0 returns the result.
template <class T>
WDM_FI void wave_sum_max(cx<T>* wbuf, int lane, double vsum, double vmax, double& osum, double& omax) {
    double* d = reinterpret_cast<double*>(wbuf);
    d[lane] = vsum;
    d[kWave + lane] = vmax;
    WDM_WAVE_SYNC();
    osum = 0.0;
    omax = 0.0;
    if (lane == 0) {
        for (int l = 0; l < kWave; ++l) {
            osum += d[l];
            omax = d[kWave + l] > omax ? d[kWave + l] : omax;
        }
    }
    WDM_WAVE_SYNC();
}

// ------------------------------------------------------------------ LDS-staged column pass
// Same work as k_col.  Each wave owns a staging buffer in LDS for one tile
// (LPW columns x N rows) and walks `tpw` tiles (stride: all waves of the
// grid).  The tile is fetched with direct global->LDS copies (no registers),
// read into registers, and the next tile's copies are issued at once, so the
// HBM reads of tile k+1 overlap the transforms and stores of tile k within the
// same wave.  A copy instruction moves 64 lanes x 16 bytes: RPI rows of LPW
// columns, LDS layout [row][column].
template <class F, class T>
struct Stage {
    static constexpr int E = 16 / int(sizeof(cx<T>));   // elements per lane copy
    static constexpr int LPR = F::LPW / E;               // lanes per row
    static constexpr int RPI = kWave / LPR;              // rows per copy instruction
    static constexpr int NI = F::N / RPI;                // copy instructions per tile
    static constexpr int ELEMS = F::N * F::LPW;          // tile elements
    static constexpr bool OK = (F::LPW % E == 0) && (kWave % LPR == 0) && (F::N % RPI == 0) && F::N2 > 1;
};

template <class T, class P, class F, int MODE, int KF, int NW>
WDM_KERNEL(NW * 64)
void k_cols(cx<T>* __restrict__ data, int ncols, int nother, unsigned sline, unsigned sother,
            const typename TwEntry<T>::type* __restrict__ tw, KTabs kt, int tpw) {
    typedef Stage<F, T> S;
    static_assert(S::OK, "tile shape not stageable");
    static_assert(MODE != COL_FMIS, "the spectrum side output uses k_col");
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, LPW = F::LPW;
    constexpr bool FMI = MODE == COL_FMI;
    constexpr int DIR = (MODE == COL_FWD || FMI) ? -1 : +1;
    static_assert(sizeof(cx<T>) * NW * (F::XBUF + S::ELEMS) <= 160 * 1024, "LDS of one CU exceeded");
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    tw = stage_table<F::N>(tw_s, tw);
    WDM_SHARED cx<T> stage[NW][S::ELEMS];
    const LaneMap m = lane_map<F, NW>();
    const int lane = m.lane;
    const int ntiles = (ncols + LPW - 1) / LPW;
    const int total = ntiles * nother;
    const int stride = int(gridDim.x) * NW;
    const unsigned step = unsigned(N2) * sline;
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    cx<T>* sb = stage[m.wv];
    // this lane's part of every copy instruction: row offset and column
    const int crow = lane / S::LPR, ccol = (lane % S::LPR) * S::E;
    auto fetch = [&](int t) {  // wave-uniform t < total
        const int o = t / ntiles, c0 = (t % ntiles) * LPW;
        const bool cok = c0 + ccol < ncols;  // a lane's E columns are all in or all out (ncols % E == 0)
        const unsigned base = unsigned(o) * sother + unsigned(c0 + ccol) + unsigned(crow) * sline;
        if (cok) {
            WDM_UNROLL
            for (int r = 0; r < S::NI; ++r)
                lds_dma16(data + (base + unsigned(r * S::RPI) * sline), sb + r * (kWave * S::E));
        }
    };
    int t = m.w;
    if (t >= total) return;
    fetch(t);
    for (int k = 0; k < tpw; ++k) {
        const int o = t / ntiles, c = (t % ntiles) * LPW + m.lw;
        const bool ok = c < ncols;
        const int cc = ok ? c : 0;
        const int tn = t + stride;
        const bool more = (k + 1 < tpw) && (tn < total);
        WDM_WAIT_VMEM();  // the tile's copies have landed
        WDM_WAVE_SYNC();
        cx<T> v[N1];
        WDM_UNROLL
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = sb[(N2 * n1 + m.j) * LPW + m.lw];
        WDM_WAIT_LDS();   // every lane has read the buffer before it is refilled
        WDM_WAVE_SYNC();
        if (more) fetch(tn);
        if constexpr (MODE == COL_MI) {
            WDM_UNROLL
            for (int n1 = 0; n1 < N1; ++n1) v[n1] = apply_kf<KF, P, T>(v[n1], kf_lane<KF, T>(kt, cc, o), kt, N2 * n1 + m.j);
        }
        cx<T> X[M][N2];
        fwd2<F, DIR, P>(v, X, m.j, xb, tw);
        if constexpr (FMI) {
            WDM_UNROLL
            for (int i = 0; i < M; ++i) {
                WDM_SCHED_FENCE();
                WDM_UNROLL
                for (int k2 = 0; k2 < N2; ++k2)
                    X[i][k2] = apply_kf<KF, P, T>(X[i][k2], kf_lane<KF, T>(kt, cc, o), kt, (i * N2 + m.j) + N1 * k2);
            }
            inv2<F, +1, P>(X, v, m.j, xb, tw);
            if (ok) {
                const unsigned sbase = unsigned(o) * sother + unsigned(cc) + unsigned(m.j) * sline;
                WDM_UNROLL
                for (int kq = 0; kq < N1; ++kq) stg(data, sbase + unsigned(kq) * step, v[kq]);
            }
        } else {
            if (ok) {
                WDM_UNROLL
                for (int i = 0; i < M; ++i)
                    WDM_UNROLL
                    for (int k2 = 0; k2 < N2; ++k2)
                        stg(data, unsigned(o) * sother + unsigned(cc) + unsigned((i * N2 + m.j) + N1 * k2) * sline,
                            X[i][k2]);
            }
        }
        if (!more) break;
        t = tn;
    }
}

// ------------------------------------------------------------------ pipelined column pass
// Same work as k_col, but each wave walks TPW tiles (stride: all waves of the
// grid) and issues the loads of tile k+1 before the transforms of tile k, so
// HBM reads overlap the FFT arithmetic of the same wave instead of relying on
// other resident waves.  Tiles past the end are idle (no loads, no stores).
template <class T, class P, class F, int MODE, int KF, int NW, int TPW>
WDM_KERNEL(NW * 64)
void k_colp(cx<T>* __restrict__ data, int ncols, int nother, unsigned sline, unsigned sother,
            const typename TwEntry<T>::type* __restrict__ tw, KTabs kt) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M;
    static_assert(MODE != COL_FMIS, "the spectrum side output uses k_col");
    constexpr bool FMI = MODE == COL_FMI;
    constexpr int DIR = (MODE == COL_FWD || FMI) ? -1 : +1;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    tw = stage_table<F::N>(tw_s, tw);
    const LaneMap m = lane_map<F, NW>();
    const int ntiles = (ncols + F::LPW - 1) / F::LPW;
    const int total = ntiles * nother;
    const int stride = int(gridDim.x) * NW;
    const unsigned step = unsigned(N2) * sline;
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    const cx<T> zero = mk(T(0), T(0));
    // tile t -> (ok, other, column)
    auto geo = [&](int t, bool& ok, int& oc, int& cc) {
        const int o = t / ntiles, c = (t % ntiles) * F::LPW + m.lw;
        ok = (t < total) && (c < ncols);
        oc = ok ? o : 0;
        cc = ok ? c : 0;
    };
    auto load = [&](cx<T>* v, bool ok, int oc, int cc) {
        const unsigned base = unsigned(oc) * sother + unsigned(cc) + unsigned(m.j) * sline;
        WDM_UNROLL
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = WDM_LDG(ok, data, base + unsigned(n1) * step);
    };
    int t = m.w;
    bool ok;
    int oc, cc;
    geo(t, ok, oc, cc);
    cx<T> v[N1];
    load(v, ok, oc, cc);
    for (int k = 0; k < TPW; ++k) {
        if (t >= total) break;  // wave-uniform
        const int tn = t + stride;
        const bool more = (k + 1 < TPW) && (tn < total);
        bool nok = false;
        int noc = 0, ncc = 0;
        cx<T> nv[N1];
        if (more) {
            geo(tn, nok, noc, ncc);
            load(nv, nok, noc, ncc);
        }
        WDM_SCHED_FENCE();
        if constexpr (MODE == COL_MI) {
            WDM_UNROLL
            for (int n1 = 0; n1 < N1; ++n1) v[n1] = apply_kf<KF, P, T>(v[n1], kf_lane<KF, T>(kt, cc, oc), kt, N2 * n1 + m.j);
        }
        cx<T> X[M][N2];
        fwd2<F, DIR, P>(v, X, m.j, xb, tw);
        if constexpr (FMI) {
            WDM_UNROLL
            for (int i = 0; i < M; ++i) {
                WDM_SCHED_FENCE();
                WDM_UNROLL
                for (int k2 = 0; k2 < N2; ++k2)
                    X[i][k2] = apply_kf<KF, P, T>(X[i][k2], kf_lane<KF, T>(kt, cc, oc), kt, (i * N2 + m.j) + N1 * k2);
            }
            inv2<F, +1, P>(X, v, m.j, xb, tw);
            if (ok) {
                const unsigned sbase = unsigned(oc) * sother + unsigned(cc) + unsigned(m.j) * sline;
                WDM_UNROLL
                for (int kq = 0; kq < N1; ++kq) stg(data, sbase + unsigned(kq) * step, v[kq]);
            }
        } else {
            if (ok) {
                WDM_UNROLL
                for (int i = 0; i < M; ++i)
                    WDM_UNROLL
                    for (int k2 = 0; k2 < N2; ++k2)
                        stg(data, unsigned(oc) * sother + unsigned(cc) + unsigned((i * N2 + m.j) + N1 * k2) * sline,
                            X[i][k2]);
            }
        }
        if (!more) break;
        WDM_UNROLL
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = nv[n1];
        ok = nok;
        oc = noc;
        cc = ncc;
        t = tn;
    }
}

// ------------------------------------------------------------------ column pass (y or z)
// Lines along an axis with element stride `sline`; columns are the contiguous
// x index c in [0, ncols); `nother` positions of the remaining axis at stride
// `sother`.  Each wave owns LPW adjacent columns at one `other` position.
// minimum waves per SIMD asked of the compiler for a pass kernel.  The fused
// FFT * drift * IFFT pass in FP32 needs 176 registers for two waves per SIMD;
// capped at 168 for three (a 40-byte spill) it runs 9 us faster at 384^3
// (272 vs 281 us).  The Poisson z pass (182 registers) is slower capped.
#ifndef WDM_MINW_FMI_F32
#define WDM_MINW_FMI_F32 3
#endif
#ifndef WDM_MINW_PFMI_F32
#define WDM_MINW_PFMI_F32 1
#endif
#ifndef WDM_MINW_XINV_F32
#define WDM_MINW_XINV_F32 1
#endif
#ifndef WDM_MINW_COL_F32
#define WDM_MINW_COL_F32 1
#endif
template <class T, class F, int MODE, int KF>
constexpr int col_min_waves() {
    if (!std::is_same<T, float>::value) return 1;
    // the Poisson z pass at 48+ values per lane, with the conflict-free
    // twiddle layout, fits three waves (384^3: 138 -> 131 us); shorter lines
    // keep the uncapped build
    if (MODE == COL_FMI && KF == KF_POISSON && F::N1 >= 48) return 3;
    if (MODE == COL_FMI) return KF == KF_POISSON ? WDM_MINW_PFMI_F32 : WDM_MINW_FMI_F32;
    if (MODE == COL_FWD || MODE == COL_INV) return WDM_MINW_COL_F32;
    return 1;
}

// WDM_COL_EARLY_LOADS: the column passes issue their data loads before
// staging the tables, so the table fetch and the barrier after it overlap the
// data's memory latency instead of preceding it.
#ifndef WDM_COL_EARLY_LOADS
#define WDM_COL_EARLY_LOADS 0
#endif

// One column pass of workgroup `bid` (already XCD-renumbered).
template <class T, class P, class F, int MODE, int KF, int NW>
WDM_FI void col_body(cx<T>* __restrict__ data, int ncols, int nother, unsigned sline, unsigned sother,
                     const typename TwEntry<T>::type* __restrict__ tw, KTabs kt, unsigned bid) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M;
    constexpr bool EARLY = WDM_COL_EARLY_LOADS && MODE != COL_MI;
    // The FP32 fused passes (drift and Poisson) of 48+ values per lane take
    // their twiddles in the conflict-free layout: its offsets are
    // compile-time constants per lane (k1 N2 + j), which frees the registers
    // the products j k1 held in these register-capped passes (z drift at
    // 384^3: 270 -> 257 us, the cap's spill 40 -> 8 bytes; the Poisson z
    // pass fits three waves with it).  Elsewhere the gather that builds the
    // layout costs more than it saves (y passes +9 us).
    constexpr int TL = (std::is_same<T, float>::value && (KF == KF_DRIFT || KF == KF_POISSON) && F::N1 >= 48 &&
                        (MODE == COL_FMI || MODE == COL_FMIS || MODE == COL_FMIA))
                           ? 1
                           : 0;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[TL ? 2 * F::N : F::N];
    const LaneMap m = lane_map<F, NW>(bid);
    const int ntiles = (ncols + F::LPW - 1) / F::LPW;
    const int wl = int(bid) * NW + m.wv;
    const int o = wl / ntiles, tile = wl % ntiles;
    const int c = tile * F::LPW + m.lw;
    const bool ok = (o < nother) && (c < ncols);
    const int oc = ok ? o : 0, cc = ok ? c : 0;
    // 32-bit element offsets from the uniform base pointer: one VGPR per
    // address (scalar-base addressing) and cheap to recompute for the stores
    const unsigned base = unsigned(oc) * sother + unsigned(cc) + unsigned(m.j) * sline;
    const unsigned step = unsigned(N2) * sline;
    const cx<T> zero = mk(T(0), T(0));
    cx<T> v[N1];
    if constexpr (EARLY) {
        WDM_UNROLL
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = WDM_LDG(ok, data, base + unsigned(n1) * step);
    }
    // k-space factor table of the line axis, staged with the twiddles
    typedef typename LineEntry<KF, T>::type LE;
    constexpr int NL = (KF == KF_DRIFT || KF == KF_POISSON) ? F::N : 1;
    WDM_SHARED LE line_s[NL];
    if constexpr (KF == KF_DRIFT || KF == KF_POISSON) {
        const LE* src = static_cast<const LE*>(kt.line);
        for (int i = int(threadIdx.x); i < NL; i += int(blockDim.x)) line_s[i] = src[i];
        kt.line = line_s;
    }
    // COL_FMIA: the line axis's k^2, staged with the tables
    WDM_SHARED double k2l_s[MODE == COL_FMIA ? F::N : 1];
    if constexpr (MODE == COL_FMIA) {
        for (int i = int(threadIdx.x); i < F::N; i += int(blockDim.x)) k2l_s[i] = kt.k2[2][i];
    }
    if constexpr (TL) tw = stage_tw2<F>(tw_s, tw);
    else tw = stage_table<F::N>(tw_s, tw);
    const KfLane<KF, T> kl = kf_lane<KF, T>(kt, cc, oc);
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    cx<T> X[M][N2];
    if constexpr (MODE == COL_MI) {
        WDM_UNROLL
        for (int n1 = 0; n1 < N1; ++n1) {
            const int k = N2 * n1 + m.j;
            v[n1] = apply_kf<KF, P, T>(WDM_LDG(ok, data, base + unsigned(n1) * step), kl, kt, k);
        }
    } else if constexpr (!EARLY) {
        WDM_UNROLL
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = WDM_LDG(ok, data, base + unsigned(n1) * step);
    }
    constexpr bool FMI = MODE == COL_FMI || MODE == COL_FMIS || MODE == COL_FMIA;
    constexpr int DIR = (MODE == COL_FWD || FMI) ? -1 : +1;
    fwd2<F, DIR, P, TL>(v, X, m.j, xb, tw);
    if constexpr (FMI) {
        if (MODE == COL_FMIS && ok) {
            cx<T>* so = static_cast<cx<T>*>(kt.spec_out);
            WDM_UNROLL
            for (int i = 0; i < M; ++i)
                WDM_UNROLL
                for (int k2 = 0; k2 < N2; ++k2)
                    stg(so, unsigned(oc) * sother + unsigned(cc) + unsigned((i * N2 + m.j) + N1 * k2) * sline, X[i][k2]);
        }
        // COL_FMIA: sum_k w_k Im(conj(h_k) Psi_k) over this lane's elements
        // (h = the factored spectrum), in FP64, a chunk of Psi at a time
        double acc = 0.0;
        const double kco = MODE == COL_FMIA ? kt.k2[0][cc] + kt.k2[1][oc] : 0.0;
        WDM_UNROLL
        for (int i = 0; i < M; ++i) {
            WDM_SCHED_FENCE();
            cx<T> q[N2];
            if constexpr (MODE == COL_FMIA) {
                const cx<T>* ps = static_cast<const cx<T>*>(kt.spec_out);
                WDM_UNROLL
                for (int k2 = 0; k2 < N2; ++k2)
                    q[k2] = WDM_LDG(ok, ps, unsigned(oc) * sother + unsigned(cc) +
                                                unsigned((i * N2 + m.j) + N1 * k2) * sline);
            }
            WDM_UNROLL
            for (int k2 = 0; k2 < N2; ++k2) {
                const int k = (i * N2 + m.j) + N1 * k2;
                X[i][k2] = apply_kf<KF, P, T>(X[i][k2], kl, kt, k);
                if constexpr (MODE == COL_FMIA) {
                    const double w = kt.wc * (kco + k2l_s[k]);
                    acc += w * (double(X[i][k2].x) * double(q[k2].y) - double(X[i][k2].y) * double(q[k2].x));
                }
            }
        }
        if constexpr (MODE == COL_FMIA) {
            double wsum, wmax;
            wave_sum_max<T>(&xbuf[m.wv][0], m.lane, ok ? acc : 0.0, 0.0, wsum, wmax);
            if (m.lane == 0 && o < nother) kt.partials[wl] = wsum;
        } else {
            (void)acc;
            (void)kco;
        }
        inv2<F, +1, P, TL>(X, v, m.j, xb, tw);
        if (ok) {
            const unsigned sbase = unsigned(oc) * sother + unsigned(cc) + unsigned(m.j) * sline;
            WDM_UNROLL
            for (int kq = 0; kq < N1; ++kq) stg(data, sbase + unsigned(kq) * step, v[kq]);
        }
    } else {
        if (ok) {
            // X[i][k2] is element (i*N2 + j) + N1*k2 of the line
            WDM_UNROLL
            for (int i = 0; i < M; ++i)
                WDM_UNROLL
                for (int k2 = 0; k2 < N2; ++k2)
                    stg(data, unsigned(oc) * sother + unsigned(cc) + unsigned((i * N2 + m.j) + N1 * k2) * sline, X[i][k2]);
        }
    }
}

template <class T, class P, class F, int MODE, int KF, int NW>
WDM_KERNEL_W(NW * 64, (col_min_waves<T, F, MODE, KF>()))
void k_col(cx<T>* __restrict__ data, int ncols, int nother, unsigned sline, unsigned sother,
           const typename TwEntry<T>::type* __restrict__ tw, KTabs kt) {
    col_body<T, P, F, MODE, KF, NW>(data, ncols, nother, sline, sother, tw, kt, xcd_block());
}

// a member of a batched column pass: its array and factor tables
template <class T>
struct ColMember {
    cx<T>* data;
    KTabs kt;
};
template <class T, class P, class F, int MODE, int KF, int NW>
WDM_KERNEL_W(NW * 64, (col_min_waves<T, F, MODE, KF>()))
void k_col_b(const ColMember<T>* __restrict__ mem, int ncols, int nother, unsigned sline, unsigned sother,
             const typename TwEntry<T>::type* __restrict__ tw) {
    unsigned member, b;
    batch_block(member, b);
    col_body<T, P, F, MODE, KF, NW>(mem[member].data, ncols, nother, sline, sother, tw, mem[member].kt, b);
}

// ------------------------------------------------------------------ x-passes
struct KickArgs {
    double mh_dt;        // (m/hbar) * dt_kick
    double t;            // kick time (for time-dependent models, folded into VModel by the host)
    int use_phi;         // self-gravity on
    int Ny;
    const double* xs;    // coordinates per axis (FP64)
    const double* ys;
    const double* zs;
    double x0, x1;       // x(i) = (1 - t) x0 + t x1, t = i * xinv (the LinRange formula)
    double xinv;
    VModel vm;
    const double* norm;  // pending mass-projection scale (device scalar) or null
};

// Kick (+ optional forward x-FFT).  Reads psi rows in real space and Phi_hat
// (half spectrum after the y/z Poisson passes); writes psi (x-transformed when
// FWD) in place.
// Potential kinds the kick is compiled for: straight-line code for the
// common single-model cases, runtime flags for any combination.
enum VKind { VK_NONE = 0, VK_QUAD = 1, VK_HERN = 2, VK_ANY = 3 };
inline int vkind_of(int flags) {
    if (flags == 0) return VK_NONE;
    if (flags == V_QUADRUPOLE) return VK_QUAD;
    if (flags == V_HERNQUIST) return VK_HERN;
    return VK_ANY;
}

// The inverse-layout input of a C2R line from its half spectrum h[0..N/2]:
// X[i][p] = spectrum element k = i N2 + j + N1 p, which is h[k] for k <= N/2
// and conj(h[N - k]) above.  Each element of h is read from memory once:
// lane j loads D[m] = h[j + N2 m] for the m with j + N2 m <= N/2 (m <= N1/2),
// and the conjugate half comes from lane N2 - j's D[N1 - 1 - m] (lane 0: its
// own D[N1 - m]) through the wave's LDS buffer.  (Reading h[N - k] from
// memory as well reads every element twice; at 384^3 the second reads miss
// L2 often enough to add 8% to the kick pass's HBM traffic.)
template <class F, class T>
WDM_FI void hermitian_row(const cx<T>* __restrict__ h, unsigned hbase, bool ok, const LaneMap& lm, cx<T>* wb,
                          cx<T> (*X)[F::N2]) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, N = F::N;
    const cx<T> zero = mk(T(0), T(0));
    if constexpr (N1 % 2 != 0) {
        // N = 1: the single element
        (void)wb;
        WDM_UNROLL
        for (int i = 0; i < M; ++i)
            WDM_UNROLL
            for (int p = 0; p < N2; ++p) X[i][p] = WDM_LDG(ok, h, hbase + unsigned(i * N2 + lm.j + N1 * p));
    } else {
        constexpr int H = N1 / 2;
        cx<T> D[H + 1];
        WDM_UNROLL
        for (int q = 0; q <= H; ++q) {
            const int k = lm.j + N2 * q;
            D[q] = (ok && 2 * k <= N) ? ldg(h, hbase + unsigned(k)) : zero;
        }
        cx<T> Mv[H];  // Mv[q] = lane (N2 - j)'s D[N1 - 1 - (H + q)] = D[H - 1 - q]
        if constexpr (N2 > 1) {
            constexpr int RB = F::XBUF / kWave;
            constexpr int R = RB < H ? RB : H;
            const int src = (lm.lane & ~(N2 - 1)) | ((N2 - lm.j) & (N2 - 1));
            WDM_UNROLL
            for (int c0 = 0; c0 < H; c0 += R) {
                WDM_UNROLL
                for (int t = 0; t < R; ++t)
                    if (c0 + t < H) wb[lm.lane * R + t] = D[H - 1 - (c0 + t)];
                WDM_WAVE_SYNC();
                WDM_UNROLL
                for (int t = 0; t < R; ++t)
                    if (c0 + t < H) Mv[c0 + t] = wb[src * R + t];
                WDM_WAVE_SYNC();
            }
        } else {
            (void)wb;
        }
        const bool j0 = lm.j == 0;
        WDM_UNROLL
        for (int i = 0; i < M; ++i)
            WDM_UNROLL
            for (int p = 0; p < N2; ++p) {
                const int q = i + M * p;  // element k = j + N2 q
                cx<T> zz;
                if (q < H) {
                    zz = D[q];
                } else if (q == H) {
                    if constexpr (N2 > 1) zz = j0 ? D[H] : mk(Mv[0].x, -Mv[0].y);
                    else zz = D[H];
                } else {
                    const cx<T> own = D[N1 - q];
                    if constexpr (N2 > 1) zz = j0 ? mk(own.x, -own.y) : mk(Mv[q - H].x, -Mv[q - H].y);
                    else zz = mk(own.x, -own.y);
                }
                X[i][p] = zz;
            }
    }
}

template <class T, class P, class F, bool FWD, int NW, int VK>
WDM_FI void x_kick_body(cx<T>* __restrict__ psi, const cx<T>* __restrict__ phihat, int nlines, int hpitch,
                        const typename TwEntry<T>::type* __restrict__ tw, const KickArgs& ka, unsigned bid) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, N = F::N;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    tw = stage_table<F::N>(tw_s, tw);
    const LaneMap m = lane_map<F, NW>(bid);
    const int l = m.w * F::LPW + m.lw;
    const bool ok = l < nlines;
    const int lc = ok ? l : 0;
    const int y = lc % ka.Ny, z = lc / ka.Ny;
    const unsigned rbase = unsigned(lc) * unsigned(N);
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    const cx<T> zero = mk(T(0), T(0));
    cx<T> X[M][N2];
    cx<T> v[N1];
    T phi[N1];
    if (ka.use_phi) {
        const unsigned hbase = unsigned(lc) * unsigned(hpitch);
        hermitian_row<F>(phihat, hbase, ok, m, &xbuf[m.wv][0], X);
        inv2<F, +1, PoissonPolicy>(X, v, m.j, xb, tw);
        WDM_UNROLL
        for (int kq = 0; kq < N1; ++kq) phi[kq] = v[kq].x;
    } else {
        WDM_UNROLL
        for (int kq = 0; kq < N1; ++kq) phi[kq] = T(0);
    }
    const LineV lv = line_V(ka.vm, ka.ys[y], ka.zs[z]);
    // psi loads only after Phi is done (otherwise the 48 loads are hoisted
    // above the C2R transform and both sets of registers are live at once)
    WDM_SCHED_FENCE();
    WDM_UNROLL
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = WDM_LDG(ok, psi, rbase + unsigned(N2 * n1 + m.j));
    // mass projection: this step's z pass scales psi by f, so the kick uses
    // the potential of the projected state, Phi f^2 (linear in |psi|^2)
    const double nf = ka.norm ? ka.norm[0] : 1.0;
    const double nf2 = nf * nf;
    WDM_UNROLL
    for (int n1 = 0; n1 < N1; ++n1) {
        if ((n1 & 3) == 0) WDM_SCHED_FENCE();
        const int x = N2 * n1 + m.j;
        double V = 0.0;
        if constexpr (VK != VK_NONE) {
            const double tx = double(x) * ka.xinv;
            const double xc = (1.0 - tx) * ka.x0 + tx * ka.x1;
            if constexpr (VK == VK_QUAD) {
                const double dx = xc - ka.vm.q_c[0];
                V = fmaT(fmaT(lv.qa, dx, lv.qb), dx, lv.qc);
            } else if constexpr (VK == VK_HERN) {
                const double dx = xc - ka.vm.hq_c[0];
                V = -ka.vm.hq_GM / (__builtin_sqrt(fmaT(dx, dx, lv.hr2)) + ka.vm.hq_a);
            } else {
                if (ka.vm.flags)
                    V = eval_V_line<T>(ka.vm, lv, xc, unsigned(lc) * N + x, ka.norm ? scale_dyn<P>(v[n1], nf) : v[n1]);
            }
        }
        T ph = phi[n1];
        WDM_OPAQUE(ph);
        const double phd = ka.norm ? double(ph) * nf2 : double(ph);
        v[n1] = kick_rot(v[n1], ka.mh_dt * (phd + V));
    }
    if constexpr (FWD) {
        fwd2<F, -1, P>(v, X, m.j, xb, tw);
        if (ok) {
            WDM_UNROLL
            for (int i = 0; i < M; ++i)
                WDM_UNROLL
                for (int k2 = 0; k2 < N2; ++k2) stg(psi, rbase + unsigned((i * N2 + m.j) + N1 * k2), X[i][k2]);
        }
    } else {
        if (ok) {
            WDM_UNROLL
            for (int n1 = 0; n1 < N1; ++n1) stg(psi, rbase + unsigned(N2 * n1 + m.j), v[n1]);
        }
    }
}

template <class T, class P, class F, bool FWD, int NW, int VK>
WDM_KERNEL(NW * 64)
void k_x_kick(cx<T>* __restrict__ psi, const cx<T>* __restrict__ phihat, int nlines, int hpitch,
              const typename TwEntry<T>::type* __restrict__ tw, KickArgs ka) {
    x_kick_body<T, P, F, FWD, NW, VK>(psi, phihat, nlines, hpitch, tw, ka, blockIdx.x);
}

// a member of a batched x pass: its arrays, kick arguments and absorber
template <class T>
struct XMember {
    cx<T>* psi;
    cx<T>* hat;           // Phi_hat in (kick), rho_hat out (inverse pass); null without self-gravity
    double* partials;     // per-wave mass partials of the inverse pass
    const double* border; // length-nx absorber or null
    KickArgs ka;
};
template <class T, class P, class F, int NW, int VK>
WDM_KERNEL(NW * 64)
void k_x_kick_b(const XMember<T>* __restrict__ mem, int nlines, int hpitch,
                const typename TwEntry<T>::type* __restrict__ tw) {
    unsigned member, b;
    batch_block(member, b);
    const XMember<T>& mm = mem[member];
    x_kick_body<T, P, F, true, NW, VK>(mm.psi, mm.hat, nlines, hpitch, tw, mm.ka, b);
}

// rho = |psi|^2 of real-space values v[] (positions x = N2*kq + j), forward R2C
// along x into the half-spectrum row h, plus FP64 sum/max partials per wave.
template <class T, class F, int NW>
WDM_FI void density_r2c(const cx<T>* v, bool ok, const LaneMap& m, cx<T>* xb, cx<T>* wbuf,
                        const typename TwEntry<T>::type* __restrict__ tw, cx<T>* __restrict__ hdata,
                        unsigned hbase, double* __restrict__ partials) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, N = F::N;
    cx<T> r[N1];
    double msum = 0.0, mmax = 0.0;
    WDM_UNROLL
    for (int kq = 0; kq < N1; ++kq) {
        const T rr = norm2(v[kq]);
        r[kq] = mk(rr, T(0));
        msum += double(rr);
        mmax = double(rr) > mmax ? double(rr) : mmax;
    }
    if (!ok) { msum = 0.0; mmax = 0.0; }
    cx<T> X[M][N2];
    fwd2<F, -1, PoissonPolicy>(r, X, m.j, xb, tw);
    if (ok) {
        WDM_UNROLL
        for (int i = 0; i < M; ++i)
            WDM_UNROLL
            for (int k2 = 0; k2 < N2; ++k2) {
                const int k = (i * N2 + m.j) + N1 * k2;
                if (2 * k <= N) stg(hdata, hbase + unsigned(k), X[i][k2]);
            }
    }
    if (partials) {
        double s, mx;
        wave_sum_max(wbuf, m.lane, msum, mmax, s, mx);
        if (m.lane == 0) {
            partials[2 * m.w] = s;
            partials[2 * m.w + 1] = mx;
        }
    }
}

// Inverse x-FFT (+ absorber along x) of psi rows, store psi in real space,
// then rho = |psi|^2 -> R2C_x -> rho_hat.  `border` is a length-Nx FP64
// vector or null (src/solver/KDK.jl:95-96; broadcast along x, Q4).
// NTS: psi is read and written with the streaming (non-temporal) hint, so
// that it does not displace the half spectrum this pass writes from the
// memory-side cache (MI355X: 256 MB Infinity Cache) before the Poisson
// passes reread it.  Worth it only when psi is larger than that cache and the
// half spectrum is not (384^3 FP32: the step 1.540 -> 1.515 ms; at 256^3,
// where psi fits, it costs 7%).
template <class T, class P, class F, int NW, bool NTS = false>
WDM_FI void x_inv_dens_body(cx<T>* __restrict__ psi, cx<T>* __restrict__ rhohat, int nlines, int hpitch,
                            const typename TwEntry<T>::type* __restrict__ tw, const double* __restrict__ border,
                            double* __restrict__ partials, unsigned bid) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, N = F::N;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    tw = stage_table<F::N>(tw_s, tw);
    const LaneMap m = lane_map<F, NW>(bid);
    const int l = m.w * F::LPW + m.lw;
    const bool ok = l < nlines;
    const int lc = ok ? l : 0;
    const unsigned rbase = unsigned(lc) * unsigned(N);
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    const cx<T> zero = mk(T(0), T(0));
    cx<T> X[M][N2];
    cx<T> v[N1];
    if constexpr (NTS) {
        // Streaming loads in row order: the M loads of one p cover M * N2
        // contiguous values, so the two 64-byte halves of each 128-byte line
        // (FP32) are requested by consecutive loads.  Issued i-major, the
        // second half comes N2 loads after the first, and with the streaming
        // hint the pass fetched 1.58x the bytes of psi it reads (384^3 FP32:
        // 247 -> 232 us in row order).
        WDM_UNROLL
        for (int p = 0; p < N2; ++p)
            WDM_UNROLL
            for (int i = 0; i < M; ++i) X[i][p] = ok ? ldg_nt(psi, rbase + unsigned((i * N2 + m.j) + N1 * p)) : zero;
    } else {
        WDM_UNROLL
        for (int i = 0; i < M; ++i)
            WDM_UNROLL
            for (int p = 0; p < N2; ++p) X[i][p] = WDM_LDG(ok, psi, rbase + unsigned((i * N2 + m.j) + N1 * p));
    }
    inv2<F, +1, P>(X, v, m.j, xb, tw);
    if (border) {
        WDM_UNROLL
        for (int kq = 0; kq < N1; ++kq) v[kq] = scale_dyn<P>(v[kq], border[N2 * kq + m.j]);
    }
    if (ok) {
        WDM_UNROLL
        for (int kq = 0; kq < N1; ++kq) {
            if constexpr (NTS) stg_nt(psi, rbase + unsigned(N2 * kq + m.j), v[kq]);
            else stg(psi, rbase + unsigned(N2 * kq + m.j), v[kq]);
        }
    }
    if (rhohat) {
        density_r2c<T, F, NW>(v, ok, m, xb, &xbuf[m.wv][0], tw, rhohat, unsigned(lc) * unsigned(hpitch), partials);
    }
}
template <class T, class P, class F, int NW, bool NTS = false>
WDM_KERNEL_W(NW * 64, (std::is_same<T, float>::value ? WDM_MINW_XINV_F32 : 1))
void k_x_inv_dens(cx<T>* __restrict__ psi, cx<T>* __restrict__ rhohat, int nlines, int hpitch,
                  const typename TwEntry<T>::type* __restrict__ tw, const double* __restrict__ border,
                  double* __restrict__ partials) {
    x_inv_dens_body<T, P, F, NW, NTS>(psi, rhohat, nlines, hpitch, tw, border, partials, blockIdx.x);
}
template <class T, class P, class F, int NW>
WDM_KERNEL_W(NW * 64, (std::is_same<T, float>::value ? WDM_MINW_XINV_F32 : 1))
void k_x_inv_dens_b(const XMember<T>* __restrict__ mem, int nlines, int hpitch,
                    const typename TwEntry<T>::type* __restrict__ tw) {
    unsigned member, b;
    batch_block(member, b);
    const XMember<T>& mm = mem[member];
    x_inv_dens_body<T, P, F, NW>(mm.psi, mm.hat, nlines, hpitch, tw, mm.border, mm.partials, b);
}

// rho_hat (and sum/max partials) from psi in real space (start of a run, or
// after the host replaced psi).
template <class T, class F, int NW>
WDM_KERNEL(NW * 64)
void k_x_dens(const cx<T>* __restrict__ psi, cx<T>* __restrict__ rhohat, int nlines, int hpitch,
              const typename TwEntry<T>::type* __restrict__ tw, double* __restrict__ partials) {
    constexpr int N1 = F::N1, N2 = F::N2, N = F::N;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    tw = stage_table<F::N>(tw_s, tw);
    const LaneMap m = lane_map<F, NW>();
    const int l = m.w * F::LPW + m.lw;
    const bool ok = l < nlines;
    const int lc = ok ? l : 0;
    const unsigned rbase = unsigned(lc) * unsigned(N);
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    const cx<T> zero = mk(T(0), T(0));
    cx<T> v[N1];
    WDM_UNROLL
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = WDM_LDG(ok, psi, rbase + unsigned(N2 * n1 + m.j));
    density_r2c<T, F, NW>(v, ok, m, xb, &xbuf[m.wv][0], tw, rhohat, unsigned(lc) * unsigned(hpitch), partials);
}

// R2C_x of a real FP64 source array (standalone Poisson solve of a host rho).
template <class T, class F, int NW>
WDM_KERNEL(NW * 64)
void k_x_r2c_real(const double* __restrict__ src, cx<T>* __restrict__ rhohat, int nlines, int hpitch,
                  const typename TwEntry<T>::type* __restrict__ tw) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, N = F::N;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    tw = stage_table<F::N>(tw_s, tw);
    const LaneMap m = lane_map<F, NW>();
    const int l = m.w * F::LPW + m.lw;
    const bool ok = l < nlines;
    const int lc = ok ? l : 0;
    const double* row = src + size_t(lc) * N;
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    cx<T> v[N1];
    WDM_UNROLL
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = mk(ok ? T(row[N2 * n1 + m.j]) : T(0), T(0));
    cx<T> X[M][N2];
    fwd2<F, -1, PoissonPolicy>(v, X, m.j, xb, tw);
    if (ok) {
        cx<T>* h = rhohat + size_t(lc) * hpitch;
        WDM_UNROLL
        for (int i = 0; i < M; ++i)
            WDM_UNROLL
            for (int k2 = 0; k2 < N2; ++k2) {
                const int k = (i * N2 + m.j) + N1 * k2;
                if (2 * k <= N) h[k] = X[i][k2];
            }
    }
}

// C2R_x of a half spectrum into a real FP64 array (Phi on the grid).
template <class T, class F, int NW>
WDM_KERNEL(NW * 64)
void k_x_c2r_real(const cx<T>* __restrict__ hsrc, double* __restrict__ dst, int nlines, int hpitch,
                  const typename TwEntry<T>::type* __restrict__ tw) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, N = F::N;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    tw = stage_table<F::N>(tw_s, tw);
    const LaneMap m = lane_map<F, NW>();
    const int l = m.w * F::LPW + m.lw;
    const bool ok = l < nlines;
    const int lc = ok ? l : 0;
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    cx<T> X[M][N2];
    cx<T> v[N1];
    hermitian_row<F>(hsrc, unsigned(lc) * unsigned(hpitch), ok, m, &xbuf[m.wv][0], X);
    inv2<F, +1, PoissonPolicy>(X, v, m.j, xb, tw);
    if (ok) {
        double* row = dst + size_t(lc) * N;
        WDM_UNROLL
        for (int kq = 0; kq < N1; ++kq) row[N2 * kq + m.j] = double(v[kq].x);
    }
}

// Plain x-FFT of psi rows in place (DIR -1 forward, +1 inverse, unnormalised).
template <class T, class P, class F, int DIR, int NW>
WDM_KERNEL(NW * 64)
void k_x_fft(cx<T>* __restrict__ psi, int nlines, const typename TwEntry<T>::type* __restrict__ tw) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, N = F::N;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    tw = stage_table<F::N>(tw_s, tw);
    const LaneMap m = lane_map<F, NW>();
    const int l = m.w * F::LPW + m.lw;
    const bool ok = l < nlines;
    const int lc = ok ? l : 0;
    const unsigned rbase = unsigned(lc) * unsigned(N);
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    const cx<T> zero = mk(T(0), T(0));
    cx<T> v[N1];
    cx<T> X[M][N2];
    WDM_UNROLL
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = WDM_LDG(ok, psi, rbase + unsigned(N2 * n1 + m.j));
    fwd2<F, DIR, P>(v, X, m.j, xb, tw);
    if (ok) {
        WDM_UNROLL
        for (int i = 0; i < M; ++i)
            WDM_UNROLL
            for (int k2 = 0; k2 < N2; ++k2) stg(psi, rbase + unsigned((i * N2 + m.j) + N1 * k2), X[i][k2]);
    }
}

// ------------------------------------------------------------------ read-only reductions
// The diagnostics of wdm_reduce / wdm_cfl as line passes that read the state
// and write only per-wave partials: no N-sized scratch.  The spectral kinetic
// sum splits by axis, sum_k k^2 |psi_hat_k|^2 = sum_a (N / n_a) sum k_a^2
// |F_a psi|^2 over the grid (Parseval along the two other axes of the
// unnormalised transform), so each pass needs only the 1-D transform of its
// own lines.

// Deterministic wave reduction (tree over the lanes in a fixed order) through
// a 64-double LDS buffer of the wave; every lane returns the result.
WDM_FI double wave_tree(double* buf, int lane, double v, bool is_max) {
    buf[lane] = v;
    WDM_WAVE_SYNC();
    WDM_UNROLL
    for (int s = kWave / 2; s > 0; s >>= 1) {
        if (lane < s) {
            const double a = buf[lane], b = buf[lane + s];
            buf[lane] = is_max ? (b > a ? b : a) : a + b;
        }
        WDM_WAVE_SYNC();
    }
    const double r = buf[0];
    WDM_WAVE_SYNC();
    return r;
}

// signed FFT-order frequency index of position k on an axis of length n
WDM_HDFI int fft_signed(int k, int n) { return 2 * k < n ? k : k - n; }

// partial slots of k_x_red, per launched wave
enum RedSlot { RS_MASS = 0, RS_RHOMAX = 1, RS_RHOPHI = 2, RS_PHIMAX = 3, RS_RHOV = 4, RS_KX = 5, RS_COUNT = 6 };

// x lines: rho = |psi|^2 (sum, max), Phi from the Poisson half spectrum by a
// C2R in registers (sum rho Phi, max |Phi|; phihat null: no self-gravity),
// V from the plan's models at the kick time in ka (sum rho V), and
// sum kx^2 |F_x psi|^2 (kx in units of kscale^(1/2), FFT order).
template <class T, class P, class F, int NW>
WDM_KERNEL(NW * 64)
void k_x_red(const cx<T>* __restrict__ psi, const cx<T>* __restrict__ phihat, int nlines, int hpitch,
             const typename TwEntry<T>::type* __restrict__ tw, KickArgs ka, double kscale,
             double* __restrict__ partials) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, N = F::N;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    WDM_SHARED double rbuf[NW][kWave];
    tw = stage_table<F::N>(tw_s, tw);
    const LaneMap m = lane_map<F, NW>();
    const int l = m.w * F::LPW + m.lw;
    const bool ok = l < nlines;
    const int lc = ok ? l : 0;
    const int y = lc % ka.Ny, z = lc / ka.Ny;
    const unsigned rbase = unsigned(lc) * unsigned(N);
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    const cx<T> zero = mk(T(0), T(0));
    cx<T> X[M][N2];
    cx<T> v[N1];
    T phi[N1];
    if (phihat) {
        const unsigned hbase = unsigned(lc) * unsigned(hpitch);
        hermitian_row<F>(phihat, hbase, ok, m, &xbuf[m.wv][0], X);
        inv2<F, +1, PoissonPolicy>(X, v, m.j, xb, tw);
        WDM_UNROLL
        for (int kq = 0; kq < N1; ++kq) phi[kq] = v[kq].x;
    } else {
        WDM_UNROLL
        for (int kq = 0; kq < N1; ++kq) phi[kq] = T(0);
    }
    const LineV lv = line_V(ka.vm, ka.ys[y], ka.zs[z]);
    WDM_SCHED_FENCE();
    WDM_UNROLL
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = WDM_LDG(ok, psi, rbase + unsigned(N2 * n1 + m.j));
    double ms = 0.0, mx = 0.0, rp = 0.0, pm = 0.0, rv = 0.0;
    WDM_UNROLL
    for (int n1 = 0; n1 < N1; ++n1) {
        const int x = N2 * n1 + m.j;
        const double r = double(norm2(v[n1]));
        ms += r;
        mx = r > mx ? r : mx;
        const double ph = double(phi[n1]);
        rp += r * ph;
        const double aph = __builtin_fabs(ph);
        pm = aph > pm ? aph : pm;
        if (ka.vm.flags) {
            const double tx = double(x) * ka.xinv;
            const double xc = (1.0 - tx) * ka.x0 + tx * ka.x1;
            rv += r * eval_V_line<T>(ka.vm, lv, xc, unsigned(lc) * N + x, v[n1]);
        }
    }
    fwd2<F, -1, P>(v, X, m.j, xb, tw);
    double ks = 0.0;
    WDM_UNROLL
    for (int i = 0; i < M; ++i)
        WDM_UNROLL
        for (int k2 = 0; k2 < N2; ++k2) {
            const int mi = fft_signed((i * N2 + m.j) + N1 * k2, N);
            ks += double(mi * mi) * double(norm2(X[i][k2]));
        }
    if (!ok) {
        ms = mx = rp = pm = rv = ks = 0.0;
    }
    double* rb = rbuf[m.wv];
    const double s_ms = wave_tree(rb, m.lane, ms, false);
    const double s_mx = wave_tree(rb, m.lane, mx, true);
    const double s_rp = wave_tree(rb, m.lane, rp, false);
    const double s_pm = wave_tree(rb, m.lane, pm, true);
    const double s_rv = wave_tree(rb, m.lane, rv, false);
    const double s_ks = wave_tree(rb, m.lane, ks, false);
    if (m.lane == 0) {
        double* o = partials + size_t(RS_COUNT) * size_t(m.w);
        o[RS_MASS] = s_ms;
        o[RS_RHOMAX] = s_mx;
        o[RS_RHOPHI] = s_rp;
        o[RS_PHIMAX] = s_pm;
        o[RS_RHOV] = s_rv;
        o[RS_KX] = s_ks * kscale;
    }
}

// y or z lines (column layout of k_col): sum k_line^2 |F_line psi|^2 per wave.
template <class T, class P, class F, int NW>
WDM_KERNEL(NW * 64)
void k_col_red(const cx<T>* __restrict__ data, int ncols, int nother, unsigned sline, unsigned sother,
               const typename TwEntry<T>::type* __restrict__ tw, double kscale, double* __restrict__ partials) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, N = F::N;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    WDM_SHARED double rbuf[NW][kWave];
    tw = stage_table<F::N>(tw_s, tw);
    const LaneMap m = lane_map<F, NW>();
    const int ntiles = (ncols + F::LPW - 1) / F::LPW;
    const int o = m.w / ntiles, tile = m.w % ntiles;
    const int c = tile * F::LPW + m.lw;
    const bool ok = (o < nother) && (c < ncols);
    const int oc = ok ? o : 0, cc = ok ? c : 0;
    const unsigned base = unsigned(oc) * sother + unsigned(cc) + unsigned(m.j) * sline;
    const unsigned step = unsigned(N2) * sline;
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    const cx<T> zero = mk(T(0), T(0));
    cx<T> v[N1];
    cx<T> X[M][N2];
    WDM_UNROLL
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = WDM_LDG(ok, data, base + unsigned(n1) * step);
    fwd2<F, -1, P>(v, X, m.j, xb, tw);
    double ks = 0.0;
    WDM_UNROLL
    for (int i = 0; i < M; ++i)
        WDM_UNROLL
        for (int k2 = 0; k2 < N2; ++k2) {
            const int mi = fft_signed((i * N2 + m.j) + N1 * k2, N);
            ks += double(mi * mi) * double(norm2(X[i][k2]));
        }
    if (!ok) ks = 0.0;
    const double s = wave_tree(rbuf[m.wv], m.lane, ks, false);
    if (m.lane == 0) partials[m.w] = s * kscale;
}

// Element-wise conversions between the plan precision and FP64/FP32 host layouts.
template <class TD, class TS>
WDM_KERNEL(256)
void k_convert(cx<TD>* __restrict__ dst, const cx<TS>* __restrict__ src, size_t n) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256)
        dst[i] = mk(TD(src[i].x), TD(src[i].y));
}

}  // namespace wdm
