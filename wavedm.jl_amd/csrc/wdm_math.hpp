// Complex arithmetic, compile-time twiddle constants and the precision policies
// used by every pass kernel.
//
// Precision policies:
//   Plain - ordinary multiplies by rounded constants.
//   Comp  - (FP32 only) every multiply by a constant or twiddle uses the
//           constant split as hi+lo floats (fma(x,hi,x*lo)), so the effective
//           constant carries ~48 bits.  A constant rounded once to FP32 and
//           applied to every cell every step would bias |psi|^2 by up to
//           ~6e-8 per use; the split removes that bias, leaving only unbiased
//           round-to-nearest errors in the data.
#pragma once

#include "wdm_port.hpp"

namespace wdm {

template <class T>
struct alignas(2 * sizeof(T)) cx {
    T x, y;
};

template <class T> WDM_HDFI cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <class T> WDM_HDFI cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <class T> WDM_HDFI cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <class T> WDM_HDFI cx<T> operator-(cx<T> a) { return mk(-a.x, -a.y); }
template <class T> WDM_HDFI cx<T> operator*(cx<T> a, T s) { return mk(a.x * s, a.y * s); }
// multiply by i*s where s = +1 or -1 (exact)
template <int S, class T> WDM_HDFI cx<T> mul_i(cx<T> a) {
    if constexpr (S > 0) return mk(-a.y, a.x);
    else return mk(a.y, -a.x);
}

// Element access at a 32-bit index from a wave-uniform base pointer: the byte
// offset is formed in 32 bits so the backend can use scalar-base addressing
// (one VGPR per address instead of a 64-bit pair).  Arrays stay below 4 GiB.
template <class T>
WDM_HDFI T ldg(const T* base, unsigned idx) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * unsigned(sizeof(T)));
}
template <class T>
WDM_HDFI void stg(T* base, unsigned idx, const T& v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * unsigned(sizeof(T))) = v;
}

// The same with the non-temporal (streaming) hint: for data that the next
// pass does not reread from the caches, so that it does not displace data the
// next pass does reread.
template <class T>
WDM_HDFI cx<T> ldg_nt(const cx<T>* base, unsigned idx) {
#ifdef WDM_GPU
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 r = __builtin_nontemporal_load(
        reinterpret_cast<const v2*>(reinterpret_cast<const char*>(base) + idx * unsigned(sizeof(cx<T>))));
    return mk(r.x, r.y);
#else
    return ldg(base, idx);
#endif
}
template <class T>
WDM_HDFI void stg_nt(cx<T>* base, unsigned idx, const cx<T>& v) {
#ifdef WDM_GPU
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 r;
    r.x = v.x;
    r.y = v.y;
    __builtin_nontemporal_store(r, reinterpret_cast<v2*>(reinterpret_cast<char*>(base) + idx * unsigned(sizeof(cx<T>))));
#else
    stg(base, idx, v);
#endif
}

WDM_HDFI float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
WDM_HDFI double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

// |a|^2 with an explicit fma: the same bits in every kernel that evaluates it
// (left to contraction, two kernels computing the density of the same psi
// could round it differently)
template <class T> WDM_HDFI T norm2(cx<T> a) { return fmaT(a.x, a.x, a.y * a.y); }

struct Plain {};
struct Comp {};

// ---------------------------------------------------------------- constexpr trig
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kRsqrt2 = 0.70710678118654752440084436210485;

// The table generators below run on the host (plan tables) and on the device
// (the allocation-free Poisson call builds its tables in place): with
// contraction off both evaluate the same IEEE operations, so the tables are
// bit-identical.
#if defined(__clang__)
#define WDM_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define WDM_NO_CONTRACT
#endif

constexpr double taylor_sin(double t) {
    WDM_NO_CONTRACT
    double t2 = t * t, term = t, s = t;
    for (int n = 1; n < 14; ++n) {
        term *= -t2 / double((2 * n) * (2 * n + 1));
        s += term;
    }
    return s;
}
constexpr double taylor_cos(double t) {
    WDM_NO_CONTRACT
    double t2 = t * t, term = 1.0, c = 1.0;
    for (int n = 1; n < 14; ++n) {
        term *= -t2 / double((2 * n - 1) * (2 * n));
        c += term;
    }
    return c;
}

struct CS {
    double c, s;
};

// cos and sin of 2*pi*num/den, reduced to the first octant with exact
// symmetries so that multiples of pi/4 come out exact.
constexpr CS cs2pi(long num, long den) {
    WDM_NO_CONTRACT
    long n = ((num % den) + den) % den;
    if ((8 * n) % den == 0) {
        const double r = kRsqrt2;
        switch ((8 * n) / den) {
            case 0: return {1.0, 0.0};
            case 1: return {r, r};
            case 2: return {0.0, 1.0};
            case 3: return {-r, r};
            case 4: return {-1.0, 0.0};
            case 5: return {-r, -r};
            case 6: return {0.0, -1.0};
            default: return {r, -r};
        }
    }
    if ((12 * n) % den == 0) {
        const double h3 = 0.86602540378443864676372317075294;  // sqrt(3)/2
        switch ((12 * n) / den) {
            case 1: return {h3, 0.5};
            case 2: return {0.5, h3};
            case 4: return {-0.5, h3};
            case 5: return {-h3, 0.5};
            case 7: return {-h3, -0.5};
            case 8: return {-0.5, -h3};
            case 10: return {0.5, -h3};
            default: return {h3, -0.5};  // 11
        }
    }
    bool neg_s = false, neg_c = false, swap = false;
    long d = den;
    if (2 * n > d) { n = d - n; neg_s = true; }            // x -> 1 - x
    if (4 * n > d) { n = d - 2 * n; d = 2 * d; neg_c = true; }  // x -> 1/2 - x
    if (8 * n > d) { n = d - 4 * n; d = 4 * d; swap = true; }   // x -> 1/4 - x
    double t = kTwoPi * double(n) / double(d);
    double c = taylor_cos(t), s = taylor_sin(t);
    if (swap) { double tmp = c; c = s; s = tmp; }
    if (neg_c) c = -c;
    if (neg_s) s = -s;
    return {c, s};
}

template <int R>
struct TwTab {
    double c[R];
    double s[R];
    constexpr TwTab() : c(), s() {
        for (int j = 0; j < R; ++j) {
            CS v = cs2pi(j, R);
            c[j] = v.c;
            s[j] = v.s;
        }
    }
};
template <int R>
inline constexpr TwTab<R> kTw{};

// ---------------------------------------------------------------- reciprocal
// FP32 reciprocal to about 1 ulp (v_rcp_f32); used where a relative error of
// a few 1e-7 is below the arithmetic's own (Poisson factor of FP32 plans).
WDM_FI float rcpf(float x) {
#ifdef WDM_GPU
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// ---------------------------------------------------------------- hi/lo splits
// v = hi + lo, hi = float(v), lo = float(v - hi).
struct HL {
    float h, l;
};
constexpr HL split_hl(double v) {
    const float h = float(v);
    return {h, float(v - double(h))};
}
// A power-of-two constant (the 0.5 of cos 60 deg) makes hi*x exact for any
// FP32 x.  Small correction terms accumulated into an fma whose product is
// exact always round away, so in a two-product sum the exact product must go
// in the last fma and the corrections must ride on the inexact one.
WDM_HDFI bool is_pow2f(float h) {
    return h != 0.0f && (__builtin_bit_cast(unsigned, h) & 0x7FFFFFu) == 0u;
}

// ---------------------------------------------------------------- constant multiplies
// x * k for a real constant k (given in FP64).
template <class P, class T> WDM_FI T kmul(T x, double k);
template <> WDM_FI double kmul<Plain, double>(double x, double k) { return x * k; }
template <> WDM_FI double kmul<Comp, double>(double x, double k) { return x * k; }
template <> WDM_FI float kmul<Plain, float>(float x, double k) { return x * float(k); }
template <> WDM_FI float kmul<Comp, float>(float x, double k) {
    const HL c = split_hl(k);
    if (c.l == 0.0f) return x * c.h;
    return fmaT(x, c.h, x * c.l);
}
template <class P, class T> WDM_FI cx<T> kmulc(cx<T> x, double k) {
    return mk(kmul<P, T>(x.x, k), kmul<P, T>(x.y, k));
}

// x * (c + i s) for FP64 constants c, s.
template <class P, class T> WDM_FI cx<T> cmulk(cx<T> a, double c, double s) {
    if constexpr (std::is_same<T, double>::value || std::is_same<P, Plain>::value) {
        const T cc = T(c), ss = T(s);
        return mk(fmaT(a.x, cc, -(a.y * ss)), fmaT(a.x, ss, a.y * cc));
    } else {
        const HL cc = split_hl(c), ss = split_hl(s);
        const float ch = cc.h, cl = cc.l, sh = ss.h, sl = ss.l;
        // re = x c - y s ; im = x s + y c  (constants: the branches fold away)
        float t = fmaT(-a.y, sl, a.x * cl);
        if (is_pow2f(sh)) {
            t = fmaT(a.x, ch, t);
            t = fmaT(-a.y, sh, t);
        } else {
            t = fmaT(-a.y, sh, t);
            t = fmaT(a.x, ch, t);
        }
        float u = fmaT(a.y, cl, a.x * sl);
        if (is_pow2f(ch)) {
            u = fmaT(a.x, sh, u);
            u = fmaT(a.y, ch, u);
        } else {
            u = fmaT(a.y, ch, u);
            u = fmaT(a.x, sh, u);
        }
        return mk(t, u);
    }
}

// x * exp(DIR * 2 pi i j / R) with compile-time R, DIR and (after unrolling)
// constant j; trivial angles become exact swaps/negations.
template <int R, int DIR, class P, class T>
WDM_FI cx<T> mul_w(cx<T> x, int j) {
    if (j == 0) return x;
    if ((4 * j) % R == 0) {
        const int q = (4 * j) / R;
        if (q == 2) return -x;
        if ((q == 1) == (DIR > 0)) return mul_i<+1>(x);
        return mul_i<-1>(x);
    }
    if ((8 * j) % R == 0) {
        const int o = (8 * j) / R;  // 1,3,5,7
        const bool cpos = (o == 1 || o == 7);
        const bool spos = ((o == 1 || o == 3) == (DIR > 0));
        const T re = (cpos ? x.x : -x.x) + (spos ? -x.y : x.y);
        const T im = (spos ? x.x : -x.x) + (cpos ? x.y : -x.y);
        return mk(kmul<P, T>(re, kRsqrt2), kmul<P, T>(im, kRsqrt2));
    }
    const double c = kTw<R>.c[j];
    const double s = DIR > 0 ? kTw<R>.s[j] : -kTw<R>.s[j];
    return cmulk<P, T>(x, c, s);
}

// Runtime twiddle table entries: FP64 {c, s}; FP32 {c_hi, s_hi, c_lo, s_lo}.
struct TwF {
    float ch, sh, cl, sl;
};
template <class T> struct TwEntry;
template <> struct TwEntry<double> { typedef cx<double> type; };
template <> struct TwEntry<float> { typedef TwF type; };

// Runtime table entries are split by split_hl_table: a "hi" part that would
// be a power of two (1, 0.5 of the angles at multiples of 30 and 90 degrees)
// is moved one ulp toward zero and the exact remainder goes into "lo".  Every
// hi * x product is then inexact for general x, so one fixed fma order is
// safe: the corrections ride on inexact products and the final fma rounds the
// whole sum once (see is_pow2f).  No per-lane operand selects are needed.
WDM_HDFI HL split_hl_table(double v) {
    float h = float(v);
    if (is_pow2f(h)) {
        unsigned b = __builtin_bit_cast(unsigned, h);
        h = __builtin_bit_cast(float, b - 1u);  // same sign, one ulp smaller magnitude
    }
    return {h, float(v - double(h))};
}

// x * (table entry)^DIR-conjugated: forward (DIR<0) uses exp(-i theta).
template <int DIR, class P>
WDM_FI cx<double> mul_tab(cx<double> a, const cx<double>& w) {
    const double s = DIR > 0 ? w.y : -w.y;
    return mk(fmaT(a.x, w.x, -(a.y * s)), fmaT(a.x, s, a.y * w.x));
}
template <int DIR, class P>
WDM_FI cx<float> mul_tab(cx<float> a, const TwF& w) {
    const float sh = DIR > 0 ? w.sh : -w.sh;
    if constexpr (std::is_same<P, Plain>::value) {
        return mk(fmaT(a.x, w.ch, -(a.y * sh)), fmaT(a.x, sh, a.y * w.ch));
    } else {
        // entries from split_hl_table: no hi part is a power of two
        const float sl = DIR > 0 ? w.sl : -w.sl;
        float t = a.x * w.cl;
        t = fmaT(-a.y, sl, t);
        t = fmaT(-a.y, sh, t);
        t = fmaT(a.x, w.ch, t);
        float u = a.x * sl;
        u = fmaT(a.y, w.cl, u);
        u = fmaT(a.y, w.ch, u);
        u = fmaT(a.x, sh, u);
        return mk(t, u);
    }
}

// Multiply by an FP64 complex factor (c, s) computed at run time.  FP32
// compensated: the data are widened and the product is formed in FP64, so
// each component is rounded to FP32 once from a result accurate to ~1e-16
// (unbiased, whatever the factor's value).
template <class P>
WDM_FI cx<double> mul_dyn(cx<double> a, double c, double s) {
    return mk(fmaT(a.x, c, -(a.y * s)), fmaT(a.x, s, a.y * c));
}
template <class P>
WDM_FI cx<float> mul_dyn(cx<float> a, double c, double s) {
    if constexpr (std::is_same<P, Plain>::value) {
        const float cc = float(c), ss = float(s);
        return mk(fmaT(a.x, cc, -(a.y * ss)), fmaT(a.x, ss, a.y * cc));
    } else {
        const double x = double(a.x), y = double(a.y);
        return mk(float(fmaT(x, c, -(y * s))), float(fmaT(x, s, y * c)));
    }
}

}  // namespace wdm
