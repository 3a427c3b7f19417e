// Register-resident DFTs of small compile-time sizes (the per-lane codelets of
// every line FFT).  X[k] = sum_n x[n] exp(DIR * 2 pi i n k / R), unnormalised;
// DIR = -1 forward, +1 inverse.  Radix 2, 3, 4, 5 are written out; composite
// sizes recurse as R = A*B (A-point DFTs on stride-B subsequences, twiddle
// W_R^(b*ka), then B-point DFTs), with natural-order input and output.
#pragma once

#include "wdm_math.hpp"

namespace wdm {

constexpr int dft_split(int R) {
    // first factor of a composite size: prefer 4 (exact twiddles inside),
    // then 8, 2, 3, 5
    if (R % 4 == 0 && R != 4 && R != 8) return (R % 8 == 0 && R >= 64) ? 8 : 4;
    if (R == 8) return 2;
    if (R % 2 == 0 && R != 2) return 2;
    if (R % 3 == 0 && R != 3) return 3;
    if (R % 5 == 0 && R != 5) return 5;
    return R;
}

template <int R, int DIR, class P, class T>
struct Dft {
    static constexpr int A = dft_split(R);
    static constexpr int B = R / A;
    static_assert(A > 1 && B > 1 && A * B == R, "unsupported DFT size");
    static WDM_FI void run(cx<T>* v) {
        cx<T> y[R];
        WDM_UNROLL
        for (int b = 0; b < B; ++b) {
            cx<T> u[A];
            WDM_UNROLL
            for (int a = 0; a < A; ++a) u[a] = v[B * a + b];
            Dft<A, DIR, P, T>::run(u);
            WDM_UNROLL
            for (int ka = 0; ka < A; ++ka) y[ka * B + b] = mul_w<R, DIR, P, T>(u[ka], b * ka);
        }
        WDM_UNROLL
        for (int ka = 0; ka < A; ++ka) {
            cx<T> w[B];
            WDM_UNROLL
            for (int b = 0; b < B; ++b) w[b] = y[ka * B + b];
            Dft<B, DIR, P, T>::run(w);
            WDM_UNROLL
            for (int kb = 0; kb < B; ++kb) v[ka + A * kb] = w[kb];
        }
    }
};

template <int DIR, class P, class T>
struct Dft<1, DIR, P, T> {
    static WDM_FI void run(cx<T>*) {}
};

template <int DIR, class P, class T>
struct Dft<2, DIR, P, T> {
    static WDM_FI void run(cx<T>* v) {
        const cx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};

template <int DIR, class P, class T>
struct Dft<3, DIR, P, T> {
    // X1 = x0 - (x1+x2)/2 + i*DIR*(sqrt3/2)(x1-x2), X2 = conjugate-sign partner
    static WDM_FI void run(cx<T>* v) {
        const cx<T> t1 = v[1] + v[2];
        const cx<T> t2 = v[0] - t1 * T(0.5);
        const cx<T> d = kmulc<P, T>(v[1] - v[2], 0.86602540378443864676372317075294);
        v[0] = v[0] + t1;
        const cx<T> id = mul_i<DIR>(d);
        v[1] = t2 + id;
        v[2] = t2 - id;
    }
};

template <int DIR, class P, class T>
struct Dft<4, DIR, P, T> {
    static WDM_FI void run(cx<T>* v) {
        const cx<T> s02 = v[0] + v[2], d02 = v[0] - v[2];
        const cx<T> s13 = v[1] + v[3], d13 = mul_i<DIR>(v[1] - v[3]);
        v[0] = s02 + s13;
        v[2] = s02 - s13;
        v[1] = d02 + d13;
        v[3] = d02 - d13;
    }
};

template <int DIR, class P, class T>
struct Dft<5, DIR, P, T> {
    static WDM_FI void run(cx<T>* v) {
        constexpr double c1 = 0.30901699437494742410229341718282;   // cos(2pi/5)
        constexpr double c2 = -0.80901699437494742410229341718282;  // cos(4pi/5)
        constexpr double s1 = 0.95105651629515357211643933337938;   // sin(2pi/5)
        constexpr double s2 = 0.58778525229247312916870595463907;   // sin(4pi/5)
        const cx<T> a1 = v[1] + v[4], a2 = v[2] + v[3];
        const cx<T> b1 = v[1] - v[4], b2 = v[2] - v[3];
        const cx<T> x0 = v[0];
        v[0] = x0 + a1 + a2;
        const cx<T> p1 = x0 + kmulc<P, T>(a1, c1) + kmulc<P, T>(a2, c2);
        const cx<T> p2 = x0 + kmulc<P, T>(a1, c2) + kmulc<P, T>(a2, c1);
        const cx<T> q1 = kmulc<P, T>(b1, s1) + kmulc<P, T>(b2, s2);
        const cx<T> q2 = kmulc<P, T>(b1, s2) - kmulc<P, T>(b2, s1);
        const cx<T> iq1 = mul_i<DIR>(q1), iq2 = mul_i<DIR>(q2);
        v[1] = p1 + iq1;
        v[4] = p1 - iq1;
        v[2] = p2 + iq2;
        v[3] = p2 - iq2;
    }
};

}  // namespace wdm
