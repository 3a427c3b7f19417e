// Wave-autonomous line FFTs.
//
// A line of length N = N1*N2 is owned by N2 consecutive lanes of one 64-lane
// wavefront; the wave holds LPW = 64/N2 lines.  Lane j of a line keeps N1
// complex values in registers.  The two-level decomposition
//
//   fwd2: lane j holds x[N2*n1 + j]  -> DFT-N1 in registers, twiddle
//         W_N^(j*k1), then M = N1/N2 rounds of an N2 x N2 transpose through a
//         small per-wave LDS buffer, each followed by a DFT-N2.  Lane j ends
//         with X[(i*N2 + j) + N1*k2] in X[i][k2].
//   inv2: the transposed factorisation, starting from that layout and ending
//         with x[N2*kq + j] in v[kq].
//
// fwd2 followed by inv2 therefore round-trips without any reordering, which is
// what lets the z-pass do forward FFT * k-space factor * inverse FFT in one
// HBM read and one write, and lets the x-passes hand values between the
// Poisson inverse and the psi forward transform in registers.  The transposes
// are confined to one wave, so no __syncthreads is needed anywhere.
#pragma once

#include "wdm_dft.hpp"

namespace wdm {

template <int N1_, int N2_>
struct Fac {
    static constexpr int N1 = N1_, N2 = N2_, N = N1_ * N2_;
    static constexpr int M = N1_ / N2_;
    static constexpr int LPW = kWave / N2_;  // lines per wave
    static_assert(N1_ % N2_ == 0, "N2 must divide N1");
    static_assert((N2_ & (N2_ - 1)) == 0 && N2_ <= kWave, "N2 must be a power of two <= 64");
    // per-line stride of the transpose buffer, padded so the LPW lines of a
    // wave spread over the LDS banks
    static constexpr int XS = N2_ * N2_ + (N2_ >= 4 ? 8 : 1);
    static constexpr int XBUF = LPW * XS;  // elements per wave
};

// One N2 x N2 transpose between the lanes of a line: lane `me` sends out[b]
// to lane b and receives in[a] = (lane a's out[me]).  XOR swizzle on the inner
// index keeps both the write and the read bank-conflict free.
template <int N2, class T>
WDM_FI void xpose(cx<T>* __restrict__ xb, int me, const cx<T>* out, cx<T>* in) {
    if constexpr (N2 == 1) {
        in[0] = out[0];
    } else {
        WDM_UNROLL
        for (int b = 0; b < N2; ++b) xb[b * N2 + (me ^ b)] = out[b];
        WDM_WAVE_SYNC();
        WDM_UNROLL
        for (int a = 0; a < N2; ++a) in[a] = xb[me * N2 + (a ^ me)];
        WDM_WAVE_SYNC();
    }
}

// Twiddle table layouts (the FFTs read W^(j k1) in fwd2 and W^(q kp) in inv2):
//   TL = 0: the plain length-N table, W^e at e;
//   TL = 1: 2N entries, W^(j k1) at k1 N2 + j and W^(q kp) at N + kp N1 + q,
//           so the N2 lanes of a line read consecutive entries (no LDS bank
//           conflicts; the plain table puts lanes j = 0..N2-1 on the same
//           banks whenever k1 is even).
template <class F, int TL>
WDM_FI int tw_fwd_index(int j, int k1) {
    return TL ? k1 * F::N2 + j : j * k1;
}
template <class F, int TL>
WDM_FI int tw_inv_index(int q, int kp) {
    return TL ? F::N + kp * F::N1 + q : q * kp;
}

// Forward-layout two-level FFT (see file comment).  `tw` is the twiddle
// table for this axis in layout TL; `xb` the line's slice of the wave LDS
// buffer.
template <class F, int DIR, class P, int TL = 0, class T>
WDM_FI void fwd2(cx<T>* v, cx<T> (*X)[F::N2], int j, cx<T>* __restrict__ xb,
                 const typename TwEntry<T>::type* __restrict__ tw) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M;
    Dft<N1, DIR, P, T>::run(v);
    WDM_UNROLL
    for (int i = 0; i < M; ++i) {
        WDM_SCHED_FENCE();
        if constexpr (N2 > 1) {
            // twiddles of this chunk only (k1 = i*N2 + b)
            WDM_UNROLL
            for (int b = 0; b < N2; ++b) {
                const int k1 = i * N2 + b;
                if (k1 > 0)
                    v[k1] = mul_tab<DIR, P>(v[k1], tw[tw_fwd_index<F, TL>(j, k1)]);
            }
        }
        cx<T> in[N2];
        xpose<N2>(xb, j, v + i * N2, in);
        Dft<N2, DIR, P, T>::run(in);
        WDM_UNROLL
        for (int k2 = 0; k2 < N2; ++k2) X[i][k2] = in[k2];
    }
}

// Inverse-layout two-level FFT: lane j holds items q = i*N2 + j with
// X[i][p] = x_hat[q + N1*p]; returns v[kq] = x[N2*kq + j].
template <class F, int DIR, class P, int TL = 0, class T>
WDM_FI void inv2(cx<T> (*X)[F::N2], cx<T>* v, int j, cx<T>* __restrict__ xb,
                 const typename TwEntry<T>::type* __restrict__ tw) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M;
    WDM_UNROLL
    for (int i = 0; i < M; ++i) {
        WDM_SCHED_FENCE();
        Dft<N2, DIR, P, T>::run(X[i]);
        if constexpr (N2 > 1) {
            const int q = i * N2 + j;
            WDM_UNROLL
            for (int kp = 1; kp < N2; ++kp)
                X[i][kp] = mul_tab<DIR, P>(X[i][kp], tw[tw_inv_index<F, TL>(q, kp)]);
        }
    }
    WDM_UNROLL
    for (int i = 0; i < M; ++i) {
        cx<T> in[N2];
        xpose<N2>(xb, j, X[i], in);
        WDM_UNROLL
        for (int a = 0; a < N2; ++a) v[i * N2 + a] = in[a];
    }
    Dft<N1, DIR, P, T>::run(v);
}

}  // namespace wdm
