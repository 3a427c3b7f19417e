// Fused x passes of the adjoint step (the backward sweep of wdm_adjoint_*).
//
// One backward step, given h = B g_{n+1} already transformed along x:
//   g'  = U^H h             y, z (conj drift, 1/N), y passes, then the x inverse here
//   s   = tau Im(conj(g') psi'),  psi' = E psi_n,  E = exp(-i theta),
//         theta = tau (Phi_n + V)
//   u   = Phi-operator(s)   R2C_x here, the Poisson y/z passes, C2R_x in the next kernel
//   g_n = conj(E) g' + 2 (u + kappa s) psi_n;  then B g_n and FFT_x for step n-1
// k_adj_xs covers [IFFT_x -> g' ; s ; parameter partials ; R2C_x(s)] and
// k_adj_xg covers [C2R_x(u) ; g_n ; * B ; FFT_x], so a backward step has the
// same eight passes as a forward step.  theta = tau (Phi + V) and its quadrant
// reduction are FP64 in both precisions (theta spans many turns); the
// rotations, s and g_n are formed in the plan precision, the parameter sums
// are accumulated in FP64.
// No counterpart in the reference (it defines no derivative for the step).
#pragma once

#include "wdm_kernels.hpp"

namespace wdm {

// parameter sums: dM, da, dcx, dcy, dcz of the Hernquist model; sum s (Phi + V)
// (dt scale); dTxx, dTyy, dTzz, dTxy, dTxz, dTyz of the quadrupole's T0
constexpr int kAdjParams = 12;

struct AdjXArgs {
    double tau;                  // (m/hbar) dt_kick
    int Ny;
    const double* xs;            // coordinates per axis (FP64)
    const double* ys;
    const double* zs;
    const double* phi;           // taped real Phi_n (FP64) or null (no self-gravity)
    const double* border;        // length-Nx absorber or null (k_adj_xg only)
    VModel vm;                   // V at the step's kick time (includes the kappa term)
    double qR[9];                // quadrupole rotation R(t) at the kick time, row-major:
                                 // V = 1/2 u^T T0 u with u = R^T (x - c)
};

// E = exp(-i theta) as (cos r, sin r, quadrant q) of theta = q pi/2 + r
template <class T>
struct AdjPhase {
    T c, s;
    int q;
};
template <class T>
WDM_FI AdjPhase<T> adj_phase(double theta) {
    AdjPhase<T> e;
    const double r = reduce_quadrant(theta, e.q);
    sincos_small(T(r), e.s, e.c);
    return e;
}
// a * exp(-i theta)
template <class T>
WDM_FI cx<T> adj_rot(cx<T> a, const AdjPhase<T>& e) {
    return rot_quadrant(mk(fmaT(a.x, e.c, a.y * e.s), fmaT(a.y, e.c, -(a.x * e.s))), e.q);
}
// a * exp(+i theta)
template <class T>
WDM_FI cx<T> adj_rot_conj(cx<T> a, const AdjPhase<T>& e) {
    const cx<T> b = adj_rot(mk(a.x, -a.y), e);
    return mk(b.x, -b.y);
}
WDM_FI float adj_sqrt(float v) { return __builtin_sqrtf(v); }
WDM_FI double adj_sqrt(double v) { return __builtin_sqrt(v); }
// Phi_n + V (FP64) at the element at x of a line (y, z fixed)
template <class T>
WDM_FI double adj_potential(const AdjXArgs& a, const LineV& lv, int x, unsigned idx, cx<T> pn) {
    const double V = a.vm.flags ? eval_V_line<T>(a.vm, lv, a.xs[x], idx, pn) : 0.0;
    return (a.phi ? a.phi[idx] : 0.0) + V;
}

// In: g rows transformed along x.  Out: g' in real space (in place), the
// half spectrum of s along x into shat (when given) and kAdjParams partial
// sums per workgroup into partials[kAdjParams * blockIdx.x + j] (the
// quadrupole's six only when QG, zeros otherwise).
template <class T, class P, class F, int NW, bool QG>
WDM_KERNEL(NW * 64)
void k_adj_xs(cx<T>* __restrict__ g, const cx<T>* __restrict__ psi_n, cx<T>* __restrict__ shat, int nlines,
              int hpitch, const typename TwEntry<T>::type* __restrict__ tw, AdjXArgs a,
              double* __restrict__ partials) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, N = F::N;
    constexpr int NP = QG ? kAdjParams : 6;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    WDM_SHARED double red[NW][NP * kWave];
    WDM_SHARED double wred[NW][NP];
    tw = stage_table<F::N>(tw_s, tw);
    const LaneMap m = lane_map<F, NW>();
    const int l = m.w * F::LPW + m.lw;
    const bool ok = l < nlines;
    const int lc = ok ? l : 0;
    const int y = lc % a.Ny, z = lc / a.Ny;
    const unsigned rbase = unsigned(lc) * unsigned(N);
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    const cx<T> zero = mk(T(0), T(0));
    cx<T> X[M][N2];
    cx<T> v[N1];
    WDM_UNROLL
    for (int i = 0; i < M; ++i)
        WDM_UNROLL
        for (int p = 0; p < N2; ++p) X[i][p] = WDM_LDG(ok, g, rbase + unsigned((i * N2 + m.j) + N1 * p));
    inv2<F, +1, P>(X, v, m.j, xb, tw);
    if (ok) {
        WDM_UNROLL
        for (int kq = 0; kq < N1; ++kq) stg(g, rbase + unsigned(N2 * kq + m.j), v[kq]);
    }
    const LineV lv = line_V(a.vm, a.ys[y], a.zs[z]);
    const bool hq = (a.vm.flags & V_HERNQUIST) != 0;
    const T hy = T(a.ys[y] - a.vm.hq_c[1]), hz = T(a.zs[z] - a.vm.hq_c[2]);
    const T hr2 = hy * hy + hz * hz;
    const T tau = T(a.tau), G = T(a.vm.hq_G), GM = T(a.vm.hq_GM), ha = T(a.vm.hq_a);
    double acc[NP] = {};
    // quadrupole: u = R^T (x - c) = R[0][.] dx + (the line's R[1][.] dy + R[2][.] dz)
    T qr0[3] = {}, qw[3] = {};
    if constexpr (QG) {
        const double dy = a.ys[y] - a.vm.q_c[1], dz = a.zs[z] - a.vm.q_c[2];
        WDM_UNROLL
        for (int b = 0; b < 3; ++b) {
            qr0[b] = T(a.qR[b]);
            qw[b] = T(a.qR[3 + b] * dy + a.qR[6 + b] * dz);
        }
    }
    cx<T> sr[N1];
    WDM_UNROLL
    for (int kq = 0; kq < N1; ++kq) {
        if ((kq & 3) == 0) WDM_SCHED_FENCE();
        const int x = N2 * kq + m.j;
        const unsigned idx = rbase + unsigned(x);
        const cx<T> pn = WDM_LDG(ok, psi_n, idx);
        const double ph = adj_potential<T>(a, lv, x, idx, pn);
        const cx<T> q = adj_rot(pn, adj_phase<T>(a.tau * ph));  // psi'
        const T sv = ok ? tau * (v[kq].x * q.y - v[kq].y * q.x) : T(0);
        sr[kq] = mk(sv, T(0));
        if (hq) {
            const T dx = T(a.xs[x] - a.vm.hq_c[0]);
            const T r = adj_sqrt(dx * dx + hr2);
            const T ira = T(1) / (r + ha);
            const T dVdr = GM * ira * ira;
            acc[0] -= double(sv * G * ira);
            acc[1] += double(sv * dVdr);
            if (r > T(0)) {
                const T f = sv * dVdr / r;
                acc[2] -= double(f * dx);
                acc[3] -= double(f * hy);
                acc[4] -= double(f * hz);
            }
        }
        acc[5] += double(sv) * ph;
        if constexpr (QG) {
            // dV/dT0: 1/2 u_a^2 on the diagonal, u_a u_b off it
            const T dxq = T(a.xs[x] - a.vm.q_c[0]);
            const T u0 = fmaT(qr0[0], dxq, qw[0]), u1 = fmaT(qr0[1], dxq, qw[1]), u2 = fmaT(qr0[2], dxq, qw[2]);
            const T hs = T(0.5) * sv;
            acc[6] += double(hs * u0 * u0);
            acc[7] += double(hs * u1 * u1);
            acc[8] += double(hs * u2 * u2);
            acc[9] += double(sv * u0 * u1);
            acc[10] += double(sv * u0 * u2);
            acc[11] += double(sv * u1 * u2);
        }
    }
    // workgroup sums in a fixed order: lanes, then waves
    WDM_UNROLL
    for (int j = 0; j < NP; ++j) red[m.wv][j * kWave + m.lane] = acc[j];
    WDM_WAVE_SYNC();
    if (m.lane < NP) {
        double s = 0.0;
        for (int k = 0; k < kWave; ++k) s += red[m.wv][m.lane * kWave + k];
        wred[m.wv][m.lane] = s;
    }
    __syncthreads();
    if (int(threadIdx.x) < kAdjParams) {
        double s = 0.0;
        if (int(threadIdx.x) < NP)
            for (int w = 0; w < NW; ++w) s += wred[w][threadIdx.x];
        partials[kAdjParams * blockIdx.x + threadIdx.x] = s;
    }
    if (shat) {
        fwd2<F, -1, PoissonPolicy>(sr, X, m.j, xb, tw);
        if (ok) {
            const unsigned hbase = unsigned(lc) * unsigned(hpitch);
            WDM_UNROLL
            for (int i = 0; i < M; ++i)
                WDM_UNROLL
                for (int k2 = 0; k2 < N2; ++k2) {
                    const int k = (i * N2 + m.j) + N1 * k2;
                    if (2 * k <= N) stg(shat, hbase + unsigned(k), X[i][k2]);
                }
        }
    }
}

// In: g' in real space, u_hat (half spectrum after the Poisson y/z passes, or
// null).  Out: g_n = conj(E) g' + 2 (u + kappa s) psi_n, times the absorber
// when given, transformed along x when FWD (the next backward step) or left in
// real space (the gradient with respect to psi_0).
template <class T, class P, class F, bool FWD, int NW>
WDM_KERNEL(NW * 64)
void k_adj_xg(cx<T>* __restrict__ g, const cx<T>* __restrict__ psi_n, const cx<T>* __restrict__ uhat, int nlines,
              int hpitch, const typename TwEntry<T>::type* __restrict__ tw, AdjXArgs a) {
    constexpr int N1 = F::N1, N2 = F::N2, M = F::M, N = F::N;
    WDM_SHARED cx<T> xbuf[NW][F::XBUF];
    WDM_SHARED typename TwEntry<T>::type tw_s[F::N];
    tw = stage_table<F::N>(tw_s, tw);
    const LaneMap m = lane_map<F, NW>();
    const int l = m.w * F::LPW + m.lw;
    const bool ok = l < nlines;
    const int lc = ok ? l : 0;
    const int y = lc % a.Ny, z = lc / a.Ny;
    const unsigned rbase = unsigned(lc) * unsigned(N);
    cx<T>* xb = &xbuf[m.wv][m.lw * F::XS];
    const cx<T> zero = mk(T(0), T(0));
    cx<T> X[M][N2];
    cx<T> v[N1];
    T u[N1];
    if (uhat) {
        const unsigned hbase = unsigned(lc) * unsigned(hpitch);
        hermitian_row<F>(uhat, hbase, ok, m, &xbuf[m.wv][0], X);
        inv2<F, +1, PoissonPolicy>(X, v, m.j, xb, tw);
        WDM_UNROLL
        for (int kq = 0; kq < N1; ++kq) u[kq] = v[kq].x;
    } else {
        WDM_UNROLL
        for (int kq = 0; kq < N1; ++kq) u[kq] = T(0);
    }
    const LineV lv = line_V(a.vm, a.ys[y], a.zs[z]);
    const bool kap = (a.vm.flags & V_KAPPA) != 0;
    const T ktau = T(a.vm.kappa * a.tau);
    WDM_SCHED_FENCE();
    WDM_UNROLL
    for (int kq = 0; kq < N1; ++kq) v[kq] = WDM_LDG(ok, g, rbase + unsigned(N2 * kq + m.j));
    WDM_UNROLL
    for (int kq = 0; kq < N1; ++kq) {
        if ((kq & 3) == 0) WDM_SCHED_FENCE();
        const int x = N2 * kq + m.j;
        const unsigned idx = rbase + unsigned(x);
        const cx<T> pn = WDM_LDG(ok, psi_n, idx);
        const AdjPhase<T> e = adj_phase<T>(a.tau * adj_potential<T>(a, lv, x, idx, pn));
        T w = u[kq];
        if (kap) {
            const cx<T> q = adj_rot(pn, e);  // psi'
            w += ktau * (v[kq].x * q.y - v[kq].y * q.x);
        }
        const cx<T> cg = adj_rot_conj(v[kq], e);  // conj(E) g'
        T rr = fmaT(T(2) * w, pn.x, cg.x);
        T ri = fmaT(T(2) * w, pn.y, cg.y);
        if (a.border) {
            const T b = T(a.border[x]);
            rr *= b;
            ri *= b;
        }
        v[kq] = mk(rr, ri);
    }
    if constexpr (FWD) {
        fwd2<F, -1, P>(v, X, m.j, xb, tw);
        if (ok) {
            WDM_UNROLL
            for (int i = 0; i < M; ++i)
                WDM_UNROLL
                for (int k2 = 0; k2 < N2; ++k2) stg(g, rbase + unsigned((i * N2 + m.j) + N1 * k2), X[i][k2]);
        }
    } else {
        if (ok) {
            WDM_UNROLL
            for (int kq = 0; kq < N1; ++kq) stg(g, rbase + unsigned(N2 * kq + m.j), v[kq]);
        }
    }
}

}  // namespace wdm
