// Adjoint SSFM (included at the end of wdm_plan.cpp; shares its helpers).
//
// Forward map of one reference-split step (src/simulation_main.jl:536-547):
//   psi' = E psi_n,  E = exp(-i tau (Phi[psi_n] + V)),  tau = (m/hbar) dt_kick
//   psi_{n+1} = B U psi',  U = F^-1 P F,  P = exp(-i (hbar/m) k^2 dt / 2)
// with Phi[psi] = 4 pi G lap^-1 (|psi|^2 - <|psi|^2>) (self-adjoint, real).
//
// Gradients are packed g = dL/dRe(psi) + i dL/dIm(psi), so dL = Re sum conj(g) dpsi.
// Backward through one step, given g_{n+1}:
//   g'  = U^H (B g_{n+1})                         (conjugate drift)
//   s   = tau Im(conj(g') psi')                   (real field)
//   u   = Phi-operator(s) + kappa s               (linearised Poisson / kappa|psi|^2)
//   g_n = conj(E) g' + 2 u psi_n
//   dL/dtheta += sum s dV/dtheta                  (Hernquist M, a, centre)
//   dL/dsigma += sum s (Phi + V) + (1/N) sum_k w_k Im(conj(F g')_k (F psi')_k),
//                w_k = (hbar/m) k^2 dt / 2        (sigma scales dt and dt_kick)
// Memory: checkpoints of psi every K steps; each segment is recomputed with the
// fused forward step while psi_n and Phi_n are taped, then walked backwards.
// A backward step is eight passes like a forward step: the y/z passes of the
// conjugate drift, k_adj_xs (x inverse, s, partials, R2C of s), the Poisson
// y/z passes, k_adj_xg (C2R of u, g_n, absorber, x forward); see
// wdm_adj_kernels.hpp.
// No counterpart in the reference (it imports Zygote, src/WaveDM.jl:42, but
// defines no rule for the step).

namespace {

// ------------------------------------------------------------------ kernels
// Elementwise kernels: workgroup b takes the cells of lines [b*kLPB, (b+1)*kLPB)
// (line = y + ny*z), its threads stride over them; 32-bit index arithmetic
// only (plans are below 2^31 cells).
constexpr int kLPB = 8;
inline unsigned adj_line_blocks(int ny, int nz) { return unsigned((ny * nz + kLPB - 1) / kLPB); }
#define WDM_FOR_CELLS(nx, ny, nz)                                                                     \
    for (unsigned j_ = threadIdx.x, lines_ = unsigned((ny) * (nz)), l0_ = blockIdx.x * unsigned(kLPB); \
         j_ < unsigned(kLPB) * unsigned(nx); j_ += blockDim.x)                                         \
        if (const unsigned line_ = l0_ + j_ / unsigned(nx); line_ < lines_)                           \
            if (const int x = int(j_ % unsigned(nx)), y = int(line_ % unsigned(ny)),                  \
                z = int(line_ / unsigned(ny));                                                        \
                true)                                                                                  \
                if (const unsigned i = line_ * unsigned(nx) + unsigned(x); true)

WDM_FI cx<double> expmi(double theta) {
    int q;
    const double r = reduce_quadrant(theta, q);
    double s, c;
    sincos_small(r, s, c);
    return rot_quadrant(mk(c, -s), q);  // exp(-i theta)
}

// deterministic block sum of NV values per thread into partials[NV*block + j]
template <int NV>
WDM_FI void block_sums(double (&acc)[NV], double (*sh)[256], double* __restrict__ partials) {
    const int t = int(threadIdx.x);
    for (int j = 0; j < NV; ++j) sh[j][t] = acc[j];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int j = 0; j < NV; ++j) sh[j][t] += sh[j][t + w];
        __syncthreads();
    }
    if (t == 0)
        for (int j = 0; j < NV; ++j) partials[NV * blockIdx.x + j] = sh[j][0];
}

constexpr int kAdjNV = kAdjParams;

// g = conj(E) g' + 2 (u + kappa s) psi  (then * border along x when given);
// the element-wise form of k_adj_xg, used for B g_N before the sweep
template <class T>
WDM_KERNEL(256)
void k_adj_g(cx<T>* __restrict__ g, const cx<T>* __restrict__ psi, const double* __restrict__ phi,
             const double* __restrict__ u, const double* __restrict__ s, int nx, int ny, int nz,
             const double* __restrict__ xs, const double* __restrict__ ys, const double* __restrict__ zs, VModel vm,
             double tau, const double* __restrict__ border) {
    WDM_FOR_CELLS(nx, ny, nz) {
        const cx<T> pv = psi[i];
        const double V = vm.flags ? eval_V<T>(vm, xs[x], ys[y], zs[z], i, pv) : 0.0;
        const cx<double> e = expmi(tau * ((phi ? phi[i] : 0.0) + V));
        const double gr = double(g[i].x), gi = double(g[i].y);
        double w = u ? u[i] : 0.0;
        if (vm.flags & V_KAPPA) w += vm.kappa * s[i];
        // e = E = exp(-i theta): conj(E) g' = (e.x - i e.y)(gr + i gi)
        double rr = e.x * gr + e.y * gi + 2.0 * w * double(pv.x);
        double ri = e.x * gi - e.y * gr + 2.0 * w * double(pv.y);
        if (border) {
            rr *= border[x];
            ri *= border[x];
        }
        g[i] = mk(T(rr), T(ri));
    }
}

// acc[j] += sum_b partials[nv*b + j] for j < nv: one block of 256 threads per
// component, each thread a fixed strided subset (Neumaier), then a fixed-order
// tree; deterministic.
WDM_KERNEL(256)
void k_adj_accumulate(const double* __restrict__ partials, int nb, int nv, double* __restrict__ acc) {
    WDM_SHARED double sh[256];
    const int t = int(threadIdx.x), j = int(blockIdx.x);
    double sum = 0.0, comp = 0.0;
    for (int b = t; b < nb; b += 256) {
        const double v = partials[nv * b + j];
        const double u = sum + v;
        comp += (std::fabs(sum) >= std::fabs(v)) ? (sum - u) + v : (v - u) + sum;
        sum = u;
    }
    sh[t] = sum + comp;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) acc[j] += sh[0];
}

WDM_KERNEL(64)
void k_adj_zero(double* __restrict__ acc, int n) {
    const int j = int(threadIdx.x);
    if (j < n) acc[j] = 0.0;
}

// ------------------------------------------------------------------ host side
struct AdjState {
    bool ready = false;
    // the plan's configuration generation and psi epoch the tape belongs to:
    // the backward sweep recomputes segments with the plan's current models
    // and coefficients and ends by restoring psi_N, so either changing in
    // between would give the gradient of another trajectory
    unsigned gen = 0;
    unsigned long long epoch = 0;
    int nsteps = 0, K = 0, nckpt = 0;
    double dt = 0.0, dt_kick = 0.0, t0 = 0.0;
    DBuf ckpt;   // (nckpt + 1) psi: psi_{kK}, k < nckpt, then psi_N
    DBuf tpsi;   // K psi of the current segment
    DBuf tphi;   // K real Phi (FP64) of the current segment
    DBuf tspec;  // K 3-D spectra F(psi') of the current segment (dt-scale term)
    DBuf g;      // adjoint field
    DBuf part, acc;
    DBuf kfull[3];
};

void adj_release(wdm_plan* p) {
    if (!p->adj) return;
    AdjState& A = *p->adj;
    for (DBuf* b : {&A.ckpt, &A.tpsi, &A.tphi, &A.tspec, &A.g, &A.part, &A.acc, &A.kfull[0], &A.kfull[1],
                    &A.kfull[2]})
        b->reset();
    delete p->adj;
    p->adj = nullptr;
}

AdjState& adj_of(wdm_plan* p) {
    if (!p->adj) p->adj = new AdjState();
    return *p->adj;
}



template <class T>
void adj_forward(wdm_plan* p, int nsteps, double dt, double dt_kick, double t0, int K) {
    if (p->kspace) fail(WDM_ERR_STATE, "wdm_adjoint_forward: psi is in k-space");
    if (p->ledger) fail(WDM_ERR_UNSUPPORTED, "wdm_adjoint_forward: the mass projection is not differentiated");
    if (p->vflags & V_TIDAL)
        fail(WDM_ERR_UNSUPPORTED, "wdm_adjoint_forward: the tidal potential (centred on the density peak) is not "
                                  "differentiated");
    AdjState& A = adj_of(p);
    A.ready = false;
    const size_t n = p->ncell;
    A.nsteps = nsteps;
    A.K = K;
    A.nckpt = (nsteps + K - 1) / K;
    A.dt = dt;
    A.dt_kick = dt_kick;
    A.t0 = t0;
    cx<T>* ck = A.ckpt.get<cx<T>>(n * size_t(A.nckpt + 1));
    for (int k = 0; k < A.nckpt; ++k) {
        hip_check(hipMemcpyAsync(ck + n * size_t(k), p->psi.as<cx<T>>(), sizeof(cx<T>) * n, hipMemcpyDeviceToDevice,
                                 p->stream),
                  "checkpoint");
        const int len = std::min(K, nsteps - k * K);
        do_step<T>(p, len, dt, dt_kick, t0 + double(k) * K * dt, WDM_SPLIT_REFERENCE);
    }
    hip_check(hipMemcpyAsync(ck + n * size_t(A.nckpt), p->psi.as<cx<T>>(), sizeof(cx<T>) * n,
                             hipMemcpyDeviceToDevice, p->stream),
              "final state");
    hip_check(hipStreamSynchronize(p->stream), "adjoint forward");
    A.gen = p->gen;
    A.epoch = p->psi_epoch;
    A.ready = true;
}

// g (device, transformed along x) <- the y and z passes of the conjugate drift
// U^H g (the x passes are fused into k_adj_xg / k_adj_xs); with Psi given,
// the z pass also sums the dt-scale spectral term, one partial per wave of
// the pass into part (adj_spec_waves of them)
inline int adj_spec_waves(const wdm_plan* p, int lpw_z) { return (p->nx + lpw_z - 1) / lpw_z * p->ny; }
template <class T>
void adj_conj_drift_yz(wdm_plan* p, cx<T>* g, const cx<T>* Psi, double* part) {
    typedef typename TwEntry<T>::type TwE;
    AdjState& A = adj_of(p);
    const AxisOps<T>* ox = ops_of<T>(p, 0);
    const AxisOps<T>* oy = ops_of<T>(p, 1);
    const AxisOps<T>* oz = ops_of<T>(p, 2);
    KTabs none = {nullptr, nullptr, nullptr, 0.0, nullptr};
    (void)ox;
    if (p->ny > 1)
        oy->col(COL_FWD, KF_NONE, g, p->nx, p->nz, unsigned(p->nx), unsigned(p->nx * p->ny), p->ax[1].tw.as<TwE>(),
                none, p->stream);
    ensure_drift(p, A.dt);
    KTabs kt = {p->ax[0].drift_conj.p, p->ax[1].drift_conj.p, p->ax[2].drift_conj.p, 1.0 / double(p->ncell),
                const_cast<cx<T>*>(Psi), nullptr};
    if (Psi) {
        for (int a = 0; a < 3; ++a) kt.k2[a] = A.kfull[a].as<double>();
        kt.wc = 0.5 * p->hbar_over_m * A.dt;
        kt.partials = part;
    }
    oz->col(Psi ? COL_FMIA : COL_FMI, KF_DRIFT, g, p->nx, p->ny, unsigned(p->nx * p->ny), unsigned(p->nx),
            p->ax[2].tw.as<TwE>(), kt, p->stream);
    if (p->ny > 1)
        oy->col(COL_INV, KF_NONE, g, p->nx, p->nz, unsigned(p->nx), unsigned(p->nx * p->ny), p->ax[1].tw.as<TwE>(),
                none, p->stream);
    hip_check(hipGetLastError(), "conjugate drift");
}

template <class T>
void adj_backward(wdm_plan* p, const void* grad, int dtype, void* grad0, double* gparams, int nparams,
                  unsigned flags) {
    typedef typename TwEntry<T>::type TwE;
    AdjState& A = adj_of(p);
    if (!A.ready) fail(WDM_ERR_STATE, "wdm_adjoint_backward: no taped forward run (call wdm_adjoint_forward)");
    if (A.gen != p->gen)
        fail(WDM_ERR_STATE, "wdm_adjoint_backward: the plan's potential models, coefficients, border, gravity or "
                            "mass projection changed since wdm_adjoint_forward (tape no longer matches)");
    if (A.epoch != p->psi_epoch)
        fail(WDM_ERR_STATE, "wdm_adjoint_backward: psi was set or stepped since wdm_adjoint_forward (tape no "
                            "longer matches)");
    const bool want_dt = (flags & WDM_ADJ_DT_SCALE) != 0;
    if (want_dt && (p->vflags & V_QUADRUPOLE) && p->quad[9] != 0.0)
        fail(WDM_ERR_UNSUPPORTED, "wdm_adjoint_backward: dt-scale gradient needs a time-independent potential");
    const size_t n = p->ncell;
    const AxisOps<T>* ox = ops_of<T>(p, 0);
    const int nlines = p->ny * p->nz;
    const unsigned nb = adj_line_blocks(p->ny, p->nz);
    cx<T>* ck = A.ckpt.as<cx<T>>();
    cx<T>* tpsi = A.tpsi.get<cx<T>>(n * size_t(A.K));
    double* tphi = p->self_gravity ? A.tphi.get<double>(n * size_t(A.K)) : nullptr;
    cx<T>* g = A.g.get<cx<T>>(n);
    // workgroup partials of every step, summed once at the end in a fixed order
    // (the result does not depend on the checkpoint interval):
    // [nsteps][nbx][kAdjNV] parameter sums of k_adj_xs, then [nsteps][nbs]
    // dt-scale spectral sums of the conjugate drift's z pass (one per wave)
    const unsigned nbx = unsigned(((nlines + ox->lpw - 1) / ox->lpw + ox->nw - 1) / ox->nw);
    const unsigned nbs = unsigned(adj_spec_waves(p, ops_of<T>(p, 2)->lpw));
    double* part = A.part.get<double>(size_t(A.nsteps) * (size_t(kAdjNV) * nbx + nbs));
    double* part_spec = part + size_t(A.nsteps) * kAdjNV * nbx;
    double* acc = A.acc.get<double>(kAdjNV);
    cx<T>* tspec = want_dt ? A.tspec.get<cx<T>>(n * size_t(A.K)) : nullptr;
    if (want_dt) {
        for (int a = 0; a < 3; ++a) {
            const int na = p->ax[a].n;
            std::vector<double> h(na);
            for (int m = 0; m < na; ++m) {
                const double k = drift_k(na, p->X[a], m);
                h[m] = k * k;
            }
            upload(A.kfull[a], h);
        }
    }
    WDM_LAUNCH((k_adj_zero), dim3(1), dim3(64), p->stream, acc, kAdjNV);
    // g_N (then * border: the first backward step starts from B g_N)
    set_psi_dev<T>(p, grad, dtype, true);  // borrow psi as the upload target
    hip_check(hipMemcpyAsync(g, p->psi.as<cx<T>>(), sizeof(cx<T>) * n, hipMemcpyDeviceToDevice, p->stream), "g copy");
    const double* border = p->border_on ? p->border.as<double>() : nullptr;
    const double* xs = p->ax[0].coord.as<double>();
    const double* ys = p->ax[1].coord.as<double>();
    const double* zs = p->ax[2].coord.as<double>();
    if (border) {
        // B g_N: reuse k_adj_g with E = 1, u = 0 (tau = 0, no potential)
        VModel none;
        std::memset(&none, 0, sizeof(none));
        WDM_LAUNCH((k_adj_g<T>), dim3(nb), dim3(256), p->stream, g, (const cx<T>*)p->psi.as<cx<T>>(),
                   (const double*)nullptr, (const double*)nullptr, (const double*)nullptr, p->nx, p->ny, p->nz, xs,
                   ys, zs, none, 0.0, border);
    }
    ox->x_fft(-1, g, nlines, p->ax[0].tw.as<TwE>(), p->stream);  // the backward steps take g along x transformed
    const double tau = A.dt_kick / p->hbar_over_m;
    for (int k = A.nckpt - 1; k >= 0; --k) {
        const int n0 = k * A.K;
        const int len = std::min(A.K, A.nsteps - n0);
        // recompute the segment, taping psi_n and Phi_n
        hip_check(hipMemcpyAsync(p->psi.as<cx<T>>(), ck + n * size_t(k), sizeof(cx<T>) * n, hipMemcpyDeviceToDevice,
                                 p->stream),
                  "restore checkpoint");
        p->kspace = false;
        p->phi_valid = false;
        p->mass_valid = false;
        for (int i = 0; i < len; ++i) {
            ensure_phi<T>(p);
            hip_check(hipMemcpyAsync(tpsi + n * size_t(i), p->psi.as<cx<T>>(), sizeof(cx<T>) * n,
                                     hipMemcpyDeviceToDevice, p->stream),
                      "tape psi");
            if (tphi)
                ox->x_c2r_real(p->rhohat.as<cx<T>>(), tphi + n * size_t(i), nlines, p->hpitch, p->ax[0].tw.as<TwE>(),
                               p->stream);
            // the step's fused z pass also tapes F(psi') for the dt-scale term;
            // the last step of the segment only runs for that spectrum
            if (i + 1 < len || tspec) {
                p->zspec_out = tspec ? static_cast<void*>(tspec + n * size_t(i)) : nullptr;
                do_step<T>(p, 1, A.dt, A.dt_kick, A.t0 + double(n0 + i) * A.dt, WDM_SPLIT_REFERENCE);
                p->zspec_out = nullptr;
            }
        }
        // walk the segment backwards
        for (int i = len - 1; i >= 0; --i) {
            const int step = n0 + i;
            const cx<T>* pn = tpsi + n * size_t(i);
            const double* fn = tphi ? tphi + n * size_t(i) : nullptr;
            AdjXArgs xa;
            xa.tau = tau;
            xa.Ny = p->ny;
            xa.xs = xs;
            xa.ys = ys;
            xa.zs = zs;
            xa.phi = fn;
            xa.border = nullptr;
            xa.vm = make_vmodel(p, A.t0 + double(step) * A.dt);
            quad_rotation(p, A.t0 + double(step) * A.dt, xa.qR);
            adj_conj_drift_yz<T>(p, g, tspec ? tspec + n * size_t(i) : nullptr, part_spec + size_t(step) * nbs);
            cx<T>* rh = p->self_gravity ? p->rhohat.as<cx<T>>() : nullptr;
            ox->adj_xs(g, pn, rh, nlines, p->hpitch, p->ax[0].tw.as<TwE>(), xa,
                       part + size_t(step) * kAdjNV * nbx, p->stream);
            if (rh) poisson_chain<T>(p, p->four_pi_G);
            xa.border = step > 0 ? border : nullptr;
            ox->adj_xg(step > 0, g, pn, rh, nlines, p->hpitch, p->ax[0].tw.as<TwE>(), xa, p->stream);
            hip_check(hipGetLastError(), "adjoint step");
        }
    }
    WDM_LAUNCH((k_adj_accumulate), dim3(kAdjNV), dim3(256), p->stream, (const double*)part,
               int(size_t(A.nsteps) * nbx), kAdjNV, acc);
    if (want_dt)
        WDM_LAUNCH((k_adj_accumulate), dim3(1), dim3(256), p->stream, (const double*)part_spec, int(A.nsteps * nbs),
                   1, acc + 5);
    // outputs; restore psi_N
    double h[kAdjNV];
    hip_check(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, p->stream), "param grads");
    hip_check(hipMemcpyAsync(p->psi.as<cx<T>>(), ck + n * size_t(A.nckpt), sizeof(cx<T>) * n, hipMemcpyDeviceToDevice,
                             p->stream),
              "restore final state");
    hip_check(hipStreamSynchronize(p->stream), "adjoint backward");
    p->kspace = false;
    p->phi_valid = false;
    p->mass_valid = false;
    A.epoch = p->psi_epoch;  // psi_N restored: the tape still matches (a second backward may follow)
    if (grad0) {
        // download g_0 through the same conversion as wdm_get_psi
        cx<T>* keep = p->psi.as<cx<T>>();
        p->psi.p = g;
        get_psi_dev<T>(p, grad0, dtype, true);
        p->psi.p = keep;
    }
    if (gparams) {
        for (int j = 0; j < nparams && j < kAdjNV; ++j) gparams[j] = h[j];
        if (nparams > 5 && !want_dt) gparams[5] = NAN;
        for (int j = kAdjNV; j < nparams; ++j) gparams[j] = NAN;
    }
}

}  // namespace

extern "C" {

int wdm_adjoint_forward(wdm_plan* p, int nsteps, double dt, double dt_kick, double t0, int ckpt_every) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (nsteps < 1 || ckpt_every < 1 || !std::isfinite(dt) || !std::isfinite(dt_kick) || !std::isfinite(t0))
        return fail_code(p, WDM_ERR_ARG, "wdm_adjoint_forward: bad nsteps, ckpt_every or time arguments");
    return guarded(p, [&] { DISPATCH(p, adj_forward, p, nsteps, dt, dt_kick, t0, ckpt_every); });
}

int wdm_adjoint_backward(wdm_plan* p, const void* grad_psi, int dtype, void* grad_psi0, double* grad_params,
                         int nparams, unsigned flags) {
    if (!p || !grad_psi) return fail_code(p, WDM_ERR_ARG, "wdm_adjoint_backward: NULL argument");
    if (dtype != WDM_C128 && dtype != WDM_C64) return fail_code(p, WDM_ERR_ARG, "wdm_adjoint_backward: bad dtype");
    if (nparams < 0 || (nparams > 0 && !grad_params))
        return fail_code(p, WDM_ERR_ARG, "wdm_adjoint_backward: bad grad_params");
    return guarded(p, [&] { DISPATCH(p, adj_backward, p, grad_psi, dtype, grad_psi0, grad_params, nparams, flags); });
}

}  // extern "C"
