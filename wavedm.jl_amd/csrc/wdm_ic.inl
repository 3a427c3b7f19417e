// Initial condition from a density and a velocity field (included at the end
// of wdm_plan.cpp).
//
// src/physics/schrodinger.jl:8-30 (solve_vector_equation): theta is the sum
// of three line integrals, each along its own axis from the first cell of that
// axis, with the cumulative trapezoid rule on x = (0:N-1) dx;
// src/ic/generation.jl:330-336 removes the mass-weighted mean velocity first
// (optional here); src/ic/generation.jl:185-187 forms
// psi = sqrt(rho) exp(i theta / (hbar/m)) (the reference's units have
// hbar/m = 1).  One thread per line walks the line in order, so every partial
// sum is formed exactly as the reference's loop forms it; the three axis
// passes add into theta in the reference's order (x, then y, then z).

namespace {

constexpr int kIcNV = 4;  // sum rho, sum rho vx, sum rho vy, sum rho vz

WDM_KERNEL(256)
void k_ic_moments(const double* __restrict__ rho, const double* __restrict__ vx, const double* __restrict__ vy,
                  const double* __restrict__ vz, int nx, int ny, int nz, double* __restrict__ partials) {
    WDM_SHARED double sh[kIcNV][256];
    double acc[kIcNV] = {0, 0, 0, 0};
    WDM_FOR_CELLS(nx, ny, nz) {
        (void)x;
        (void)y;
        (void)z;
        const double r = rho[i];
        acc[0] += r;
        acc[1] += r * vx[i];
        acc[2] += r * vy[i];
        acc[3] += r * vz[i];
    }
    block_sums<kIcNV>(acc, sh, partials);
}

// theta (+)= 1/2 sum_{l<=i} (x_l - x_{l-1}) ((v_l - v0) + (v_{l-1} - v0)) along
// axis `axis`, one thread per line; first = 1 initialises theta
WDM_KERNEL(256)
void k_ic_scan(const double* __restrict__ v, double* __restrict__ theta, int nx, int ny, int nz, int axis, double d,
               double v0, int first) {
    const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
    const unsigned lines = axis == 0 ? unsigned(ny) * unsigned(nz)
                           : axis == 1 ? unsigned(nx) * unsigned(nz)
                                       : unsigned(nx) * unsigned(ny);
    const unsigned stride = axis == 0 ? 1u : axis == 1 ? unsigned(nx) : unsigned(nx) * unsigned(ny);
    for (unsigned line = blockIdx.x * 256u + threadIdx.x; line < lines; line += gridDim.x * 256u) {
        unsigned base;
        if (axis == 0) base = line * unsigned(nx);
        else if (axis == 1) base = (line % unsigned(nx)) + (line / unsigned(nx)) * unsigned(nx) * unsigned(ny);
        else base = line;
        double acc = 0.0;
        double prev = v[base] - v0;
        if (first) theta[base] = 0.0;
        for (int l = 1; l < n; ++l) {
            const unsigned idx = base + unsigned(l) * stride;
            const double cur = v[idx] - v0;
            acc += (double(l) * d - double(l - 1) * d) * (cur + prev);
            prev = cur;
            if (first) theta[idx] = 0.5 * acc;
            else theta[idx] += 0.5 * acc;
        }
    }
}

template <class T>
WDM_KERNEL(256)
void k_ic_assemble(const double* __restrict__ rho, const double* __restrict__ theta, cx<T>* __restrict__ psi,
                   size_t n, double inv_hm) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
        int q;
        const double r = reduce_quadrant(-theta[i] * inv_hm, q);  // exp(i th) = exp(-i (-th))
        double s, c;
        sincos_small(r, s, c);
        const cx<double> e = rot_quadrant(mk(c, -s), q);
        const double a = __builtin_sqrt(rho[i]);
        psi[i] = mk(T(a * e.x), T(a * e.y));
    }
}

template <class T>
void ic_velocity(wdm_plan* p, const double* rho, const double* const v[3], int remove_bulk) {
    const size_t n = p->ncell;
    DBuf tmp;
    double* d = tmp.get<double>(4 * n);
    double* drho = d;
    double* dv[3] = {d + n, d + 2 * n, d + 3 * n};
    hip_check(hipMemcpyAsync(drho, rho, sizeof(double) * n, hipMemcpyHostToDevice, p->stream), "ic rho");
    for (int a = 0; a < 3; ++a)
        hip_check(hipMemcpyAsync(dv[a], v[a], sizeof(double) * n, hipMemcpyHostToDevice, p->stream), "ic v");
    double v0[3] = {0.0, 0.0, 0.0};
    if (remove_bulk) {
        const unsigned nb = adj_line_blocks(p->ny, p->nz);
        double* part = p->diag.get<double>(size_t(kIcNV) * nb);
        WDM_LAUNCH((k_ic_moments), dim3(nb), dim3(256), p->stream, (const double*)drho, (const double*)dv[0],
                   (const double*)dv[1], (const double*)dv[2], p->nx, p->ny, p->nz, part);
        hip_check(hipGetLastError(), "ic moments");
        std::vector<double> h(size_t(kIcNV) * nb);
        hip_check(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, p->stream),
                  "ic moments copy");
        hip_check(hipStreamSynchronize(p->stream), "ic moments sync");
        Neumaier S[kIcNV];
        for (unsigned b = 0; b < nb; ++b)
            for (int j = 0; j < kIcNV; ++j) S[j].add(h[size_t(kIcNV) * b + j]);
        if (!(S[0].value() > 0.0)) fail(WDM_ERR_ARG, "wdm_ic_velocity: the density sums to zero");
        for (int a = 0; a < 3; ++a) v0[a] = S[1 + a].value() / S[0].value();
    }
    double* theta = p->scratch_real.get<double>(n);
    for (int a = 0; a < 3; ++a) {
        const unsigned lines = a == 0 ? unsigned(p->ny * p->nz) : a == 1 ? unsigned(p->nx * p->nz)
                                                                         : unsigned(p->nx * p->ny);
        WDM_LAUNCH((k_ic_scan), dim3((lines + 255) / 256), dim3(256), p->stream, (const double*)dv[a], theta, p->nx,
                   p->ny, p->nz, a, p->dx[a], v0[a], a == 0 ? 1 : 0);
        hip_check(hipGetLastError(), "ic scan");
    }
    const unsigned nb = unsigned(std::min<size_t>((n + 255) / 256, 65536));
    WDM_LAUNCH((k_ic_assemble<T>), dim3(nb), dim3(256), p->stream, (const double*)drho, (const double*)theta,
               p->psi.as<cx<T>>(), n, 1.0 / p->hbar_over_m);
    hip_check(hipGetLastError(), "ic assemble");
    hip_check(hipStreamSynchronize(p->stream), "ic sync");
    ++p->psi_epoch;
    p->kspace = false;
    p->phi_valid = false;
    p->mass_valid = false;
    p->ledger_stale = true;
}

// The acceleration field of an initial condition (src/ic/generation.jl:246-273,
// periodic branch): Phi = 4 pi G lap^-1 (rho_halo - <rho_halo>) on the plan's
// lattice, a = -grad_central(Phi) (central differences, periodic on the
// faces, as wdm_acceleration); and generate_velocity_field's magnitude
// (:285-291): a_CDM = |a| on the interior cells [2:end-1]^3, 0 on the faces,
// times 0 at the cell (div(N,2), div(N,2), div(N,2)) (1-based).  The interior
// uses no face neighbour, so a_CDM does not depend on the face rule.
WDM_KERNEL(256)
void k_ic_acc_cdm(const double* __restrict__ phi, double* __restrict__ out, int nx, int ny, int nz, double hx,
                  double hy, double hz) {
    WDM_FOR_CELLS(nx, ny, nz) {
        double a = 0.0;
        if (x > 0 && x < nx - 1 && y > 0 && y < ny - 1 && z > 0 && z < nz - 1) {
            const unsigned sy = unsigned(nx), sz = unsigned(nx) * unsigned(ny);
            const double ax = -(phi[i + 1] - phi[i - 1]) * hx;
            const double ay = -(phi[i + sy] - phi[i - sy]) * hy;
            const double az = -(phi[i + sz] - phi[i - sz]) * hz;
            a = __builtin_sqrt(ax * ax + ay * ay + az * az);
        }
        if (x == nx / 2 - 1 && y == ny / 2 - 1 && z == nz / 2 - 1) a *= 0.0;
        out[i] = a;
    }
}

template <class T>
void ic_acceleration(wdm_plan* p, const double* rho, double* phi_out, double* const acc_out[3], double* a_cdm) {
    poisson_api<T>(p, rho, nullptr, p->four_pi_G);  // rhohat = Phi_hat of rho_halo
    typedef typename TwEntry<T>::type TwE;
    double* phi = p->scratch_real.get<double>(p->ncell);
    ops_of<T>(p, 0)->x_c2r_real(p->rhohat.as<cx<T>>(), phi, p->ny * p->nz, p->hpitch, p->ax[0].tw.as<TwE>(),
                                 p->stream);
    hip_check(hipGetLastError(), "C2R pass");
    if (phi_out)
        hip_check(hipMemcpyAsync(phi_out, phi, sizeof(double) * p->ncell, hipMemcpyDeviceToHost, p->stream), "phi copy");
    double* g = p->diag.get<double>(p->ncell);
    for (int a = 0; a < 3; ++a) {
        if (!acc_out[a]) continue;
        hip_check(hipStreamSynchronize(p->stream), "acceleration sync");  // g is reused per axis
        WDM_LAUNCH((k_neg_cdiff), dim3(adj_line_blocks(p->ny, p->nz)), dim3(256), p->stream, (const double*)phi, g,
                   p->nx, p->ny, p->nz, a, 0.5 / p->dx[a]);
        hip_check(hipGetLastError(), "acceleration");
        hip_check(hipMemcpyAsync(acc_out[a], g, sizeof(double) * p->ncell, hipMemcpyDeviceToHost, p->stream),
                  "acceleration copy");
    }
    if (a_cdm) {
        hip_check(hipStreamSynchronize(p->stream), "a_CDM sync");
        WDM_LAUNCH((k_ic_acc_cdm), dim3(adj_line_blocks(p->ny, p->nz)), dim3(256), p->stream, (const double*)phi, g,
                   p->nx, p->ny, p->nz, 0.5 / p->dx[0], 0.5 / p->dx[1], 0.5 / p->dx[2]);
        hip_check(hipGetLastError(), "a_CDM");
        hip_check(hipMemcpyAsync(a_cdm, g, sizeof(double) * p->ncell, hipMemcpyDeviceToHost, p->stream), "a_CDM copy");
    }
    hip_check(hipStreamSynchronize(p->stream), "ic acceleration sync");
}

}  // namespace

extern "C" {

int wdm_ic_acceleration(wdm_plan* p, const double* rho_halo, double* phi, double* ax, double* ay, double* az,
                        double* a_cdm) {
    if (!p || !rho_halo) return fail_code(p, WDM_ERR_ARG, "wdm_ic_acceleration: NULL argument");
    if (p->kspace) return fail_code(p, WDM_ERR_STATE, "wdm_ic_acceleration: psi is in k-space");
    double* const acc[3] = {ax, ay, az};
    return guarded(p, [&] { DISPATCH(p, ic_acceleration, p, rho_halo, phi, acc, a_cdm); });
}

int wdm_ic_velocity(wdm_plan* p, const double* rho, const double* vx, const double* vy, const double* vz,
                    int remove_bulk) {
    if (!p || !rho || !vx || !vy || !vz) return fail_code(p, WDM_ERR_ARG, "wdm_ic_velocity: NULL argument");
    const double* v[3] = {vx, vy, vz};
    return guarded(p, [&] { DISPATCH(p, ic_velocity, p, rho, v, remove_bulk); });
}

}  // extern "C"
