// Per-step diagnostics of the reference loop on the device (included at the end
// of wdm_plan.cpp, after the adjoint, whose block sums it shares).
//
// src/simulation_main.jl:550-558,641-656 and src/visualization/runtime.jl:210-238
// evaluate, on the host copy of psi after every step:
//   rho = |psi|^2;  rho_max, id = findmax(rho)          first maximum, x-fastest
//   TotalMass = sum(rho) dV
//   R  = std(r[:, :, nz/2], aweights(rho[:, :, nz/2]))  r = |x|; StatsBase's
//        default for analytic weights is the uncorrected sum w (r - mean)^2 / sum w
//   R1..R9 = quantile(|x - x_id|, weights(rho), 0.1:0.1:0.9)
//   rho v = Im(conj(psi) grad psi) (grad_central);  Momentum = sum(rho v) dV
//   PE_abs = -1/2 sum_int rho Phi_all dV,  KE = 1/2 sum_int |rho v|^2 / rho dV,
//   QE = 1/2 sum_int |grad |psi||^2 dV,    Virial = 2 KE + 2 QE - PE_abs
// sum_int: cells 1..n-2 of every axis (0-based), where the central difference
// needs no boundary rule.  The momentum sums every cell; on the faces it uses
// the periodic neighbour (grad_central comes from PhysicalFDM, whose boundary
// rule is not part of the WaveDM tree).
//
// Layout: one pass over psi (7-point stencil) for the sums and the peak, one
// pass for the radial histogram about the peak.  Both walk whole x lines, one
// line per wave, blocks striding over lines, so every load is a contiguous run
// of the line.  All sums are FP64 in a fixed order (deterministic); the
// histogram weights are 64-bit fixed point, so its integer atomics are
// order-independent too.
//
// Radii: h_p = p (W - w_1) + w_1 with w_1 = rho_max (the peak is the only cell
// at distance 0), the quantile is the first bin whose cumulative weight
// exceeds h_p, interpolated linearly inside it.  The sorted weighted quantile
// lies between the two cell distances that bracket h_p, and the one above it is
// in that bin, so the two differ by at most one bin width plus the gap between
// those distances.

namespace {

constexpr int kDiagNV = 10;  // mass, momentum x3, rho Phi, KE, QE, plane w, w r, w r^2
constexpr int kDiagBins = WDM_DIAG_BINS;
constexpr int kDiagWaves = 4;  // waves per block (one x line each)

inline unsigned diag_blocks(int ny, int nz) {
    const unsigned lines = unsigned(ny) * unsigned(nz);
    return std::max(1u, std::min((lines + kDiagWaves - 1) / kDiagWaves, 1024u));
}

// sqrt(r) and 1/r for finite r > 0 to a few 1e-15 relative: an FP32 hardware
// seed (about 1 ulp) and one FP64 Newton step, a handful of FP64 operations
// instead of the correctly rounded sequences.  Values outside the FP32 range
// (and 0) take the IEEE operation.
WDM_FI float diag_rsqf(float x) {
#ifdef WDM_GPU
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / __builtin_sqrtf(x);
#endif
}
WDM_FI double diag_sqrt(double r) {
    const float rf = float(r);
    if (!(rf > 1.0e-37f) || !(rf < 1.0e37f)) return __builtin_sqrt(r);
    const double y = double(diag_rsqf(rf));
    const double sq = r * y;
    return fmaT(0.5 * y, fmaT(-sq, sq, r), sq);
}
WDM_FI double diag_recip(double r) {
    const float rf = float(r);
    if (!(rf > 1.0e-37f) || !(rf < 1.0e37f)) return 1.0 / r;
    const double y = double(rcpf(rf));
    return fmaT(y, fmaT(-r, y, 1.0), y);
}

template <class T>
WDM_FI double rho_of(cx<T> v) {
    const double a = double(v.x), b = double(v.y);
    return a * a + b * b;
}

// Pass 1: sums and the first maximum.  partials[kDiagNV * b + j]; pmax[2b] =
// block maximum of rho, pmax[2b + 1] = its smallest cell index.  phi (may be
// NULL) is the self-gravity potential; vm (kappa term removed) adds V_ext.
// PE_abs needs Phi_all only at the cell itself, so V_ext is evaluated here
// instead of being added to a stored field.
// four blocks per CU (<= 128 registers): the grid of diag_blocks (at most
// 1024 blocks) then runs in one round on 256 CUs instead of 1.33
template <class T>
WDM_KERNEL_W(256, 4)
void k_diag_cells(const cx<T>* __restrict__ psi, const double* __restrict__ phi, int nx, int ny, int nz,
                  double hx, double hy, double hz, const double* __restrict__ xs, const double* __restrict__ ys,
                  const double* __restrict__ zs, VModel vm, int kmid, double* __restrict__ partials,
                  double* __restrict__ pmax) {
    WDM_SHARED double sh[kDiagNV][256];
    WDM_SHARED double smx[256];
    WDM_SHARED unsigned sid[256];
    const int t = int(threadIdx.x);
    const int lane = t & (kWave - 1);
    double acc[kDiagNV];
    for (int j = 0; j < kDiagNV; ++j) acc[j] = 0.0;
    double mx = -1.0;
    unsigned id = 0xffffffffu;
    const unsigned lines = unsigned(ny) * unsigned(nz);
    const unsigned sy = unsigned(nx), sz = unsigned(nx) * unsigned(ny);
    for (unsigned line = blockIdx.x * kDiagWaves + unsigned(t / kWave); line < lines;
         line += gridDim.x * kDiagWaves) {
        const int y = int(line % unsigned(ny)), z = int(line / unsigned(ny));
        const unsigned base = line * unsigned(nx);
        const unsigned oym = y == 0 ? sz : 0u, oyp = y == ny - 1 ? sz : 0u;  // periodic wraps
        const unsigned ozm = z == 0 ? unsigned(nz) * sz : 0u, ozp = z == nz - 1 ? unsigned(nz) * sz : 0u;
        const bool line_in = y > 0 && y < ny - 1 && z > 0 && z < nz - 1;
        const bool plane = z == kmid;
        for (int x = lane; x < nx; x += kWave) {
            const unsigned i = base + unsigned(x);
            const cx<T> c = psi[i];
            const double cr = double(c.x), ci = double(c.y);
            const double rho = cr * cr + ci * ci;
            acc[0] += rho;
            if (rho > mx) {  // indices rise along a thread's walk: keeps the first
                mx = rho;
                id = i;
            }
            const cx<T> n[6] = {psi[x == 0 ? i + unsigned(nx - 1) : i - 1u],
                                psi[x == nx - 1 ? i + 1u - unsigned(nx) : i + 1u],
                                psi[i + oym - sy], psi[i - oyp + sy],
                                psi[i + ozm - sz], psi[i - ozp + sz]};
            const double h[3] = {hx, hy, hz};
            double rv[3];
            WDM_UNROLL
            for (int d = 0; d < 3; ++d) {
                const double gr = (double(n[2 * d + 1].x) - double(n[2 * d].x)) * h[d];
                const double gi = (double(n[2 * d + 1].y) - double(n[2 * d].y)) * h[d];
                rv[d] = cr * gi - ci * gr;
                acc[1 + d] += rv[d];
            }
            if (line_in && x > 0 && x < nx - 1) {
                if (phi || vm.flags) {  // Phi_all = Phi + V_ext
                    const double v = vm.flags ? eval_V<T>(vm, xs[x], ys[y], zs[z], i, c) : 0.0;
                    acc[4] += rho * ((phi ? phi[i] : 0.0) + v);
                }
                acc[5] += (rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]) * diag_recip(rho);
                // (|psi+| - |psi-|)^2 = rho+ + rho- - 2 sqrt(rho+ rho-): one
                // root per axis instead of two.  The cancellation costs about
                // 1e-16 rho / (|psi+| - |psi-|)^2 relative, negligible for the
                // cells that carry the sum.
                double q = 0.0;
                WDM_UNROLL
                for (int d = 0; d < 3; ++d) {
                    const double rp = rho_of(n[2 * d + 1]), rm = rho_of(n[2 * d]);
                    const double dd = (rp + rm) - 2.0 * diag_sqrt(rp * rm);
                    q += dd * (h[d] * h[d]);
                }
                acc[6] += q;
            }
            if (plane) {
                const double r = __builtin_sqrt(xs[x] * xs[x] + ys[y] * ys[y] + zs[z] * zs[z]);
                acc[7] += rho;
                acc[8] += rho * r;
                acc[9] += rho * r * r;
            }
        }
    }
    block_sums<kDiagNV>(acc, sh, partials);
    smx[t] = mx;
    sid[t] = id;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            const double o = smx[t + w];
            if (o > smx[t] || (o == smx[t] && sid[t + w] < sid[t])) {
                smx[t] = o;
                sid[t] = sid[t + w];
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        pmax[2 * blockIdx.x] = smx[0];
        pmax[2 * blockIdx.x + 1] = double(sid[0]);
    }
}

// Pass 2: radial histogram of |x - c| weighted by round(rho * qscale), one
// LDS histogram per block, written out whole (hpart[kDiagBins * b + bin]).
template <class T>
WDM_KERNEL(256)
void k_diag_hist(const cx<T>* __restrict__ psi, int nx, int ny, int nz, const double* __restrict__ xs,
                 const double* __restrict__ ys, const double* __restrict__ zs, double c0, double c1, double c2,
                 double inv_width, double qscale, unsigned long long* __restrict__ hpart) {
    WDM_SHARED unsigned long long hs[kDiagBins];
    const int t = int(threadIdx.x);
    const int lane = t & (kWave - 1);
    for (int b = t; b < kDiagBins; b += 256) hs[b] = 0ull;
    __syncthreads();
    const unsigned lines = unsigned(ny) * unsigned(nz);
    for (unsigned line = blockIdx.x * kDiagWaves + unsigned(t / kWave); line < lines;
         line += gridDim.x * kDiagWaves) {
        const int y = int(line % unsigned(ny)), z = int(line / unsigned(ny));
        const double dy = ys[y] - c1, dz = zs[z] - c2;
        const double r2yz = dy * dy + dz * dz;
        const unsigned base = line * unsigned(nx);
        // lanes take cells nx/64 apart (when nx is a multiple of 64): cells
        // next to each other along x often fall in one bin, and 64 lanes
        // adding to one LDS counter serialise
        const int chunk = nx % kWave == 0 ? nx / kWave : 0;
        for (int xi = lane; xi < nx; xi += kWave) {
            const int x = chunk ? (xi % kWave) * chunk + xi / kWave : xi;
            const double rho = rho_of(psi[base + unsigned(x)]);
            if (rho > 0.0) {  // zero weights take no part (StatsBase drops them)
                const double dx = xs[x] - c0;
                const double r = diag_sqrt(dx * dx + r2yz);
                const double fb = r * inv_width;
                const int bin = fb < double(kDiagBins - 1) ? int(fb) : kDiagBins - 1;
                atomicAdd(&hs[bin], (unsigned long long)__builtin_rint(rho * qscale));
            }
        }
    }
    __syncthreads();
    unsigned long long* out = hpart + size_t(kDiagBins) * blockIdx.x;
    for (int b = t; b < kDiagBins; b += 256) out[b] = hs[b];
}

// out[kDiagBins * blockIdx.y + bin] = sum of the rows [rows_per * blockIdx.y,
// +rows_per) of hpart (nrows rows of kDiagBins), one thread per bin.  Run
// twice: per slice of the block histograms, then over the slices.  Integer
// sums: the result does not depend on the grouping.
constexpr int kDiagSlices = 32;
WDM_KERNEL(256)
void k_diag_hist_sum(const unsigned long long* __restrict__ hpart, int nrows, int rows_per,
                     unsigned long long* __restrict__ out) {
    const int bin = int(blockIdx.x * 256 + threadIdx.x);
    if (bin >= kDiagBins) return;
    const int r0 = int(blockIdx.y) * rows_per;
    const int r1 = r0 + rows_per < nrows ? r0 + rows_per : nrows;
    unsigned long long s = 0ull;
    for (int r = r0; r < r1; ++r) s += hpart[size_t(kDiagBins) * r + bin];
    out[size_t(kDiagBins) * blockIdx.y + bin] = s;
}

WDM_KERNEL(256)
void k_fill_zero(double* __restrict__ a, int nx, int ny, int nz) {
    WDM_FOR_CELLS(nx, ny, nz) {
        (void)y;
        (void)z;
        a[i] = 0.0;
    }
}

// phi += V_ext (static, Hernquist, quadrupole; the kappa term is not part of Phi_all)
template <class T>
WDM_KERNEL(256)
void k_add_vext(double* __restrict__ phi, int nx, int ny, int nz, const double* __restrict__ xs,
                const double* __restrict__ ys, const double* __restrict__ zs, VModel vm) {
    WDM_FOR_CELLS(nx, ny, nz) phi[i] += eval_V<T>(vm, xs[x], ys[y], zs[z], i, mk(T(0), T(0)));
}

// out = -(f[i+1] - f[i-1]) * h along one axis (stride s, length n), periodic on the faces
WDM_KERNEL(256)
void k_neg_cdiff(const double* __restrict__ f, double* __restrict__ out, int nx, int ny, int nz, int axis,
                 double h) {
    WDM_FOR_CELLS(nx, ny, nz) {
        const int c = axis == 0 ? x : axis == 1 ? y : z;
        const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
        const unsigned s = axis == 0 ? 1u : axis == 1 ? unsigned(nx) : unsigned(nx) * unsigned(ny);
        const unsigned ip = c == n - 1 ? i - unsigned(n - 1) * s : i + s;
        const unsigned im = c == 0 ? i + unsigned(n - 1) * s : i - s;
        out[i] = -(f[ip] - f[im]) * h;
    }
}

double coord_at(const wdm_plan* p, int a, int i) {
    const int n = p->ax[a].n;
    return (p->flags & WDM_GRID_CONSISTENT) ? -p->X[a] + i * (2.0 * p->X[a] / n)
                                            : linrange_at(-p->X[a], p->X[a], n, i);
}

// Phi_all of the current psi into dst (FP64, device)
template <class T>
void phi_all(wdm_plan* p, double* dst) {
    phi_real<T>(p, dst);
    VModel vm = make_vmodel(p, p->last_kick_t);
    vm.flags &= ~V_KAPPA;
    if (vm.flags) {
        WDM_LAUNCH((k_add_vext<T>), dim3(adj_line_blocks(p->ny, p->nz)), dim3(256), p->stream, dst, p->nx, p->ny,
                   p->nz, p->ax[0].coord.as<double>(), p->ax[1].coord.as<double>(), p->ax[2].coord.as<double>(),
                   vm);
        hip_check(hipGetLastError(), "Phi_all");
    }
}

struct Neumaier {
    double s = 0.0, c = 0.0;
    void add(double v) {
        const double t = s + v;
        c += std::fabs(s) >= std::fabs(v) ? (s - t) + v : (v - t) + s;
        s = t;
    }
    double value() const { return s + c; }
};

// Weighted quantile from the binned cumulative weight (see the file header).
double hist_quantile(const std::vector<unsigned long long>& hist, double w1, double total, double p, double width) {
    const double h = p * (total - w1) + w1;
    double cum = 0.0;
    for (int b = 0; b < kDiagBins; ++b) {
        const double nxt = cum + double(hist[b]);
        if (nxt > h) return width * (double(b) + (h - cum) / double(hist[b]));
        cum = nxt;
    }
    for (int b = kDiagBins - 1; b >= 0; --b)  // h at the total: the largest distance present
        if (hist[b]) return width * double(b + 1);
    return NAN;
}

template <class T>
void diagnostics(wdm_plan* p, double* out, int nout) {
    if (p->kspace) fail(WDM_ERR_STATE, "wdm_diagnostics: psi is in k-space");
    double r[WDM_DIAG_COUNT];
    for (double& v : r) v = 0.0;
    const bool want_phi = nout > WDM_DIAG_PE_ABS;
    double* phi = nullptr;
    VModel vm;
    std::memset(&vm, 0, sizeof(vm));
    if (want_phi) {
        if (p->self_gravity) {
            phi = p->scratch_real.get<double>(p->ncell);
            phi_real<T>(p, phi);
        }
        vm = make_vmodel(p, p->last_kick_t);
        vm.flags &= ~V_KAPPA;
    }
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const unsigned nb = diag_blocks(ny, nz);
    const size_t part_bytes = sizeof(double) * size_t(nb) * (kDiagNV + 2);
    const size_t hpart_bytes = sizeof(unsigned long long) * size_t(kDiagBins) * (nb + kDiagSlices + 1);
    char* buf = p->diag.get<char>(part_bytes + hpart_bytes);
    double* part = reinterpret_cast<double*>(buf);
    double* pmax = part + size_t(nb) * kDiagNV;
    unsigned long long* hpart = reinterpret_cast<unsigned long long*>(buf + part_bytes);
    unsigned long long* hslice = hpart + size_t(kDiagBins) * nb;
    unsigned long long* hist = hslice + size_t(kDiagBins) * kDiagSlices;
    const double* xs = p->ax[0].coord.as<double>();
    const double* ys = p->ax[1].coord.as<double>();
    const double* zs = p->ax[2].coord.as<double>();
    const int kmid = nz > 1 ? nz / 2 - 1 : 0;  // Julia's div(end, 2), 1-based
    WDM_LAUNCH((k_diag_cells<T>), dim3(nb), dim3(256), p->stream, p->psi.as<cx<T>>(), (const double*)phi, nx, ny,
               nz, 0.5 / p->dx[0], 0.5 / p->dx[1], 0.5 / p->dx[2], xs, ys, zs, vm, kmid, part, pmax);
    hip_check(hipGetLastError(), "diagnostics pass");
    std::vector<double> hp(size_t(nb) * (kDiagNV + 2));
    hip_check(hipMemcpyAsync(hp.data(), part, part_bytes, hipMemcpyDeviceToHost, p->stream), "diagnostics copy");
    hip_check(hipStreamSynchronize(p->stream), "diagnostics sync");
    Neumaier S[kDiagNV];
    double rmax = -1.0, id = 0.0;
    for (unsigned b = 0; b < nb; ++b) {
        for (int j = 0; j < kDiagNV; ++j) S[j].add(hp[size_t(kDiagNV) * b + j]);
        const double m = hp[size_t(nb) * kDiagNV + 2 * b], mi = hp[size_t(nb) * kDiagNV + 2 * b + 1];
        if (m > rmax || (m == rmax && mi < id)) {
            rmax = m;
            id = mi;
        }
    }
    const double dV = p->dx[0] * p->dx[1] * p->dx[2];
    const size_t idx = size_t(id);
    const int ix = int(idx % size_t(nx)), iy = int((idx / size_t(nx)) % size_t(ny)), iz = int(idx / (size_t(nx) * ny));
    r[WDM_DIAG_MASS] = S[0].value() * dV;
    r[WDM_DIAG_RHOMAX] = rmax;
    r[WDM_DIAG_PEAK_IX] = ix;
    r[WDM_DIAG_PEAK_IY] = iy;
    r[WDM_DIAG_PEAK_IZ] = iz;
    {
        const double W = S[7].value(), mu = S[8].value() / W;
        r[WDM_DIAG_R] = std::sqrt(std::max(0.0, S[9].value() / W - mu * mu));
    }
    for (int d = 0; d < 3; ++d) r[WDM_DIAG_MOMX + d] = S[1 + d].value() * dV;
    r[WDM_DIAG_PE_ABS] = -0.5 * S[4].value() * dV;
    r[WDM_DIAG_KE] = 0.5 * S[5].value() * dV;
    r[WDM_DIAG_QE] = 0.5 * S[6].value() * dV;
    r[WDM_DIAG_VIRIAL] = 2.0 * r[WDM_DIAG_KE] + 2.0 * r[WDM_DIAG_QE] - r[WDM_DIAG_PE_ABS];
    if (nout > WDM_DIAG_R1) {
        if (!(rmax > 0.0) || !std::isfinite(rmax)) {
            for (int q = 0; q < 9; ++q) r[WDM_DIAG_R1 + q] = NAN;
        } else {
            const double c[3] = {coord_at(p, 0, ix), coord_at(p, 1, iy), coord_at(p, 2, iz)};
            double far2 = 0.0;
            for (int a = 0; a < 3; ++a) {
                const double lo = coord_at(p, a, 0), hi = coord_at(p, a, p->ax[a].n - 1);
                const double e = std::max(std::fabs(c[a] - lo), std::fabs(hi - c[a]));
                far2 += e * e;
            }
            const double width = std::max(std::sqrt(far2), 1e-300) * (1.0 + 1e-12) / kDiagBins;
            // fixed point: every weight <= 2^K and the total < 2^62
            int lg = 0;
            while ((size_t(1) << lg) < p->ncell) ++lg;
            const double qscale = std::ldexp(1.0, 62 - lg) / rmax;
            WDM_LAUNCH((k_diag_hist<T>), dim3(nb), dim3(256), p->stream, p->psi.as<cx<T>>(), nx, ny, nz, xs, ys, zs,
                       c[0], c[1], c[2], 1.0 / width, qscale, hpart);
            hip_check(hipGetLastError(), "diagnostics histogram");
            const int per = int((nb + kDiagSlices - 1) / kDiagSlices);
            WDM_LAUNCH((k_diag_hist_sum), dim3((kDiagBins + 255) / 256, kDiagSlices), dim3(256), p->stream,
                       (const unsigned long long*)hpart, int(nb), per, hslice);
            WDM_LAUNCH((k_diag_hist_sum), dim3((kDiagBins + 255) / 256, 1), dim3(256), p->stream,
                       (const unsigned long long*)hslice, kDiagSlices, kDiagSlices, hist);
            hip_check(hipGetLastError(), "diagnostics histogram sum");
            std::vector<unsigned long long> hh(kDiagBins);
            hip_check(hipMemcpyAsync(hh.data(), hist, sizeof(unsigned long long) * kDiagBins, hipMemcpyDeviceToHost,
                                     p->stream),
                      "histogram copy");
            hip_check(hipStreamSynchronize(p->stream), "histogram sync");
            double total = 0.0;
            for (unsigned long long v : hh) total += double(v);
            const double w1 = __builtin_rint(rmax * qscale);
            for (int q = 0; q < 9; ++q) r[WDM_DIAG_R1 + q] = hist_quantile(hh, w1, total, 0.1 * (q + 1), width);
        }
    }
    for (int i = 0; i < nout && i < WDM_DIAG_COUNT; ++i) out[i] = r[i];
}

template <class T>
void acceleration(wdm_plan* p, double* const host[3]) {
    if (p->kspace) fail(WDM_ERR_STATE, "wdm_acceleration: psi is in k-space");
    double* phi = p->scratch_real.get<double>(p->ncell);
    phi_all<T>(p, phi);
    double* g = p->diag.get<double>(p->ncell);
    for (int a = 0; a < 3; ++a) {
        WDM_LAUNCH((k_neg_cdiff), dim3(adj_line_blocks(p->ny, p->nz)), dim3(256), p->stream, (const double*)phi, g,
                   p->nx, p->ny, p->nz, a, 0.5 / p->dx[a]);
        hip_check(hipGetLastError(), "acceleration");
        hip_check(hipMemcpyAsync(host[a], g, sizeof(double) * p->ncell, hipMemcpyDeviceToHost, p->stream),
                  "acceleration copy");
        hip_check(hipStreamSynchronize(p->stream), "acceleration sync");
    }
}

// acc += |psi|^2 (FP64), the loop's time-averaged density buffer
template <class T>
WDM_KERNEL(256)
void k_density_accumulate(const cx<T>* __restrict__ psi, double* __restrict__ acc, int nx, int ny, int nz) {
    WDM_FOR_CELLS(nx, ny, nz) {
        (void)y;
        (void)z;
        acc[i] += rho_of(psi[i]);
    }
}

template <class T>
void density_accumulate(wdm_plan* p) {
    if (p->kspace) fail(WDM_ERR_STATE, "wdm_density_accumulate: psi is in k-space");
    double* acc = p->rho_avg.as<double>();
    if (p->rho_avg_count == 0) {
        acc = p->rho_avg.get<double>(p->ncell);
        WDM_LAUNCH((k_fill_zero), dim3(adj_line_blocks(p->ny, p->nz)), dim3(256), p->stream, acc, p->nx, p->ny, p->nz);
    }
    WDM_LAUNCH((k_density_accumulate<T>), dim3(adj_line_blocks(p->ny, p->nz)), dim3(256), p->stream,
               (const cx<T>*)p->psi.as<cx<T>>(), acc, p->nx, p->ny, p->nz);
    hip_check(hipGetLastError(), "density accumulate");
    hip_check(hipStreamSynchronize(p->stream), "density accumulate sync");
    ++p->rho_avg_count;
}

}  // namespace

extern "C" {

int wdm_density_accumulate(wdm_plan* p) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    return guarded(p, [&] { DISPATCH(p, density_accumulate, p); });
}

int wdm_density_average(wdm_plan* p, double* host_out, long long* count) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    return guarded(p, [&] {
        if (count) *count = p->rho_avg_count;
        if (!host_out) return;
        if (p->rho_avg_count == 0) fail(WDM_ERR_STATE, "wdm_density_average: nothing accumulated");
        std::vector<double> h(p->ncell);
        hip_check(hipMemcpyAsync(h.data(), p->rho_avg.as<double>(), sizeof(double) * p->ncell, hipMemcpyDeviceToHost,
                                 p->stream),
                  "density average copy");
        hip_check(hipStreamSynchronize(p->stream), "density average sync");
        const double inv = 1.0 / double(p->rho_avg_count);
        for (size_t i = 0; i < p->ncell; ++i) host_out[i] = h[i] * inv;
    });
}

int wdm_density_reset(wdm_plan* p) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    p->rho_avg_count = 0;
    p->rho_avg.reset();
    return WDM_OK;
}

int wdm_density_sum(wdm_plan* p, double* host_sum, long long* count) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    return guarded(p, [&] {
        if (count) *count = p->rho_avg_count;
        if (!host_sum) return;
        if (p->rho_avg_count == 0) fail(WDM_ERR_STATE, "wdm_density_sum: nothing accumulated");
        hip_check(hipMemcpyAsync(host_sum, p->rho_avg.as<double>(), sizeof(double) * p->ncell, hipMemcpyDeviceToHost,
                                 p->stream),
                  "density sum copy");
        hip_check(hipStreamSynchronize(p->stream), "density sum sync");
    });
}

int wdm_density_restore(wdm_plan* p, const double* host_sum, long long count) {
    if (!p || !host_sum || count <= 0) return fail_code(p, WDM_ERR_ARG, "wdm_density_restore: bad argument");
    return guarded(p, [&] {
        double* acc = p->rho_avg.get<double>(p->ncell);
        hip_check(hipMemcpyAsync(acc, host_sum, sizeof(double) * p->ncell, hipMemcpyHostToDevice, p->stream),
                  "density restore copy");
        hip_check(hipStreamSynchronize(p->stream), "density restore sync");
        p->rho_avg_count = count;
    });
}

int wdm_diagnostics(wdm_plan* p, double* out, int nout) {
    if (!p || !out || nout <= 0) return fail_code(p, WDM_ERR_ARG, "wdm_diagnostics: bad argument");
    return guarded(p, [&] { DISPATCH(p, diagnostics, p, out, nout); });
}

int wdm_acceleration(wdm_plan* p, double* ax, double* ay, double* az) {
    if (!p || !ax || !ay || !az) return fail_code(p, WDM_ERR_ARG, "wdm_acceleration: NULL argument");
    double* const h[3] = {ax, ay, az};
    return guarded(p, [&] { DISPATCH(p, acceleration, p, h); });
}

}  // extern "C"
