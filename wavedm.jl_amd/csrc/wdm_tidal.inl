// Host side of the opt-in MW tidal potential (included at the end of
// wdm_plan.cpp; kernel side in wdm_tidal.hpp).  Per kick the host needs the
// density peak of the state being kicked (src/simulation_main.jl:554 keeps
// rho_max_id from the previous iteration's psi, which is that state) and the
// first moments of rho in cell-index units (the sum(ramps * rho) of
// cancel_field_gradient_at_center!): one read-only pass over psi
// (k_tidal_stats), reduced on the host in a fixed order, then the centre
// value and central differences of the interpolated table (host copy).

namespace {

constexpr int kTidalNV = 6;  // max rho, its index, sum rho, sum rho i, sum rho j, sum rho k

// Per-block: the first maximum of rho in x-fastest order and the moments.
template <class T>
WDM_KERNEL(256)
void k_tidal_stats(const cx<T>* __restrict__ psi, int nx, int ny, int nz, double* __restrict__ partials) {
    WDM_SHARED double sv[kTidalNV][256];
    double mx = -1.0, mi = 0.0, s0 = 0.0, si = 0.0, sj = 0.0, sk = 0.0;
    WDM_FOR_CELLS(nx, ny, nz) {
        const double r = rho_of(psi[i]);
        if (r > mx) {  // cells visited in increasing index per thread: keeps the first
            mx = r;
            mi = double(i);
        }
        s0 += r;
        si += r * double(x);
        sj += r * double(y);
        sk += r * double(z);
    }
    const int t = int(threadIdx.x);
    sv[0][t] = mx;
    sv[1][t] = mi;
    sv[2][t] = s0;
    sv[3][t] = si;
    sv[4][t] = sj;
    sv[5][t] = sk;
    __syncthreads();
    if (t == 0) {
        double M = -1.0, I = 0.0, S[4] = {0, 0, 0, 0};
        for (int l = 0; l < 256; ++l) {
            if (sv[0][l] > M || (sv[0][l] == M && sv[1][l] < I)) {
                M = sv[0][l];
                I = sv[1][l];
            }
            for (int q = 0; q < 4; ++q) S[q] += sv[2 + q][l];
        }
        double* o = partials + size_t(kTidalNV) * blockIdx.x;
        o[0] = M;
        o[1] = I;
        for (int q = 0; q < 4; ++q) o[2 + q] = S[q];
    }
}

}  // namespace

struct TidalState {
    DBuf table;                     // device copy of MW_Phi
    std::vector<double> host_table; // host copy (centre values)
    TidalModel m;                   // table geometry; per-kick fields set by tidal_prepare
    std::vector<double> traj_t, traj_xyz;
    DBuf part;
};

namespace {

void tidal_release(wdm_plan* p) {
    p->tidal_kick.table = nullptr;  // no kick constants without a table
    if (!p->tidal) return;
    p->tidal->table.reset();
    p->tidal->part.reset();
    delete p->tidal;
    p->tidal = nullptr;
}

// the trajectory sample of a kick at time t: the last with time <= t (the
// first before the table starts)
const double* tidal_offset(const TidalState& s, double t) {
    size_t k = 0;
    for (size_t q = 0; q < s.traj_t.size(); ++q)
        if (s.traj_t[q] <= t) k = q;
    return &s.traj_xyz[3 * k];
}

// Per-kick constants of the tidal term for the current psi (real space).
template <class T>
void tidal_prepare(wdm_plan* p, double t) {
    TidalState& s = *p->tidal;
    const unsigned nb = adj_line_blocks(p->ny, p->nz);
    double* part = s.part.get<double>(size_t(kTidalNV) * nb);
    WDM_LAUNCH((k_tidal_stats<T>), dim3(nb), dim3(256), p->stream, (const cx<T>*)p->psi.as<cx<T>>(), p->nx, p->ny,
               p->nz, part);
    hip_check(hipGetLastError(), "tidal stats");
    std::vector<double> h(size_t(kTidalNV) * nb);
    hip_check(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, p->stream),
              "tidal stats copy");
    hip_check(hipStreamSynchronize(p->stream), "tidal stats sync");
    double M = -1.0, I = 0.0, S[4] = {0, 0, 0, 0};
    for (unsigned b = 0; b < nb; ++b) {
        const double* o = &h[size_t(kTidalNV) * b];
        if (o[0] > M || (o[0] == M && o[1] < I)) {
            M = o[0];
            I = o[1];
        }
        for (int q = 0; q < 4; ++q) S[q] += o[2 + q];
    }
    if (!(S[0] > 0.0)) fail(WDM_ERR_STATE, "tidal potential: psi has no mass");
    const size_t ci = size_t(I);
    const int c[3] = {int(ci % size_t(p->nx)), int((ci / size_t(p->nx)) % size_t(p->ny)),
                      int(ci / (size_t(p->nx) * size_t(p->ny)))};
    TidalModel& m = s.m;
    const double* o = tidal_offset(s, t);
    for (int a = 0; a < 3; ++a) m.off[a] = o[a];
    // host evaluation on the host copy of the table
    TidalModel hm = m;
    hm.table = s.host_table.data();
    auto pot_at = [&](int i, int j, int k) {
        return tidal_interp(hm, coord_at(p, 0, i) + m.off[0], coord_at(p, 1, j) + m.off[1],
                            coord_at(p, 2, k) + m.off[2]);
    };
    m.c0 = pot_at(c[0], c[1], c[2]);
    const int n[3] = {p->nx, p->ny, p->nz};
    for (int a = 0; a < 3; ++a) {
        int up[3] = {c[0], c[1], c[2]}, dn[3] = {c[0], c[1], c[2]};
        up[a] = (c[a] + 1) % n[a];
        dn[a] = (c[a] - 1 + n[a]) % n[a];
        m.g[a] = (pot_at(up[0], up[1], up[2]) - pot_at(dn[0], dn[1], dn[2])) / 2.0;
        m.half[a] = n[a] / 2;
    }
    // sum_a g_a sum rho (i_a + 1 - div(N_a, 2)) / sum rho
    double pe = 0.0;
    for (int a = 0; a < 3; ++a) pe += m.g[a] * (S[1 + a] + S[0] * double(1 - m.half[a]));
    m.pe = pe / S[0];
    m.table = s.table.as<double>();
    p->tidal_kick = m;
}

}  // namespace

extern "C" {

int wdm_set_tidal_model(wdm_plan* p, const double* table, int tnx, int tny, int tnz, const double* lo,
                        const double* hi, const double* traj_t, const double* traj_xyz, int ntraj) {
    if (p) ++p->gen;
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (!table) {
        p->vflags &= ~V_TIDAL;
        tidal_release(p);
        return WDM_OK;
    }
    if (tnx < 2 || tny < 2 || tnz < 2 || !lo || !hi || !traj_t || !traj_xyz || ntraj < 1)
        return fail_code(p, WDM_ERR_ARG, "wdm_set_tidal_model: bad table or trajectory");
    for (int a = 0; a < 3; ++a)
        if (!(hi[a] > lo[a])) return fail_code(p, WDM_ERR_ARG, "wdm_set_tidal_model: hi must exceed lo");
    for (int q = 1; q < ntraj; ++q)
        if (!(traj_t[q] >= traj_t[q - 1]))
            return fail_code(p, WDM_ERR_ARG, "wdm_set_tidal_model: trajectory times must not decrease");
    return guarded(p, [&] {
        hip_check(hipStreamSynchronize(p->stream), "tidal sync");
        if (!p->tidal) p->tidal = new TidalState();
        TidalState& s = *p->tidal;
        const int n[3] = {tnx, tny, tnz};
        const size_t cnt = size_t(tnx) * size_t(tny) * size_t(tnz);
        s.host_table.assign(table, table + cnt);
        upload(s.table, s.host_table);
        std::memset(&s.m, 0, sizeof(s.m));
        for (int a = 0; a < 3; ++a) {
            s.m.n[a] = n[a];
            s.m.lo[a] = lo[a];
            s.m.step[a] = (hi[a] - lo[a]) / double(n[a] - 1);
            s.m.inv[a] = double(n[a] - 1) / (hi[a] - lo[a]);
        }
        s.m.table = s.table.as<double>();
        s.traj_t.assign(traj_t, traj_t + ntraj);
        s.traj_xyz.assign(traj_xyz, traj_xyz + 3 * size_t(ntraj));
        p->tidal_kick.table = nullptr;  // set by the next kick
        p->vflags |= V_TIDAL;
    });
}

}  // extern "C"
