// Autotuner and batched ensemble (included at the end of wdm_plan.cpp; shares
// its helpers).  No counterpart in the reference, which steps one box at a
// time (src/simulation_main.jl:536-551).
//
// wdm_autotune: for every axis of a plan, time each compiled pass-table variant
// (factorisation N1 x N2, waves per workgroup; wdm_ops.hpp) on trial steps of
// the plan's own state, and keep the fastest per axis.
//
// wdm_ensemble_*: many independent plans (uneven grids) stepped together.
// Members of one shape, precision, pass table and potential kind form a
// batch: each pass of the step is one launch over all of them (the member
// index in the grid's y, each member's arrays and factor tables read from a
// small device table), 8 launches per shape instead of 8 per member.  Members
// that fit no batch step on their own streams.  One step of the ensemble
// (every batch on its first member's stream and every other member on its
// own, forked from and joined into a root stream) is captured once into a HIP
// graph and launched once per step.

namespace {

// per-axis cost of a pass-timing vector: x = kick + inverse/density,
// y = psi y passes + Poisson y passes, z = drift + Poisson z
void axis_costs(const double* ms, double cost[3]) {
    cost[0] = ms[WDM_PASS_XKICK] + ms[WDM_PASS_XINV];
    cost[1] = ms[WDM_PASS_YFWD] + ms[WDM_PASS_YINV] + ms[WDM_PASS_PYFWD] + ms[WDM_PASS_PYINV];
    cost[2] = ms[WDM_PASS_ZDRIFT] + ms[WDM_PASS_PZ];
}

template <class T>
void set_axis_ops(wdm_plan* p, int a, const AxisOps<T>* ops) {
    ++p->gen;
    p->ax[a].ops = ops;
    if (a == 0) update_partials<T>(p);
}

template <class T>
void autotune(wdm_plan* p, int trial, double dt, int* chosen, double* best_ms) {
    if (p->kspace) fail(WDM_ERR_STATE, "wdm_autotune: psi is in k-space");
    const size_t n = p->ncell;
    // keep the state: trial steps run on a copy of psi
    cx<T>* keep = p->scratch.get<cx<T>>(n);
    hip_check(hipMemcpyAsync(keep, p->psi.as<cx<T>>(), sizeof(cx<T>) * n, hipMemcpyDeviceToDevice, p->stream),
              "autotune save");
    const AxisOps<T>* orig[3] = {ops_of<T>(p, 0), ops_of<T>(p, 1), ops_of<T>(p, 2)};
    const AxisOps<T>* best[3] = {orig[0], orig[1], orig[2]};
    double bcost[3] = {INFINITY, INFINITY, INFINITY};
    int bvar[3] = {0, 0, 0};
    for (int v = 0; v < kMaxVariants; ++v) {
        const AxisOps<T>* cand[3];
        bool any = false;
        for (int a = 0; a < 3; ++a) {
            cand[a] = axis_ops(p->ax[a].n, (T*)nullptr, v);
            any = any || cand[a] != nullptr;
        }
        if (!any) continue;
        for (int a = 0; a < 3; ++a) set_axis_ops<T>(p, a, cand[a] ? cand[a] : orig[a]);
        double ms[WDM_PASS_COUNT] = {0};
        for (int rep = 0; rep < 2; ++rep) {  // first round warms the code objects up
            hip_check(hipMemcpyAsync(p->psi.as<cx<T>>(), keep, sizeof(cx<T>) * n, hipMemcpyDeviceToDevice, p->stream),
                      "autotune restore");
            p->phi_valid = false;
            p->mass_valid = false;
            std::fill(ms, ms + WDM_PASS_COUNT, 0.0);
            timed_steps(p, trial, dt, 0.5 * dt, 0.0, ms, nullptr);
        }
        double cost[3];
        axis_costs(ms, cost);
        for (int a = 0; a < 3; ++a)
            if (cand[a] && cost[a] < bcost[a]) {
                bcost[a] = cost[a];
                best[a] = cand[a];
                bvar[a] = v;
            }
    }
    for (int a = 0; a < 3; ++a) set_axis_ops<T>(p, a, best[a]);
    hip_check(hipMemcpyAsync(p->psi.as<cx<T>>(), keep, sizeof(cx<T>) * n, hipMemcpyDeviceToDevice, p->stream),
              "autotune restore");
    hip_check(hipStreamSynchronize(p->stream), "autotune sync");
    p->phi_valid = false;
    p->mass_valid = false;
    for (int a = 0; a < 3; ++a) {
        if (chosen) chosen[a] = bvar[a];
        if (best_ms) best_ms[a] = bcost[a] / trial;
    }
}

template <class T>
void select_variants(wdm_plan* p, const int* v) {
    const AxisOps<T>* ops[3];
    for (int a = 0; a < 3; ++a) {
        ops[a] = axis_ops(p->ax[a].n, (T*)nullptr, v[a]);
        if (!ops[a])
            fail(WDM_ERR_ARG, "wdm_select_variants: axis " + std::to_string(a) + " (length " +
                                  std::to_string(p->ax[a].n) + ") has no variant " + std::to_string(v[a]));
    }
    hip_check(hipStreamSynchronize(p->stream), "select sync");
    for (int a = 0; a < 3; ++a) set_axis_ops<T>(p, a, ops[a]);
}

}  // namespace

struct wdm_ensemble {
    std::vector<wdm_plan*> m;
    hipStream_t root = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<double> g_dt, g_dtk;  // the step the graph holds
    std::vector<unsigned> g_gen;      // members' configuration at capture
    // batches of the current step and the device tables of their members'
    // arguments (separate for the graph, which keeps pointing at them, and
    // for eager steps)
    struct Batch {
        std::vector<int> idx;  // member indices
        int vk = 0;
        size_t cols_off = 0, x_off = 0;  // byte offsets of the batch's tables in args
    };
    std::vector<Batch> batches;
    std::vector<char> solo;  // 1: member steps on its own stream
    DBuf g_args, e_args;
};

namespace {

void ens_drop_graph(wdm_ensemble* e) {
#ifdef WDM_GPU
    if (e->exec) hipGraphExecDestroy(e->exec);
    if (e->graph) hipGraphDestroy(e->graph);
#endif
    e->exec = nullptr;
    e->graph = nullptr;
    e->g_dt.clear();
    e->g_dtk.clear();
    e->g_gen.clear();
}

std::vector<unsigned> ens_gens(const wdm_ensemble* e) {
    std::vector<unsigned> g;
    for (const wdm_plan* p : e->m) g.push_back(p->gen);
    return g;
}

bool time_dependent_v(const wdm_plan* p) {
    return ((p->vflags & V_QUADRUPOLE) && p->quad[9] != 0.0) || (p->vflags & V_TIDAL);
}

// number of column passes of a batched step: psi y FWD, z drift, y INV,
// Poisson y FWD, z, y INV
constexpr int kBatchCols = 6;

// a member can join a batch: the reference-split step with the fused kernels
// of the default build, no per-step host work
bool batchable(const wdm_plan* p) {
    return !p->ledger && !time_dependent_v(p) && !p->timer && !p->zspec_out;
}

// group the members into batches of >= 2 (same shape, precision, pass tables,
// gravity, absorber and potential kind); the rest step alone
void ens_group(wdm_ensemble* e) {
    const size_t nm = e->m.size();
    e->batches.clear();
    e->solo.assign(nm, 1);
    std::vector<char> used(nm, 0);
    for (size_t i = 0; i < nm; ++i) {
        if (used[i] || !batchable(e->m[i])) continue;
        const wdm_plan* a = e->m[i];
        wdm_ensemble::Batch b;
        b.idx.push_back(int(i));
        b.vk = vkind_of(make_vmodel(a, 0.0).flags);
        for (size_t j = i + 1; j < nm; ++j) {
            const wdm_plan* c = e->m[j];
            if (used[j] || !batchable(c)) continue;
            bool same = c->precision == a->precision && c->nx == a->nx && c->ny == a->ny && c->nz == a->nz &&
                        c->hpitch == a->hpitch && c->self_gravity == a->self_gravity &&
                        c->border_on == a->border_on && vkind_of(make_vmodel(c, 0.0).flags) == b.vk;
            for (int x = 0; x < 3 && same; ++x) same = c->ax[x].ops == a->ax[x].ops;
            if (same) b.idx.push_back(int(j));
        }
        if (b.idx.size() < 2) continue;
        for (int k : b.idx) {
            used[size_t(k)] = 1;
            e->solo[size_t(k)] = 0;
        }
        e->batches.push_back(b);
    }
}

// the device tables of every batch's member arguments for one step (kick at
// t0, time-independent potentials), built on the host and uploaded into buf
template <class T>
size_t batch_table_bytes(const wdm_ensemble::Batch& b) {
    return (sizeof(ColMember<T>) * kBatchCols + sizeof(XMember<T>)) * b.idx.size();
}
template <class T>
void batch_fill(wdm_ensemble* e, wdm_ensemble::Batch& b, const double* dt, const double* dtk, double t0,
                std::vector<char>& host, size_t off) {
    const size_t nb = b.idx.size();
    b.cols_off = off;
    b.x_off = off + sizeof(ColMember<T>) * kBatchCols * nb;
    ColMember<T>* cm = reinterpret_cast<ColMember<T>*>(host.data() + b.cols_off);
    XMember<T>* xm = reinterpret_cast<XMember<T>*>(host.data() + b.x_off);
    for (size_t k = 0; k < nb; ++k) {
        const size_t i = size_t(b.idx[k]);
        wdm_plan* p = e->m[i];
        cx<T>* psi = p->psi.as<cx<T>>();
        cx<T>* hat = p->self_gravity ? p->rhohat.as<cx<T>>() : nullptr;
        const KTabs none = {nullptr, nullptr, nullptr, 0.0, nullptr, nullptr};
        const KTabs drift = {p->ax[0].drift.p, p->ax[1].drift.p, p->ax[2].drift.p, 1.0 / double(p->ncell), nullptr,
                             nullptr};
        const double ncell = double(p->nx) * double(p->ny) * double(p->nz);
        const KTabs pois = {p->ax[0].k2.p, p->ax[1].k2.p, p->ax[2].k2.p, -p->four_pi_G / ncell, nullptr, nullptr};
        cm[0 * nb + k] = {psi, none};
        cm[1 * nb + k] = {psi, drift};
        cm[2 * nb + k] = {psi, none};
        cm[3 * nb + k] = {hat, none};
        cm[4 * nb + k] = {hat, pois};
        cm[5 * nb + k] = {hat, none};
        XMember<T> x;
        std::memset(&x, 0, sizeof(x));
        x.psi = psi;
        x.hat = hat;
        x.partials = p->partials.as<double>();
        x.border = p->border_on ? p->border.as<double>() : nullptr;
        x.ka = make_kick<T>(p, dtk[i], t0);
        xm[k] = x;
    }
    (void)dt;
}

template <class T>
void batch_issue(wdm_ensemble* e, const wdm_ensemble::Batch& b, const char* args, hipStream_t s) {
    typedef typename TwEntry<T>::type TwE;
    const wdm_plan* p = e->m[size_t(b.idx[0])];
    const int nmem = int(b.idx.size());
    const AxisOps<T>* ox = ops_of<T>(p, 0);
    const AxisOps<T>* oy = ops_of<T>(p, 1);
    const AxisOps<T>* oz = ops_of<T>(p, 2);
    const ColMember<T>* cm = reinterpret_cast<const ColMember<T>*>(args + b.cols_off);
    const XMember<T>* xm = reinterpret_cast<const XMember<T>*>(args + b.x_off);
    const int nx = p->nx, ny = p->ny, nz = p->nz, hp = p->hpitch, nh = nx / 2 + 1;
    const TwE* twx = p->ax[0].tw.as<TwE>();
    const TwE* twy = p->ax[1].tw.as<TwE>();
    const TwE* twz = p->ax[2].tw.as<TwE>();
    ox->x_kick_b(b.vk, xm, nmem, ny * nz, hp, twx, s);
    if (ny > 1) oy->col_b(COL_FWD, KF_NONE, cm + 0 * nmem, nmem, nx, nz, unsigned(nx), unsigned(nx * ny), twy, s);
    oz->col_b(COL_FMI, KF_DRIFT, cm + 1 * nmem, nmem, nx, ny, unsigned(nx * ny), unsigned(nx), twz, s);
    if (ny > 1) oy->col_b(COL_INV, KF_NONE, cm + 2 * nmem, nmem, nx, nz, unsigned(nx), unsigned(nx * ny), twy, s);
    ox->x_inv_dens_b(xm, nmem, ny * nz, hp, twx, s);
    if (p->self_gravity) {
        if (ny > 1)
            oy->col_b(COL_FWD, KF_PNONE, cm + 3 * nmem, nmem, nh, nz, unsigned(hp), unsigned(hp * ny), twy, s);
        oz->col_b(COL_FMI, KF_POISSON, cm + 4 * nmem, nmem, nh, ny, unsigned(hp * ny), unsigned(hp), twz, s);
        if (ny > 1)
            oy->col_b(COL_INV, KF_PNONE, cm + 5 * nmem, nmem, nh, nz, unsigned(hp), unsigned(hp * ny), twy, s);
    }
    hip_check(hipGetLastError(), "batched step");
}

// group the members, build and upload the batches' argument tables into buf
void ens_batches(wdm_ensemble* e, const double* dt, const double* dtk, double t0, DBuf& buf) {
    ens_group(e);
    size_t total = 0;
    for (auto& b : e->batches) {
        const wdm_plan* p = e->m[size_t(b.idx[0])];
        total += p->precision == WDM_F32 ? batch_table_bytes<float>(b) : batch_table_bytes<double>(b);
        total = (total + 255) / 256 * 256;
    }
    if (!total) return;
    std::vector<char> host(total, 0);
    size_t off = 0;
    for (auto& b : e->batches) {
        const wdm_plan* p = e->m[size_t(b.idx[0])];
        if (p->precision == WDM_F32) {
            batch_fill<float>(e, b, dt, dtk, t0, host, off);
            off += batch_table_bytes<float>(b);
        } else {
            batch_fill<double>(e, b, dt, dtk, t0, host, off);
            off += batch_table_bytes<double>(b);
        }
        off = (off + 255) / 256 * 256;
    }
    hip_check(hipMemcpy(buf.get<char>(total), host.data(), total, hipMemcpyHostToDevice), "batch tables");
}

// host-side state of a batched member after a step (as do_step leaves it)
void batch_member_done(wdm_plan* p, double t) {
    p->kspace = false;
    p->mass_valid = p->self_gravity;
    p->phi_valid = p->self_gravity;
    p->last_kick_t = t;
}

void ens_prepare(wdm_ensemble* e, const double* dt) {
    for (size_t i = 0; i < e->m.size(); ++i) {
        wdm_plan* p = e->m[i];
        if (p->kspace) fail(WDM_ERR_STATE, "wdm_ensemble_step: member " + std::to_string(i) + " is in k-space");
        DISPATCH(p, ledger_begin, p);  // outside any capture: it reads the mass back
        DISPATCH(p, ensure_phi, p);
        ensure_drift(p, dt[i]);
    }
}

// step s of every member (kick at t0 + s dt_i) on its own stream, forked from
// and joined into root
void ens_issue(wdm_ensemble* e, const double* dt, const double* dtk, double t0, int s, const DBuf& args) {
    hip_check(hipEventRecord(e->fork, e->root), "fork");
    for (size_t i = 0; i < e->m.size(); ++i) {
        if (!e->solo[i]) continue;
        wdm_plan* p = e->m[i];
        hip_check(hipStreamWaitEvent(p->stream, e->fork, 0), "fork wait");
        DISPATCH(p, do_step, p, 1, dt[i], dtk[i], t0 + s * dt[i], WDM_SPLIT_REFERENCE);
        hip_check(hipEventRecord(e->join[i], p->stream), "join");
    }
    // each batch on its first member's stream, so the batches overlap too
    for (const auto& b : e->batches) {
        const size_t i0 = size_t(b.idx[0]);
        wdm_plan* p = e->m[i0];
        hip_check(hipStreamWaitEvent(p->stream, e->fork, 0), "fork wait");
        if (p->precision == WDM_F32) batch_issue<float>(e, b, args.as<char>(), p->stream);
        else batch_issue<double>(e, b, args.as<char>(), p->stream);
        hip_check(hipEventRecord(e->join[i0], p->stream), "join");
        for (int k : b.idx) batch_member_done(e->m[size_t(k)], t0 + s * dt[size_t(k)]);
    }
    for (size_t i = 0; i < e->m.size(); ++i)
        if (e->solo[i]) hip_check(hipStreamWaitEvent(e->root, e->join[i], 0), "join wait");
    for (const auto& b : e->batches)
        hip_check(hipStreamWaitEvent(e->root, e->join[size_t(b.idx[0])], 0), "join wait");
}

void ens_step(wdm_ensemble* e, int nsteps, const double* dt, const double* dtk, double t0, bool use_graph) {
    const size_t nm = e->m.size();
    for (wdm_plan* p : e->m) ++p->psi_epoch;  // graph replays step psi without passing through do_step
    ens_prepare(e, dt);
    for (wdm_plan* p : e->m) hip_check(hipStreamSynchronize(p->stream), "member sync");
    if (use_graph)
        for (wdm_plan* p : e->m)
            if (time_dependent_v(p))
                fail(WDM_ERR_UNSUPPORTED, "wdm_ensemble_step: the graph holds one step; a member has a "
                                          "time-dependent potential (use eager mode)");
#ifdef WDM_GPU
    if (use_graph) {
        const bool same = e->exec && e->g_dt == std::vector<double>(dt, dt + nm) &&
                          e->g_dtk == std::vector<double>(dtk, dtk + nm) && e->g_gen == ens_gens(e);
        if (!same) {
            ens_drop_graph(e);
            // a member with the mass projection ends every step with the final
            // scaling of psi, after which the next step recomputes Phi from
            // the returned psi (as wdm_step does per call): the captured step
            // must start with that Poisson solve
            for (wdm_plan* p : e->m)
                if (p->ledger) p->phi_valid = false;
            ens_batches(e, dt, dtk, t0, e->g_args);
            hip_check(hipStreamBeginCapture(e->root, hipStreamCaptureModeRelaxed), "begin capture");
            try {
                ens_issue(e, dt, dtk, t0, 0, e->g_args);
            } catch (...) {
                hipGraph_t g = nullptr;
                hipStreamEndCapture(e->root, &g);
                if (g) hipGraphDestroy(g);
                throw;
            }
            hip_check(hipStreamEndCapture(e->root, &e->graph), "end capture");
            hip_check(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0), "graph instantiate");
            e->g_dt.assign(dt, dt + nm);
            e->g_dtk.assign(dtk, dtk + nm);
            e->g_gen = ens_gens(e);
        }
        for (int s = 0; s < nsteps; ++s) hip_check(hipGraphLaunch(e->exec, e->root), "graph launch");
        hip_check(hipStreamSynchronize(e->root), "ensemble sync");
        return;
    }
#else
    (void)use_graph;
#endif
    ens_batches(e, dt, dtk, t0, e->e_args);
    for (int s = 0; s < nsteps; ++s) ens_issue(e, dt, dtk, t0, s, e->e_args);
    hip_check(hipStreamSynchronize(e->root), "ensemble sync");
}

template <class F>
int ens_guarded(F&& f) {
    return guarded(nullptr, f);  // message: wdm_last_error(NULL)
}

}  // namespace

extern "C" {

int wdm_autotune(wdm_plan* p, int trial_steps, double dt, int* variants_out, double* ms_out) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (trial_steps < 1 || !(dt > 0) || !std::isfinite(dt))
        return fail_code(p, WDM_ERR_ARG, "wdm_autotune: trial_steps must be >= 1 and dt > 0");
    return guarded(p, [&] { DISPATCH(p, autotune, p, trial_steps, dt, variants_out, ms_out); });
}

int wdm_select_variants(wdm_plan* p, const int* variants) {
    if (!p || !variants) return fail_code(p, WDM_ERR_ARG, "wdm_select_variants: NULL argument");
    return guarded(p, [&] { DISPATCH(p, select_variants, p, variants); });
}

int wdm_ensemble_create(wdm_ensemble** out, wdm_plan* const* plans, int n) {
    if (!out || !plans || n < 1) return fail_code(nullptr, WDM_ERR_ARG, "wdm_ensemble_create: bad arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!plans[i]) return fail_code(nullptr, WDM_ERR_ARG, "wdm_ensemble_create: NULL member");
        for (int j = 0; j < i; ++j)
            if (plans[j] == plans[i]) return fail_code(nullptr, WDM_ERR_ARG, "wdm_ensemble_create: repeated member");
    }
    wdm_ensemble* e = new (std::nothrow) wdm_ensemble();
    if (!e) return fail_code(nullptr, WDM_ERR_OOM, "wdm_ensemble_create: out of host memory");
    const int rc = guarded(nullptr, [&] {
        e->m.assign(plans, plans + n);
        hip_check(hipStreamCreate(&e->root), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming), "hipEventCreate");
        e->join.resize(n, nullptr);
        for (int i = 0; i < n; ++i) hip_check(hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming), "event");
    });
    if (rc != WDM_OK) {
        wdm_ensemble_destroy(e);
        return rc;
    }
    *out = e;
    return WDM_OK;
}

int wdm_ensemble_destroy(wdm_ensemble* e) {
    if (!e) return WDM_OK;
    if (e->root) hipStreamSynchronize(e->root);
    ens_drop_graph(e);
    e->g_args.reset();
    e->e_args.reset();
    for (hipEvent_t ev : e->join)
        if (ev) hipEventDestroy(ev);
    if (e->fork) hipEventDestroy(e->fork);
    if (e->root) hipStreamDestroy(e->root);
    delete e;
    return WDM_OK;
}

int wdm_ensemble_step(wdm_ensemble* e, int nsteps, const double* dt, const double* dt_kick, double t0, int use_graph) {
    if (!e || !dt || !dt_kick || nsteps < 0) return fail_code(nullptr, WDM_ERR_ARG, "wdm_ensemble_step: bad arguments");
    for (size_t i = 0; i < e->m.size(); ++i)
        if (!std::isfinite(dt[i]) || !std::isfinite(dt_kick[i]))
            return fail_code(nullptr, WDM_ERR_ARG, "wdm_ensemble_step: non-finite dt");
    if (nsteps == 0) return WDM_OK;
    return ens_guarded([&] { ens_step(e, nsteps, dt, dt_kick, t0, use_graph != 0); });
}

int wdm_ensemble_layout(wdm_ensemble* e, int* nbatches, int* nsolo) {
    if (!e || !nbatches || !nsolo) return fail_code(nullptr, WDM_ERR_ARG, "wdm_ensemble_layout: NULL argument");
    *nbatches = int(e->batches.size());
    int n = 0;
    for (char c : e->solo) n += c;
    *nsolo = e->solo.empty() ? int(e->m.size()) : n;
    return WDM_OK;
}

int wdm_ensemble_autotune(wdm_ensemble* e, int trial_steps, const double* dt) {
    if (!e || !dt || trial_steps < 1) return fail_code(nullptr, WDM_ERR_ARG, "wdm_ensemble_autotune: bad arguments");
    return ens_guarded([&] {
        ens_drop_graph(e);
        // members of one shape and precision share the first member's choice
        std::vector<std::pair<std::vector<int>, std::vector<int>>> seen;
        for (size_t i = 0; i < e->m.size(); ++i) {
            wdm_plan* p = e->m[i];
            std::vector<int> key = {p->nx, p->ny, p->nz, p->precision};
            const std::vector<int>* hit = nullptr;
            for (auto& s : seen)
                if (s.first == key) hit = &s.second;
            if (hit) {
                DISPATCH(p, select_variants, p, hit->data());
            } else {
                int v[3];
                DISPATCH(p, autotune, p, trial_steps, dt[i], v, nullptr);
                seen.push_back({key, std::vector<int>(v, v + 3)});
            }
        }
    });
}

}  // extern "C"
