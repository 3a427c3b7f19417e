// Plan object and C ABI (include/wavedm.h).  Compiled as HIP for the product
// library and as plain C++ for the CPU emulation test build.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/wavedm.h"
#include "wdm_ops.hpp"

using namespace wdm;

namespace {

thread_local std::string g_err;

struct WdmError {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw WdmError{code, msg}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(WDM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T* dalloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) fail(WDM_ERR_OOM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
    return static_cast<T*>(p);
}

struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void reset() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* get(size_t n) {
        const size_t b = std::max<size_t>(n * sizeof(T), 1);
        if (bytes < b) {
            reset();
            p = dalloc<char>(b);
            bytes = b;
        }
        return static_cast<T*>(p);
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Julia LinRange(a, b, n)[j] = (1 - t) a + t b, t = j/(n-1)
double linrange_at(double a, double b, int n, int j) {
    if (n == 1) return a;
    const double t = double(j) / double(n - 1);
    return (1.0 - t) * a + t * b;
}

// drift-lattice frequency of FFT index m along one axis (src/solver/fft_operators.jl:9-13):
// f = LinRange(-n/(4X), n/(4X) - 1/(2X), n)[(m - n div 2) mod n]; k = 2 pi f.
double drift_k(int n, double X, int m) {
    if (n == 1) return 0.0;
    const int j = ((m - n / 2) % n + n) % n;
    const double f = linrange_at(-n / 4.0 / X, n / 4.0 / X - 1.0 / 2.0 / X, n, j);
    return 2.0 * M_PI * f;
}

}  // namespace

struct AxisState {
    int n = 0;
    const void* ops = nullptr;  // AxisOps<T>*
    DBuf tw;                    // twiddle table
    DBuf coord;                 // FP64 coordinates
    DBuf drift;                 // complex FP64 drift factors for dt_cached
    DBuf drift_conj;            // their complex conjugates (adjoint)
    DBuf k2;                    // k^2 (FFT order; x axis: half spectrum), FP64 or FP32
};

// Start/stop HIP events attached to each pass launch (wdm_step_timed).
struct PassTimer {
    std::vector<hipEvent_t> ev;  // 2 per launch
    std::vector<int> id;
    size_t used = 0;             // launches recorded
    unsigned mask = ~0u;         // passes that get events (bit WDM_PASS_*)
    void reserve(size_t launches) {
        while (ev.size() < 2 * launches) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) break;
            ev.push_back(e);
        }
        id.resize(ev.size() / 2);
    }
    ~PassTimer() {
        for (hipEvent_t e : ev) hipEventDestroy(e);
    }
};

namespace {
struct AdjState;
void adj_release(wdm_plan* p);
}  // namespace
struct TidalState;
namespace {
void tidal_release(wdm_plan* p);
template <class T>
void tidal_prepare(wdm_plan* p, double t);
}  // namespace

struct wdm_plan {
    int nx, ny, nz;
    size_t ncell;
    double X[3];
    double dx[3];
    int precision;
    unsigned flags;
    double hbar_over_m = 1.0;
    double four_pi_G = 4.0 * M_PI;
    bool self_gravity = true;
    int hpitch = 0;
    // the inverse x pass streams psi past the caches (psi larger than the
    // memory-side cache, the half spectrum not: it then stays cached for the
    // Poisson passes)
    bool stream_psi = false;
    hipStream_t stream = nullptr;
    AxisState ax[3];
    DBuf psi, rhohat, partials, redout, border, vstatic, scratch, scratch_real;
    DBuf diag;  // diagnostics partials and histograms
    DBuf rho_avg;  // sum of |psi|^2 over wdm_density_accumulate calls (FP64)
    long long rho_avg_count = 0;
    // opt-in mass projection (wdm_set_mass_ledger): [reference mass, pending scale]
    bool ledger = false, ledger_stale = true;
    DBuf ledger_buf;
    size_t npartials = 0;
    bool border_on = false;
    // state
    bool kspace = false;    // psi holds the 3-D spectrum (after wdm_kick)
    bool phi_valid = false; // rhohat holds Phi_hat of the current psi
    bool mass_valid = false;
    double dt_cached = NAN, hbar_cached = NAN;
    // potential models
    int vflags = 0;
    double hq[5] = {0, 0, 0, 0, 0};
    double quad[13] = {0};
    double kappa = 0.0;
    double last_kick_t = 0.0;
    PassTimer* timer = nullptr;
    AdjState* adj = nullptr;  // adjoint tape (wdm_adjoint_*), created on first use
    TidalState* tidal = nullptr;  // MW tidal table and trajectory (wdm_set_tidal_model)
    TidalModel tidal_kick;        // the tidal term's constants at the last kick
    void* zspec_out = nullptr; // when set, the fused z pass also stores the 3-D spectrum here
    unsigned gen = 0;          // bumped by every change a captured step would not see
    unsigned long long psi_epoch = 0;  // bumped whenever psi is replaced or stepped
    std::string err;
};

namespace {

// Arm the events for the next launch (call before it) / record which pass it
// was (call after it).
void arm(wdm_plan* p, int pass) {
#ifdef WDM_GPU
    if (!p->timer) return;
    PassTimer& t = *p->timer;
    if (!((t.mask >> pass) & 1u) || t.used >= t.id.size()) return;
    launch_events().start = t.ev[2 * t.used];
    launch_events().stop = t.ev[2 * t.used + 1];
#else
    (void)p;
    (void)pass;
#endif
}
void mark(wdm_plan* p, int pass) {
#ifdef WDM_GPU
    launch_events() = LaunchEvents();
#endif
    if (!p->timer) return;
    PassTimer& t = *p->timer;
    if (!((t.mask >> pass) & 1u) || t.used >= t.id.size()) return;
    t.id[t.used++] = pass;
}

template <class T>
const AxisOps<T>* ops_of(const wdm_plan* p, int a) {
    return static_cast<const AxisOps<T>*>(p->ax[a].ops);
}

// Twiddle table of an axis of length n: W_n^j for j < n (FP64 {c, s}; FP32
// {c_hi, s_hi, c_lo, s_lo} split by split_hl_table).  Host side, into h[0..n).
template <class T>
WDM_HDFI typename TwEntry<T>::type twiddle_entry(int j, int n) {
    const CS v = cs2pi(j, n);
    if constexpr (std::is_same<T, double>::value) {
        return mk(v.c, v.s);
    } else {
        const HL c = split_hl_table(v.c), s = split_hl_table(v.s);
        TwF w;
        w.ch = c.h;
        w.cl = c.l;
        w.sh = s.h;
        w.sl = s.l;
        return w;
    }
}
template <class T>
void fill_twiddles(typename TwEntry<T>::type* h, int n) {
    for (int j = 0; j < n; ++j) h[j] = twiddle_entry<T>(j, n);
}

template <class T>
void build_twiddles(AxisState& a) {
    typedef typename TwEntry<T>::type TwE;
    std::vector<TwE> h(a.n);
    fill_twiddles<T>(h.data(), a.n);
    TwE* d = a.tw.get<TwE>(a.n);
    hip_check(hipMemcpy(d, h.data(), sizeof(TwE) * a.n, hipMemcpyHostToDevice), "twiddle upload");
}

void upload(DBuf& b, const std::vector<double>& h) {
    // synchronous copy; callers replacing a table that queued kernels may
    // still read synchronise the plan stream first
    double* d = b.get<double>(h.size());
    hip_check(hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice), "table upload");
}

void ensure_drift(wdm_plan* p, double dt) {
    if (p->dt_cached == dt && p->hbar_cached == p->hbar_over_m) return;
    hip_check(hipStreamSynchronize(p->stream), "drift table sync");
    for (int a = 0; a < 3; ++a) {
        const int n = p->ax[a].n;
        std::vector<double> h(2 * n);
        for (int m = 0; m < n; ++m) {
            const double k = drift_k(n, p->X[a], m);
            const double th = p->hbar_over_m * k * k * dt / 2.0;
            h[2 * m] = std::cos(th);
            h[2 * m + 1] = -std::sin(th);
        }
        upload(p->ax[a].drift, h);
        for (int m = 0; m < n; ++m) h[2 * m + 1] = -h[2 * m + 1];
        upload(p->ax[a].drift_conj, h);
    }
    p->dt_cached = dt;
    p->hbar_cached = p->hbar_over_m;
}

// k^2 of the Poisson lattice (period 2X) for axis a of length n: the x axis
// holds the half spectrum m = 0..n/2, y and z the FFT order.  Host side, into
// h[0..k2_count(a, n)) in the plan precision T.
WDM_HDFI int k2_count(int a, int n) { return a == 0 ? n / 2 + 1 : n; }
template <class T>
WDM_HDFI T k2_entry(int a, int n, int m, double X) {
    WDM_NO_CONTRACT
    const int q = (a == 0 || 2 * m < n) ? m : m - n;
    const double k = 2.0 * M_PI * q / (2.0 * X);
    return T(k * k);
}
template <class T>
void fill_k2(T* h, int a, int n, double X) {
    for (int m = 0; m < k2_count(a, n); ++m) h[m] = k2_entry<T>(a, n, m, X);
}

void build_k2(wdm_plan* p) {
    for (int a = 0; a < 3; ++a) {
        const int n = p->ax[a].n, cnt = k2_count(a, n);
        if (p->precision == WDM_F32) {
            std::vector<float> f(cnt);
            fill_k2<float>(f.data(), a, n, p->X[a]);
            float* d = p->ax[a].k2.get<float>(f.size());
            hip_check(hipMemcpy(d, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice), "k2 upload");
        } else {
            std::vector<double> h(cnt);
            fill_k2<double>(h.data(), a, n, p->X[a]);
            upload(p->ax[a].k2, h);
        }
    }
}

template <class T>
void launch_check(wdm_plan* p, const char* what) {
    (void)p;
    hip_check(hipGetLastError(), what);
}

// ---- pass sequences ------------------------------------------------------
// The Poisson chain on a half spectrum that holds R2C_x(rho): y FWD, z (FWD *
// -scale/(N k^2) * INV, k = 0 -> 0), y INV.  Buffer-level, shared by the plan
// (poisson_chain) and the allocation-free entry point, so both issue the same
// launches; hook(pass, before) brackets each launch (pass timers).
template <class T, class Hook>
void poisson_passes(const AxisOps<T>* oy, const AxisOps<T>* oz, int nx, int ny, int nz, int hp, cx<T>* rh,
                    const typename TwEntry<T>::type* twy, const typename TwEntry<T>::type* twz, const void* k2x,
                    const void* k2y, const void* k2z, double scale, hipStream_t s, Hook&& hook) {
    const int nh = nx / 2 + 1;
    const double ncell = double(nx) * double(ny) * double(nz);
    KTabs none = {nullptr, nullptr, nullptr, 0.0, nullptr};
    if (ny > 1) {
        hook(WDM_PASS_PYFWD, true);
        oy->col(COL_FWD, KF_PNONE, rh, nh, nz, unsigned(hp), unsigned(hp * ny), twy, none, s);
        hook(WDM_PASS_PYFWD, false);
    }
    // Phi_k = -scale * rho_k / k^2, with the 1/N of the unnormalised inverse folded in
    KTabs kt = {k2x, k2y, k2z, -scale / ncell, nullptr};
    hook(WDM_PASS_PZ, true);
    oz->col(COL_FMI, KF_POISSON, rh, nh, ny, unsigned(hp * ny), unsigned(hp), twz, kt, s);
    hook(WDM_PASS_PZ, false);
    if (ny > 1) {
        hook(WDM_PASS_PYINV, true);
        oy->col(COL_INV, KF_PNONE, rh, nh, nz, unsigned(hp), unsigned(hp * ny), twy, none, s);
        hook(WDM_PASS_PYINV, false);
    }
}

template <class T>
void poisson_chain(wdm_plan* p, double scale) {
    // rhohat holds R2C_x(rho)
    typedef typename TwEntry<T>::type TwE;
    poisson_passes<T>(ops_of<T>(p, 1), ops_of<T>(p, 2), p->nx, p->ny, p->nz, p->hpitch, p->rhohat.as<cx<T>>(),
                      p->ax[1].tw.as<TwE>(), p->ax[2].tw.as<TwE>(), p->ax[0].k2.p, p->ax[1].k2.p, p->ax[2].k2.p,
                      scale, p->stream, [p](int pass, bool before) {
                          if (before) arm(p, pass);
                          else mark(p, pass);
                      });
    launch_check<T>(p, "poisson chain");
}

template <class T>
void compute_phi(wdm_plan* p) {
    typedef typename TwEntry<T>::type TwE;
    const AxisOps<T>* ox = ops_of<T>(p, 0);
    const int nlines = p->ny * p->nz;
    ox->x_dens(p->psi.as<cx<T>>(), p->rhohat.as<cx<T>>(), nlines, p->hpitch, p->ax[0].tw.as<TwE>(),
               p->partials.as<double>(), p->stream);
    hip_check(hipGetLastError(), "density pass");
    poisson_chain<T>(p, p->four_pi_G);
    p->phi_valid = true;
    p->mass_valid = true;
}

void rotation_matrix(const double* n_in, double ang, double R[3][3]) {
    double n[3] = {n_in[0], n_in[1], n_in[2]};
    const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (nn == 0.0) {
        n[0] = 0; n[1] = 0; n[2] = 1;
    } else {
        for (double& v : n) v /= nn;
    }
    const double c = std::cos(ang), s = std::sin(ang), C = 1.0 - c;
    R[0][0] = c + n[0] * n[0] * C;
    R[0][1] = n[0] * n[1] * C - n[2] * s;
    R[0][2] = n[0] * n[2] * C + n[1] * s;
    R[1][0] = n[1] * n[0] * C + n[2] * s;
    R[1][1] = c + n[1] * n[1] * C;
    R[1][2] = n[1] * n[2] * C - n[0] * s;
    R[2][0] = n[2] * n[0] * C - n[1] * s;
    R[2][1] = n[2] * n[1] * C + n[0] * s;
    R[2][2] = c + n[2] * n[2] * C;
}

// R(t) of the rotating quadrupole (row-major; identity without the model)
void quad_rotation(const wdm_plan* p, double t, double* R9) {
    double R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (p->vflags & V_QUADRUPOLE) {
        const double period = p->quad[9];
        const double ang = period != 0.0 ? 2.0 * M_PI * t / period : 0.0;
        rotation_matrix(p->quad + 6, ang, R);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R9[3 * i + j] = R[i][j];
}

VModel make_vmodel(const wdm_plan* p, double t) {
    VModel vm;
    std::memset(&vm, 0, sizeof(vm));
    vm.flags = p->vflags;
    vm.vstatic = p->vstatic.as<double>();
    const double G = p->four_pi_G / (4.0 * M_PI);
    vm.hq_GM = G * p->hq[0];
    vm.hq_G = G;
    vm.hq_a = p->hq[1];
    vm.hq_c[0] = p->hq[2];
    vm.hq_c[1] = p->hq[3];
    vm.hq_c[2] = p->hq[4];
    if (p->vflags & V_QUADRUPOLE) {
        const double* q = p->quad;
        const double T0[3][3] = {{q[0], q[3], q[4]}, {q[3], q[1], q[5]}, {q[4], q[5], q[2]}};
        const double period = q[9];
        const double ang = period != 0.0 ? 2.0 * M_PI * t / period : 0.0;
        double R[3][3];
        rotation_matrix(q + 6, ang, R);
        double RT[3][3] = {{0}}, Tt[3][3] = {{0}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) RT[i][j] += R[i][k] * T0[k][j];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) Tt[i][j] += RT[i][k] * R[j][k];
        vm.qT[0] = Tt[0][0];
        vm.qT[1] = Tt[1][1];
        vm.qT[2] = Tt[2][2];
        vm.qT[3] = Tt[0][1];
        vm.qT[4] = Tt[0][2];
        vm.qT[5] = Tt[1][2];
        vm.q_c[0] = q[10];
        vm.q_c[1] = q[11];
        vm.q_c[2] = q[12];
    }
    vm.kappa = p->kappa;
    vm.gnx = p->nx;
    vm.gny = p->ny;
    vm.tidal = p->tidal_kick;
    // the tidal term exists once a kick has computed its constants
    if ((vm.flags & V_TIDAL) && !p->tidal_kick.table) vm.flags &= ~V_TIDAL;
    return vm;
}

template <class T>
KickArgs make_kick(const wdm_plan* p, double dt_kick, double t) {
    KickArgs ka;
    ka.mh_dt = dt_kick / p->hbar_over_m;
    ka.t = t;
    ka.use_phi = p->self_gravity ? 1 : 0;
    ka.Ny = p->ny;
    ka.xs = p->ax[0].coord.as<double>();
    ka.ys = p->ax[1].coord.as<double>();
    ka.zs = p->ax[2].coord.as<double>();
    // x coordinate in-kernel: LinRange(-X, X, nx) (or the consistent grid)
    if (p->flags & WDM_GRID_CONSISTENT) {
        ka.x0 = -p->X[0];
        ka.x1 = p->X[0];
        ka.xinv = 1.0 / double(p->nx);
    } else {
        ka.x0 = -p->X[0];
        ka.x1 = p->X[0];
        ka.xinv = p->nx > 1 ? 1.0 / double(p->nx - 1) : 0.0;
    }
    ka.vm = make_vmodel(p, t);
    ka.norm = p->ledger ? p->ledger_buf.as<double>() + 1 : nullptr;
    return ka;
}

template <class T>
void ledger_begin(wdm_plan* p);
template <class T>
void ledger_update(wdm_plan* p);
template <class T>
void ledger_finish(wdm_plan* p);

template <class T>
void ensure_phi(wdm_plan* p) {
    if (p->self_gravity && !p->phi_valid) compute_phi<T>(p);
}

// kick x-pass (FWD: fused forward x-FFT)
template <class T>
void x_kick(wdm_plan* p, bool fwd, double dt_kick, double t) {
    typedef typename TwEntry<T>::type TwE;
    const AxisOps<T>* ox = ops_of<T>(p, 0);
    if (p->vflags & V_TIDAL) tidal_prepare<T>(p, t);  // this kick's peak, moments and trajectory offset
    const KickArgs ka = make_kick<T>(p, dt_kick, t);
    arm(p, WDM_PASS_XKICK);
    ox->x_kick(fwd, p->psi.as<cx<T>>(), p->rhohat.as<cx<T>>(), p->ny * p->nz, p->hpitch, p->ax[0].tw.as<TwE>(), ka,
               p->stream);
    hip_check(hipGetLastError(), "kick pass");
    mark(p, WDM_PASS_XKICK);
    p->last_kick_t = t;
    p->phi_valid = false;
    p->mass_valid = false;
}

template <class T>
void y_pass(wdm_plan* p, int mode) {
    typedef typename TwEntry<T>::type TwE;
    if (p->ny == 1) return;
    const AxisOps<T>* oy = ops_of<T>(p, 1);
    KTabs none = {nullptr, nullptr, nullptr, 0.0, nullptr};
    arm(p, mode == COL_FWD ? WDM_PASS_YFWD : WDM_PASS_YINV);
    oy->col(mode, KF_NONE, p->psi.as<cx<T>>(), p->nx, p->nz, unsigned(p->nx), unsigned(p->nx * p->ny),
            p->ax[1].tw.as<TwE>(), none, p->stream);
    hip_check(hipGetLastError(), "y pass");
    mark(p, mode == COL_FWD ? WDM_PASS_YFWD : WDM_PASS_YINV);
}

template <class T>
void z_pass(wdm_plan* p, int mode, double dt) {
    typedef typename TwEntry<T>::type TwE;
    const AxisOps<T>* oz = ops_of<T>(p, 2);
    KTabs kt = {nullptr, nullptr, nullptr, 0.0, nullptr};
    if (mode == COL_FMI || mode == COL_MI) {
        ensure_drift(p, dt);
        kt = {p->ax[0].drift.p, p->ax[1].drift.p, p->ax[2].drift.p, 1.0 / double(p->ncell),
              mode == COL_FMI ? p->zspec_out : nullptr, p->ledger ? p->ledger_buf.as<double>() + 1 : nullptr};
        if (mode == COL_FMI && p->zspec_out) mode = COL_FMIS;
    } else if (p->nz == 1) {
        return;
    }
    arm(p, WDM_PASS_ZDRIFT);
    oz->col(mode, KF_NONE, p->psi.as<cx<T>>(), p->nx, p->ny, unsigned(p->nx * p->ny), unsigned(p->nx),
            p->ax[2].tw.as<TwE>(), kt, p->stream);
    hip_check(hipGetLastError(), "z pass");
    mark(p, WDM_PASS_ZDRIFT);
}

// inverse x-pass + border + density (+ Poisson chain when self-gravity is on)
template <class T>
void x_inv_close(wdm_plan* p) {
    typedef typename TwEntry<T>::type TwE;
    const AxisOps<T>* ox = ops_of<T>(p, 0);
    arm(p, WDM_PASS_XINV);
    (p->stream_psi && p->self_gravity ? ox->x_inv_dens_stream : ox->x_inv_dens)(
        p->psi.as<cx<T>>(), p->self_gravity ? p->rhohat.as<cx<T>>() : nullptr, p->ny * p->nz, p->hpitch,
        p->ax[0].tw.as<TwE>(), p->border_on ? p->border.as<double>() : nullptr, p->partials.as<double>(), p->stream);
    hip_check(hipGetLastError(), "inverse x pass");
    mark(p, WDM_PASS_XINV);
    p->kspace = false;
    p->mass_valid = p->self_gravity;
    if (p->self_gravity) {
        poisson_chain<T>(p, p->four_pi_G);
        p->phi_valid = true;
    }
    ledger_update<T>(p);
}

template <class T>
void do_kick(wdm_plan* p, double dt_kick, double t) {
    if (p->kspace) fail(WDM_ERR_STATE, "wdm_kick: psi is in k-space (call wdm_drift first)");
    ++p->psi_epoch;
    ledger_begin<T>(p);
    ensure_phi<T>(p);
    x_kick<T>(p, true, dt_kick, t);
    y_pass<T>(p, COL_FWD);
    z_pass<T>(p, COL_FWD, 0.0);
    p->kspace = true;
}

template <class T>
void do_drift(wdm_plan* p, double dt) {
    if (!p->kspace) fail(WDM_ERR_STATE, "wdm_drift: psi is not in k-space (call wdm_kick first)");
    ++p->psi_epoch;
    z_pass<T>(p, COL_MI, dt);
    y_pass<T>(p, COL_INV);
    x_inv_close<T>(p);
    ledger_finish<T>(p);
}

template <class T>
void do_step(wdm_plan* p, int nsteps, double dt, double dt_kick, double t0, int split) {
    if (p->kspace) fail(WDM_ERR_STATE, "wdm_step: psi is in k-space (call wdm_drift first)");
    if (nsteps <= 0) return;
    ++p->psi_epoch;
    ledger_begin<T>(p);
    ensure_phi<T>(p);
    ensure_drift(p, dt);
    if (split == WDM_SPLIT_REFERENCE) {
        for (int s = 0; s < nsteps; ++s) {
            x_kick<T>(p, true, dt_kick, t0 + s * dt);
            y_pass<T>(p, COL_FWD);
            z_pass<T>(p, COL_FMI, dt);
            y_pass<T>(p, COL_INV);
            x_inv_close<T>(p);
        }
        ledger_finish<T>(p);
    } else {
        typedef typename TwEntry<T>::type TwE;
        const AxisOps<T>* ox = ops_of<T>(p, 0);
        x_kick<T>(p, false, 0.5 * dt, t0);
        ox->x_fft(-1, p->psi.as<cx<T>>(), p->ny * p->nz, p->ax[0].tw.as<TwE>(), p->stream);
        for (int s = 0; s < nsteps; ++s) {
            y_pass<T>(p, COL_FWD);
            z_pass<T>(p, COL_FMI, dt);
            y_pass<T>(p, COL_INV);
            x_inv_close<T>(p);
            const double t = t0 + (s + 1) * dt;
            if (s + 1 < nsteps) x_kick<T>(p, true, dt, t);
            else x_kick<T>(p, false, 0.5 * dt, t);
        }
        ledger_finish<T>(p);  // the closing half kick used Phi f^2; psi still needs f
    }
}

template <class T>
void phi_real(wdm_plan* p, double* dst_dev) {
    typedef typename TwEntry<T>::type TwE;
    ensure_phi<T>(p);
    if (!p->self_gravity) {
        hip_check(hipMemsetAsync(dst_dev, 0, sizeof(double) * p->ncell, p->stream), "memset phi");
        return;
    }
    const AxisOps<T>* ox = ops_of<T>(p, 0);
    ox->x_c2r_real(p->rhohat.as<cx<T>>(), dst_dev, p->ny * p->nz, p->hpitch, p->ax[0].tw.as<TwE>(), p->stream);
    hip_check(hipGetLastError(), "C2R pass");
}

// Final reduction of per-wave partial pairs (sum, max): Neumaier-compensated
// sum in a fixed order, deterministic.  One block of 256 threads.
WDM_KERNEL(256)
void k_reduce_partials(const double* __restrict__ partials, int n, int second_sum, double* __restrict__ out) {
    WDM_SHARED double ssum[256], scmp[256], smax[256];
    const int t = int(threadIdx.x);
    double s = 0.0, comp = 0.0, mx = 0.0;
    for (int i = t; i < n; i += 256) {
        const double v = partials[2 * i];
        const double tt = s + v;
        if (__builtin_fabs(s) >= __builtin_fabs(v)) comp += (s - tt) + v;
        else comp += (v - tt) + s;
        s = tt;
        const double mv = partials[2 * i + 1];
        mx = second_sum ? mx + mv : (mv > mx ? mv : mx);
    }
    ssum[t] = s;
    scmp[t] = comp;
    smax[t] = mx;
    __syncthreads();
    if (t == 0) {
        double S = 0.0, C = 0.0, MX = 0.0;
        for (int i = 0; i < 256; ++i) {
            const double v = ssum[i];
            const double tt = S + v;
            if (__builtin_fabs(S) >= __builtin_fabs(v)) C += (S - tt) + v;
            else C += (v - tt) + S;
            S = tt;
            C += scmp[i];
            MX = second_sum ? MX + smax[i] : (smax[i] > MX ? smax[i] : MX);
        }
        out[0] = S + C;
        out[1] = MX;
    }
}

unsigned red_blocks(size_t n) { return unsigned(std::min<size_t>((n + 255) / 256, 1024)); }

template <class T>
WDM_KERNEL(256)
void k_psi_sums(const cx<T>* __restrict__ psi, size_t n, double* __restrict__ partials) {
    WDM_SHARED double s1[256], s2[256];
    const int t = int(threadIdx.x);
    double s = 0.0, mx = 0.0;
    for (size_t i = size_t(blockIdx.x) * 256 + t; i < n; i += size_t(gridDim.x) * 256) {
        const double v = double(norm2(psi[i]));
        s += v;
        mx = v > mx ? v : mx;
    }
    s1[t] = s;
    s2[t] = mx;
    __syncthreads();
    if (t == 0) {
        double S = 0.0, MX = 0.0;
        for (int i = 0; i < 256; ++i) {
            S += s1[i];
            MX = s2[i] > MX ? s2[i] : MX;
        }
        partials[2 * blockIdx.x] = S;
        partials[2 * blockIdx.x + 1] = MX;
    }
}

// out[0] = compensated sum of the first slots; out[1] = max (or sum) of the second
void finalize(wdm_plan* p, int n, double out[2], bool second_sum = false) {
    double* o = p->redout.get<double>(2);
    WDM_LAUNCH((k_reduce_partials), dim3(1), dim3(256), p->stream, p->partials.as<double>(), n,
               second_sum ? 1 : 0, o);
    hip_check(hipGetLastError(), "reduce");
    hip_check(hipMemcpyAsync(out, o, 2 * sizeof(double), hipMemcpyDeviceToHost, p->stream), "reduce copy");
    hip_check(hipStreamSynchronize(p->stream), "reduce sync");
}

template <class T>
void psi_sums(wdm_plan* p, double out[2]) {
    const unsigned nb = red_blocks(p->ncell);
    p->partials.get<double>(std::max<size_t>(p->npartials, 2 * nb));
    WDM_LAUNCH((k_psi_sums<T>), dim3(nb), dim3(256), p->stream, p->psi.as<cx<T>>(), p->ncell,
               p->partials.as<double>());
    hip_check(hipGetLastError(), "psi sums");
    finalize(p, int(nb), out);
}

// ---------------------------------------------------------------- mass projection
// Opt-in (wdm_set_mass_ledger).  ledger[0] = the reference mass (the mass of
// psi when the projection was switched on or psi was last set), ledger[1] =
// the pending scale f = sqrt(M_ref / M) computed from the per-wave mass
// partials of the inverse x pass.  The next step applies it where an FP64
// factor multiplies every element before a single rounding, the drift factor
// of the z pass (psi_k *= P f), so the correction survives FP32 rounding
// (a relative change of f - 1 ~ 1e-9 applied to FP32 values with ordinary
// rounding would round back to the same values); the kick before it uses
// Phi f^2 and kappa f^2, the potential of the projected state.  The last
// scale of a call is applied to the returned state by k_ledger_apply.  The sum runs in a fixed
// order (per-thread compensated, then pairwise).
WDM_KERNEL(1024)
void k_ledger_update(const double* __restrict__ partials, int n, double* __restrict__ ledger) {
    WDM_SHARED double sh[1024];
    const int t = int(threadIdx.x);
    double s = 0.0, comp = 0.0;
    for (int i = t; i < n; i += 1024) {
        const double v = partials[2 * i];
        const double tt = s + v;
        if (__builtin_fabs(s) >= __builtin_fabs(v)) comp += (s - tt) + v;
        else comp += (v - tt) + s;
        s = tt;
    }
    sh[t] = s + comp;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {  // pairwise, fixed order
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) {
        const double M = sh[0], ref = ledger[0];
        ledger[1] = (M > 0.0 && ref > 0.0) ? __builtin_sqrt(ref / M) : 1.0;
    }
}

// End of a step sequence: psi *= f with unbiased stochastic rounding (the
// product is formed in FP64 and rounded to the neighbour above or below with
// probability given by the remainder, from a hash of the element index), so a
// correction far below one FP32 ulp still moves the mass to M_ref on average.
// FP64 plans round to nearest.  Phi_hat is not rescaled: the next call derives
// it from the returned psi, so the state between calls is psi alone (a run
// restored from psi continues exactly).
// The index hash is the "lowbias32" integer hash (Chris Wellons'
// hash-prospector constants).
WDM_FI unsigned ledger_hash(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
// the neighbouring float of a finite r toward +inf (up) or -inf
WDM_FI float ledger_next(float r, bool up) {
    unsigned b = __builtin_bit_cast(unsigned, r);
    if ((b & 0x7fffffffu) == 0u) b = up ? 1u : 0x80000001u;
    else if ((b >> 31) == 0u) b = up ? b + 1u : b - 1u;
    else b = up ? b - 1u : b + 1u;
    return __builtin_bit_cast(float, b);
}
WDM_FI float ledger_round(double y, unsigned key) {
    const float r = float(y);
    const double d = y - double(r);
    if (d == 0.0) return r;
    const float nb = ledger_next(r, d > 0.0);
    const double p = d / (double(nb) - double(r));  // in (0, 1/2]
    const double u = double(ledger_hash(key) >> 8) * (1.0 / 16777216.0);
    return u < p ? nb : r;
}
WDM_FI cx<float> ledger_scale(cx<float> a, double f, size_t i) {
    return mk(ledger_round(double(a.x) * f, unsigned(2 * i)), ledger_round(double(a.y) * f, unsigned(2 * i + 1)));
}
WDM_FI cx<double> ledger_scale(cx<double> a, double f, size_t) { return a * f; }

template <class T>
WDM_KERNEL(256)
void k_ledger_apply(cx<T>* __restrict__ psi, size_t n, const double* __restrict__ ledger) {
    const double f = ledger[1];
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256)
        psi[i] = ledger_scale(psi[i], f, i);
}

WDM_KERNEL(64)
void k_ledger_set(double* __restrict__ ledger, double m_ref, int set_ref) {
    if (threadIdx.x == 0) {
        if (set_ref) ledger[0] = m_ref;
        ledger[1] = 1.0;
    }
}

template <class T>
void ledger_begin(wdm_plan* p) {
    if (!p->ledger || !p->ledger_stale) return;
    if (p->kspace) fail(WDM_ERR_STATE, "mass projection: psi is in k-space");
    double ms[2];
    psi_sums<T>(p, ms);
    WDM_LAUNCH((k_ledger_set), dim3(1), dim3(64), p->stream, p->ledger_buf.get<double>(2), ms[0], 1);
    hip_check(hipGetLastError(), "ledger set");
    p->ledger_stale = false;
}

template <class T>
void ledger_update(wdm_plan* p) {
    if (!p->ledger || !p->mass_valid) return;
    WDM_LAUNCH((k_ledger_update), dim3(1), dim3(1024), p->stream, (const double*)p->partials.as<double>(),
               int(p->npartials / 2), p->ledger_buf.as<double>());
    hip_check(hipGetLastError(), "ledger update");
}

// Apply the pending scale to the psi a call returns, then clear it so that
// the next z pass does not apply it again; Phi_hat (of the unscaled psi) is
// dropped and recomputed from the returned psi by the next call.
template <class T>
void ledger_finish(wdm_plan* p) {
    if (!p->ledger) return;
    const unsigned nb = unsigned(std::min<size_t>((p->ncell + 255) / 256, 65536));
    WDM_LAUNCH((k_ledger_apply<T>), dim3(nb), dim3(256), p->stream, p->psi.as<cx<T>>(), p->ncell,
               (const double*)p->ledger_buf.as<double>());
    WDM_LAUNCH((k_ledger_set), dim3(1), dim3(64), p->stream, p->ledger_buf.as<double>(), 0.0, 0);
    hip_check(hipGetLastError(), "ledger finish");
    p->phi_valid = false;
}

// Fixed-order final reduction of per-wave slot partials: block s reduces slot s
// of nwaves records of nslots doubles (Neumaier-compensated sum, or max when
// bit s of max_mask is set) into out[s].
WDM_KERNEL(256)
void k_reduce_slots(const double* __restrict__ partials, int nwaves, int nslots, unsigned max_mask,
                    double* __restrict__ out) {
    WDM_SHARED double ssum[256], scmp[256];
    const int t = int(threadIdx.x), slot = int(blockIdx.x);
    const bool is_max = (max_mask >> slot) & 1u;
    double sum = 0.0, comp = 0.0;
    for (int i = t; i < nwaves; i += 256) {
        const double v = partials[size_t(i) * size_t(nslots) + size_t(slot)];
        if (is_max) {
            sum = v > sum ? v : sum;
        } else {
            const double tt = sum + v;
            if (__builtin_fabs(sum) >= __builtin_fabs(v)) comp += (sum - tt) + v;
            else comp += (v - tt) + sum;
            sum = tt;
        }
    }
    ssum[t] = sum;
    scmp[t] = comp;
    __syncthreads();
    if (t == 0) {
        double S = 0.0, C = 0.0;
        for (int i = 0; i < 256; ++i) {
            const double v = ssum[i];
            if (is_max) {
                S = v > S ? v : S;
            } else {
                const double tt = S + v;
                if (__builtin_fabs(S) >= __builtin_fabs(v)) C += (S - tt) + v;
                else C += (v - tt) + S;
                S = tt;
                C += scmp[i];
            }
        }
        out[slot] = is_max ? S : S + C;
    }
}

// waves launched by an x pass / a column pass of a table
template <class T>
size_t x_waves(const AxisOps<T>* o, int nlines) {
    const size_t w = (size_t(nlines) + o->lpw - 1) / o->lpw;
    return (w + o->nw - 1) / o->nw * o->nw;
}
template <class T>
size_t col_waves(const AxisOps<T>* o, int ncols, int nother) {
    const size_t w = (size_t(ncols) + o->lpw - 1) / o->lpw * size_t(nother);
    return (w + o->nw - 1) / o->nw * o->nw;
}

// Diagnostics of the current psi by read-only passes (no N-sized scratch):
// k_x_red for mass, max rho, sum rho Phi, max |Phi|, sum rho V and the kx part
// of the kinetic sum; k_col_red for the ky and kz parts.  Phi is the plan's
// Phi_hat (computed first when psi changed), V the models at the last kick.
template <class T>
void reduce_all(wdm_plan* p, double* out, int nout) {
    typedef typename TwEntry<T>::type TwE;
    if (p->kspace) fail(WDM_ERR_STATE, "wdm_reduce: psi is in k-space");
    double r[WDM_RED_COUNT] = {0};
    const double dV = p->dx[0] * p->dx[1] * p->dx[2];
    const bool sg = p->self_gravity;
    if (sg && nout > WDM_RED_WSELF) ensure_phi<T>(p);
    const AxisOps<T>* ox = ops_of<T>(p, 0);
    const int nlines = p->ny * p->nz;
    KickArgs ka = make_kick<T>(p, 0.0, p->last_kick_t);
    ka.norm = nullptr;
    const double kx0 = 2.0 * M_PI / (2.0 * p->X[0]);
    double* part = p->partials.as<double>();
    double* o = p->redout.as<double>();
    ox->x_red(p->psi.as<cx<T>>(), (sg && nout > WDM_RED_WSELF) ? p->rhohat.as<cx<T>>() : nullptr, nlines,
              p->hpitch, p->ax[0].tw.as<TwE>(), ka, kx0 * kx0, part, p->stream);
    hip_check(hipGetLastError(), "x reduction");
    const unsigned maxmask = (1u << RS_RHOMAX) | (1u << RS_PHIMAX);
    WDM_LAUNCH((k_reduce_slots), dim3(RS_COUNT), dim3(256), p->stream, (const double*)part,
               int(x_waves(ox, nlines)), int(RS_COUNT), maxmask, o);
    double h[RS_COUNT + 2] = {0};
    size_t nh = RS_COUNT;
    if (nout > WDM_RED_KINETIC) {
        // ky and kz parts (their partials follow the x slots' final values in redout)
        for (int a = 1; a < 3; ++a) {
            const int n = p->ax[a].n;
            if (n == 1) {  // no k != 0 on this axis: its slot is an empty sum
                WDM_LAUNCH((k_reduce_slots), dim3(1), dim3(256), p->stream, (const double*)part, 0, 1, 0u,
                           o + RS_COUNT + (a - 1));
                continue;
            }
            const AxisOps<T>* oa = ops_of<T>(p, a);
            const double ka0 = 2.0 * M_PI / (2.0 * p->X[a]);
            const unsigned sl = a == 1 ? unsigned(p->nx) : unsigned(p->nx * p->ny);
            const unsigned so = a == 1 ? unsigned(p->nx * p->ny) : unsigned(p->nx);
            const int nother = a == 1 ? p->nz : p->ny;
            oa->col_red(p->psi.as<cx<T>>(), p->nx, nother, sl, so, p->ax[a].tw.as<TwE>(), ka0 * ka0, part, p->stream);
            hip_check(hipGetLastError(), "column reduction");
            WDM_LAUNCH((k_reduce_slots), dim3(1), dim3(256), p->stream, (const double*)part,
                       int(col_waves(oa, p->nx, nother)), 1, 0u, o + RS_COUNT + (a - 1));
        }
        nh = RS_COUNT + 2;
    }
    hip_check(hipGetLastError(), "reduce slots");
    hip_check(hipMemcpyAsync(h, o, sizeof(double) * nh, hipMemcpyDeviceToHost, p->stream), "reduce copy");
    hip_check(hipStreamSynchronize(p->stream), "reduce sync");
    r[WDM_RED_MASS] = h[RS_MASS] * dV;
    r[WDM_RED_PSIMAX] = std::sqrt(h[RS_RHOMAX]);
    if (nout > WDM_RED_KINETIC) {
        // sum_k k^2 |psi_hat|^2 / N = sum_a S_a / n_a (S_a: the axis-a pass's sum)
        const double sk = h[RS_KX] / double(p->nx) + h[RS_COUNT] / double(p->ny) + h[RS_COUNT + 1] / double(p->nz);
        r[WDM_RED_KINETIC] = 0.5 * p->hbar_over_m * p->hbar_over_m * sk * dV;
    }
    if (nout > WDM_RED_WSELF) {
        r[WDM_RED_WSELF] = sg ? 0.5 * h[RS_RHOPHI] * dV : 0.0;
        r[WDM_RED_VEXT] = h[RS_RHOV] * dV;
        r[WDM_RED_PHIMAX] = sg ? h[RS_PHIMAX] : 0.0;
    }
    for (int i = 0; i < nout && i < WDM_RED_COUNT; ++i) out[i] = r[i];
}

template <class T>
WDM_KERNEL(256)
void k_ic_solitons(cx<T>* __restrict__ psi, int nx, int ny, int nz, const double* __restrict__ xs,
                   const double* __restrict__ ys, const double* __restrict__ zs, const double* __restrict__ prm,
                   int nsol, double Lx, double Ly, double Lz) {
    const size_t n = size_t(nx) * ny * nz;
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
        const int x = int(i % nx), y = int((i / nx) % ny), z = int(i / (size_t(nx) * ny));
        const double px = xs[x], py = ys[y], pz = zs[z];
        double re = 0.0, im = 0.0;
        for (int s = 0; s < nsol; ++s) {
            const double* q = prm + 9 * s;
            double dx = px - q[0], dy = py - q[1], dz = pz - q[2];
            dx -= Lx * __builtin_rint(dx / Lx);
            dy -= Ly * __builtin_rint(dy / Ly);
            dz -= Lz * __builtin_rint(dz / Lz);
            const double u = (dx * dx + dy * dy + dz * dz) / (q[3] * q[3]);
            const double b = 1.0 + 0.091 * u;
            const double b2 = b * b, b4 = b2 * b2;
            const double amp = __builtin_sqrt(q[4]) / b4;  // sqrt(rho_c / b^8)
            const double ph = q[5] + q[6] * px + q[7] * py + q[8] * pz;
            int quad;
            const double r = reduce_quadrant(ph, quad);
            double sn, cs;
            sincos_small(r, sn, cs);
            // exp(i ph) = i^quad * (cs + i sn)
            const cx<double> e = rot_quadrant(mk(cs, -sn), quad);  // (cs - i sn) * (-i)^quad = conj(exp(i ph))
            re += amp * e.x;
            im -= amp * e.y;
        }
        psi[i] = mk(T(re), T(im));
    }
}

template <class T>
void ic_solitons(wdm_plan* p, const double* params, int nsol) {
    double* d = p->scratch_real.get<double>(std::max<size_t>(size_t(9) * nsol, 1));
    hip_check(hipMemcpyAsync(d, params, sizeof(double) * 9 * nsol, hipMemcpyHostToDevice, p->stream), "ic params");
    const unsigned nb = unsigned(std::min<size_t>((p->ncell + 255) / 256, 65536));
    WDM_LAUNCH((k_ic_solitons<T>), dim3(nb), dim3(256), p->stream, p->psi.as<cx<T>>(), p->nx, p->ny, p->nz,
               p->ax[0].coord.as<double>(), p->ax[1].coord.as<double>(), p->ax[2].coord.as<double>(),
               (const double*)d, nsol, 2.0 * p->X[0], 2.0 * p->X[1], 2.0 * p->X[2]);
    hip_check(hipGetLastError(), "ic kernel");
    hip_check(hipStreamSynchronize(p->stream), "ic sync");
    ++p->psi_epoch;
    p->kspace = false;
    p->phi_valid = false;
    p->mass_valid = false;
    p->ledger_stale = true;
}

template <class TD, class TS>
void convert_dev(void* dst, const void* src, size_t n, hipStream_t s) {
    WDM_LAUNCH((k_convert<TD, TS>), dim3(unsigned(std::min<size_t>((n + 255) / 256, 65535))), dim3(256), s,
               static_cast<cx<TD>*>(dst), static_cast<const cx<TS>*>(src), n);
    hip_check(hipGetLastError(), "convert");
}

template <class T>
void set_psi_dev(wdm_plan* p, const void* src, int dtype, bool host) {
    const size_t n = p->ncell;
    cx<T>* d = p->psi.as<cx<T>>();
    const bool same = (dtype == WDM_C128) == std::is_same<T, double>::value;
    if (host) {
        if (same) {
            hip_check(hipMemcpyAsync(d, src, n * sizeof(cx<T>), hipMemcpyHostToDevice, p->stream), "set_psi");
        } else {
            std::vector<cx<T>> tmp(n);
            if (dtype == WDM_C128) {
                const double* s = static_cast<const double*>(src);
                for (size_t i = 0; i < n; ++i) tmp[i] = mk(T(s[2 * i]), T(s[2 * i + 1]));
            } else {
                const float* s = static_cast<const float*>(src);
                for (size_t i = 0; i < n; ++i) tmp[i] = mk(T(s[2 * i]), T(s[2 * i + 1]));
            }
            hip_check(hipMemcpyAsync(d, tmp.data(), n * sizeof(cx<T>), hipMemcpyHostToDevice, p->stream), "set_psi");
            hip_check(hipStreamSynchronize(p->stream), "set_psi sync");
        }
    } else {
        if (same) hip_check(hipMemcpyAsync(d, src, n * sizeof(cx<T>), hipMemcpyDeviceToDevice, p->stream), "set_psi");
        else if (dtype == WDM_C128) convert_dev<T, double>(d, src, n, p->stream);
        else convert_dev<T, float>(d, src, n, p->stream);
    }
    hip_check(hipStreamSynchronize(p->stream), "set_psi sync");
    ++p->psi_epoch;
    p->kspace = false;
    p->phi_valid = false;
    p->mass_valid = false;
    p->ledger_stale = true;
}

template <class T>
void get_psi_dev(wdm_plan* p, void* dst, int dtype, bool host) {
    const size_t n = p->ncell;
    const cx<T>* d = p->psi.as<cx<T>>();
    const bool same = (dtype == WDM_C128) == std::is_same<T, double>::value;
    if (host) {
        if (same) {
            hip_check(hipMemcpyAsync(dst, d, n * sizeof(cx<T>), hipMemcpyDeviceToHost, p->stream), "get_psi");
            hip_check(hipStreamSynchronize(p->stream), "get_psi sync");
        } else {
            std::vector<cx<T>> tmp(n);
            hip_check(hipMemcpyAsync(tmp.data(), d, n * sizeof(cx<T>), hipMemcpyDeviceToHost, p->stream), "get_psi");
            hip_check(hipStreamSynchronize(p->stream), "get_psi sync");
            if (dtype == WDM_C128) {
                double* o = static_cast<double*>(dst);
                for (size_t i = 0; i < n; ++i) { o[2 * i] = double(tmp[i].x); o[2 * i + 1] = double(tmp[i].y); }
            } else {
                float* o = static_cast<float*>(dst);
                for (size_t i = 0; i < n; ++i) { o[2 * i] = float(tmp[i].x); o[2 * i + 1] = float(tmp[i].y); }
            }
        }
    } else {
        if (same) hip_check(hipMemcpyAsync(dst, d, n * sizeof(cx<T>), hipMemcpyDeviceToDevice, p->stream), "get_psi");
        else if (dtype == WDM_C128) convert_dev<double, T>(dst, d, n, p->stream);
        else convert_dev<float, T>(dst, d, n, p->stream);
        hip_check(hipStreamSynchronize(p->stream), "get_psi sync");
    }
}

template <class T>
void poisson_api(wdm_plan* p, const double* rho, double* phi_out, double scale) {
    typedef typename TwEntry<T>::type TwE;
    if (rho == nullptr && phi_out == nullptr) {
        // recompute Phi_hat of psi for the next kick
        if (p->kspace) fail(WDM_ERR_STATE, "wdm_poisson: psi is in k-space");
        if (scale != 0.0 && scale != p->four_pi_G) fail(WDM_ERR_ARG, "wdm_poisson: device result needs 4 pi G");
        compute_phi<T>(p);
        return;
    }
    const AxisOps<T>* ox = ops_of<T>(p, 0);
    const int nlines = p->ny * p->nz;
    if (scale == 0.0) scale = p->four_pi_G;
    if (rho) {
        double* src = p->scratch_real.get<double>(p->ncell);
        hip_check(hipMemcpyAsync(src, rho, sizeof(double) * p->ncell, hipMemcpyHostToDevice, p->stream), "rho upload");
        ox->x_r2c_real(src, p->rhohat.as<cx<T>>(), nlines, p->hpitch, p->ax[0].tw.as<TwE>(), p->stream);
        hip_check(hipGetLastError(), "R2C pass");
    } else {
        if (p->kspace) fail(WDM_ERR_STATE, "wdm_poisson: psi is in k-space");
        ox->x_dens(p->psi.as<cx<T>>(), p->rhohat.as<cx<T>>(), nlines, p->hpitch, p->ax[0].tw.as<TwE>(),
                   p->partials.as<double>(), p->stream);
        hip_check(hipGetLastError(), "density pass");
    }
    poisson_chain<T>(p, scale);
    // rhohat now holds a user-scaled potential: not the kick's Phi unless it is
    // the plan's own |psi|^2 with the plan's coefficient
    p->phi_valid = (rho == nullptr && scale == p->four_pi_G);
    if (phi_out) {
        double* dst = p->scratch_real.get<double>(p->ncell);
        ox->x_c2r_real(p->rhohat.as<cx<T>>(), dst, nlines, p->hpitch, p->ax[0].tw.as<TwE>(), p->stream);
        hip_check(hipGetLastError(), "C2R pass");
        hip_check(hipMemcpyAsync(phi_out, dst, sizeof(double) * p->ncell, hipMemcpyDeviceToHost, p->stream),
                  "phi download");
    }
    hip_check(hipStreamSynchronize(p->stream), "poisson sync");
}

// partial-sum slots: one pair per launched wave of the x passes of the current
// x-axis table (every launched wave writes its pair, including the idle tail
// of the last group)
template <class T>
void update_partials(wdm_plan* p) {
    const AxisOps<T>* ox = ops_of<T>(p, 0);
    const size_t xw = x_waves(ox, p->ny * p->nz);
    p->npartials = 2 * xw;
    // capacity: the step's mass pairs, the reductions' slots (k_x_red,
    // k_col_red of either column axis) and the diagnostics' blocks
    size_t cap = std::max<size_t>(size_t(RS_COUNT) * xw, 2 * 1024);
    if (p->ax[1].ops) cap = std::max(cap, col_waves(ops_of<T>(p, 1), p->nx, p->nz));
    if (p->ax[2].ops) cap = std::max(cap, col_waves(ops_of<T>(p, 2), p->nx, p->ny));
    p->partials.get<double>(cap);
}

template <class T>
void init_plan(wdm_plan* p) {
    for (int a = 0; a < 3; ++a) {
        AxisState& ax = p->ax[a];
        ax.ops = axis_ops(ax.n, (T*)nullptr);
        if (!ax.ops) fail(WDM_ERR_UNSUPPORTED, "axis length " + std::to_string(ax.n) + " is not supported");
        build_twiddles<T>(ax);
    }
    p->psi.get<cx<T>>(p->ncell);
#ifndef WDM_HPITCH_ALIGN
#define WDM_HPITCH_ALIGN 8
#endif
    // half-spectrum row pitch in elements (multiples of 4 and 16 measured
    // slower or no faster for the Poisson column passes at 384^3)
    p->hpitch = ((p->nx / 2 + 1) + WDM_HPITCH_ALIGN - 1) / WDM_HPITCH_ALIGN * WDM_HPITCH_ALIGN;
    p->rhohat.get<cx<T>>(size_t(p->hpitch) * p->ny * p->nz);
    {
        // MI355X's Infinity Cache (memory-side, shared by all XCDs)
        constexpr size_t kMallBytes = size_t(256) << 20;
        const size_t psi_bytes = sizeof(cx<T>) * p->ncell;
        const size_t hat_bytes = sizeof(cx<T>) * size_t(p->hpitch) * p->ny * p->nz;
        p->stream_psi = psi_bytes > kMallBytes && hat_bytes < kMallBytes;
    }
    update_partials<T>(p);
    p->redout.get<double>(RS_COUNT + 2);
}

int fail_code(wdm_plan* p, int code, const std::string& msg) {
    if (p) p->err = msg;
    g_err = msg;
    return code;
}

template <class F>
int guarded(wdm_plan* p, F&& f) {
    try {
        f();
        if (p) p->err.clear();
        return WDM_OK;
    } catch (const WdmError& e) {
        return fail_code(p, e.code, e.msg);
    } catch (const std::bad_alloc&) {
        return fail_code(p, WDM_ERR_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail_code(p, WDM_ERR_HIP, e.what());
    }
}

#define DISPATCH(p, FN, ...)                                   \
    do {                                                       \
        if ((p)->precision == WDM_F64) FN<double>(__VA_ARGS__); \
        else FN<float>(__VA_ARGS__);                           \
    } while (0)

// nsteps reference-split steps with HIP events around every pass launch:
// acc[pass] += summed kernel ms, *total = first-to-last event time.  Set-up
// work (Phi of the initial psi, drift tables) runs before the timed window.
void timed_steps(wdm_plan* p, int nsteps, double dt, double dt_kick, double t0, double* acc, double* total,
                 unsigned mask = ~0u) {
    PassTimer timer;
    timer.mask = mask;
    DISPATCH(p, ensure_phi, p);
    ensure_drift(p, dt);
    hip_check(hipStreamSynchronize(p->stream), "timed sync");
    int npass = 0;
    for (int i = 0; i < WDM_PASS_COUNT; ++i) npass += (mask >> i) & 1u;
    timer.reserve(size_t(npass) * nsteps);
    p->timer = &timer;
    try {
        DISPATCH(p, do_step, p, nsteps, dt, dt_kick, t0, WDM_SPLIT_REFERENCE);
    } catch (...) {
        p->timer = nullptr;
        throw;
    }
    p->timer = nullptr;
    hip_check(hipStreamSynchronize(p->stream), "timed sync");
    for (size_t i = 0; i < timer.used; ++i) {
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, timer.ev[2 * i], timer.ev[2 * i + 1]), "hipEventElapsedTime");
        if (timer.id[i] >= 0 && timer.id[i] < WDM_PASS_COUNT) acc[timer.id[i]] += ms;
    }
    if (total) {
        float ms = 0.f;
        if (timer.used > 0)
            hip_check(hipEventElapsedTime(&ms, timer.ev[0], timer.ev[2 * timer.used - 1]), "hipEventElapsedTime");
        *total = ms;
    }
}

}  // namespace

extern "C" {

int wdm_version(void) { return WDM_ABI_VERSION; }

const char* wdm_last_error(const wdm_plan* plan) {
    if (plan) return plan->err.c_str();
    return g_err.c_str();
}

int wdm_plan_create(wdm_plan** out, int nx, int ny, int nz, double xmax, double ymax, double zmax, int precision,
                    unsigned flags) {
    if (!out) return fail_code(nullptr, WDM_ERR_ARG, "wdm_plan_create: out is NULL");
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1) return fail_code(nullptr, WDM_ERR_ARG, "wdm_plan_create: bad grid size");
    if (!(xmax > 0) || !(ymax > 0) || !(zmax > 0))
        return fail_code(nullptr, WDM_ERR_ARG, "wdm_plan_create: box half-widths must be positive");
    if (precision != WDM_F64 && precision != WDM_F32)
        return fail_code(nullptr, WDM_ERR_ARG, "wdm_plan_create: precision must be WDM_F64 or WDM_F32");
    if (double(nx) * ny * nz >= 2147483648.0)
        return fail_code(nullptr, WDM_ERR_UNSUPPORTED, "wdm_plan_create: grid exceeds 2^31 cells");
    {
        // The pass kernels address psi and the half spectrum with 32-bit byte
        // offsets from a uniform base (ldg/stg): both arrays must stay below 4 GiB.
        const double elem = precision == WDM_F64 ? 16.0 : 8.0;
        const double hp = double(((nx / 2 + 1) + WDM_HPITCH_ALIGN - 1) / WDM_HPITCH_ALIGN * WDM_HPITCH_ALIGN);
        if (double(nx) * ny * nz * elem >= 4294967296.0 || hp * ny * nz * elem >= 4294967296.0)
            return fail_code(nullptr, WDM_ERR_UNSUPPORTED,
                             "wdm_plan_create: psi or the half spectrum would reach 4 GiB (32-bit byte offsets)");
    }
    wdm_plan* p = new (std::nothrow) wdm_plan();
    if (!p) return fail_code(nullptr, WDM_ERR_OOM, "wdm_plan_create: out of host memory");
    p->nx = nx;
    p->ny = ny;
    p->nz = nz;
    p->ncell = size_t(nx) * ny * nz;
    p->X[0] = xmax;
    p->X[1] = ymax;
    p->X[2] = zmax;
    p->precision = precision;
    p->flags = flags;
    const int rc = guarded(nullptr, [&] {
        hip_check(hipStreamCreate(&p->stream), "hipStreamCreate");
        const int ns[3] = {nx, ny, nz};
        for (int a = 0; a < 3; ++a) {
            p->ax[a].n = ns[a];
            std::vector<double> c(ns[a]);
            for (int i = 0; i < ns[a]; ++i)
                c[i] = (flags & WDM_GRID_CONSISTENT) ? -p->X[a] + i * (2.0 * p->X[a] / ns[a])
                                                     : linrange_at(-p->X[a], p->X[a], ns[a], i);
            if (flags & WDM_GRID_CONSISTENT) p->dx[a] = 2.0 * p->X[a] / ns[a];
            else p->dx[a] = ns[a] > 1 ? c[1] - c[0] : 2.0 * p->X[a];
            upload(p->ax[a].coord, c);
        }
        if (precision == WDM_F64) init_plan<double>(p);
        else init_plan<float>(p);
        build_k2(p);
        hip_check(hipStreamSynchronize(p->stream), "plan init");
    });
    if (rc != WDM_OK) {
        wdm_plan_destroy(p);
        return rc;
    }
    *out = p;
    return WDM_OK;
}

int wdm_plan_destroy(wdm_plan* p) {
    if (!p) return WDM_OK;
    if (p->stream) hipStreamSynchronize(p->stream);
    adj_release(p);
    tidal_release(p);
    for (auto& a : p->ax) {
        a.tw.reset();
        a.coord.reset();
        a.drift.reset();
        a.drift_conj.reset();
        a.k2.reset();
    }
    for (DBuf* b : {&p->psi, &p->rhohat, &p->partials, &p->redout, &p->border, &p->vstatic, &p->scratch,
                    &p->scratch_real, &p->diag, &p->rho_avg, &p->ledger_buf})
        b->reset();
    if (p->stream) hipStreamDestroy(p->stream);
    delete p;
    return WDM_OK;
}

int wdm_set_coeffs(wdm_plan* p, double hbar_over_m, double four_pi_G) {
    if (p) ++p->gen;
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (!(hbar_over_m > 0) || !std::isfinite(four_pi_G))
        return fail_code(p, WDM_ERR_ARG, "wdm_set_coeffs: hbar_over_m must be > 0 and four_pi_G finite");
    p->hbar_over_m = hbar_over_m;
    if (four_pi_G != p->four_pi_G) p->phi_valid = false;
    p->four_pi_G = four_pi_G;
    return WDM_OK;
}

int wdm_set_self_gravity(wdm_plan* p, int on) {
    if (p) ++p->gen;
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (!on && p->ledger)
        return fail_code(p, WDM_ERR_UNSUPPORTED, "wdm_set_self_gravity: the mass projection needs self-gravity");
    p->self_gravity = on != 0;
    p->phi_valid = false;
    return WDM_OK;
}

int wdm_set_mass_ledger(wdm_plan* p, int on) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    ++p->gen;
    if (on && (!p->self_gravity || p->border_on))
        return fail_code(p, WDM_ERR_UNSUPPORTED,
                         "wdm_set_mass_ledger: needs self-gravity (its density pass gives the mass) and no absorber");
    if (on && !p->ledger) {
        // [M_ref, f] = [0, 1] until the next kick or step sets the reference: a
        // drift in between (after a kick taken without the projection) scales by 1
        const int rc = guarded(p, [&] {
            WDM_LAUNCH((k_ledger_set), dim3(1), dim3(64), p->stream, p->ledger_buf.get<double>(2), 0.0, 1);
            hip_check(hipGetLastError(), "ledger init");
            hip_check(hipStreamSynchronize(p->stream), "ledger init sync");
        });
        if (rc != WDM_OK) return rc;
    }
    p->ledger = on != 0;
    p->ledger_stale = true;
    return WDM_OK;
}

int wdm_get_mass_ledger(wdm_plan* p, int* on, double* m_ref) {
    if (!p || !on || !m_ref) return fail_code(p, WDM_ERR_ARG, "wdm_get_mass_ledger: NULL argument");
    *on = p->ledger ? 1 : 0;
    *m_ref = NAN;
    if (!p->ledger || p->ledger_stale) return WDM_OK;
    return guarded(p, [&] {
        double h[2];
        hip_check(hipMemcpyAsync(h, p->ledger_buf.as<double>(), sizeof(h), hipMemcpyDeviceToHost, p->stream),
                  "ledger read");
        hip_check(hipStreamSynchronize(p->stream), "ledger read sync");
        *m_ref = h[0];
    });
}

int wdm_set_mass_reference(wdm_plan* p, double m_ref) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (!p->ledger) return fail_code(p, WDM_ERR_STATE, "wdm_set_mass_reference: the mass projection is off");
    if (!(m_ref > 0) || !std::isfinite(m_ref))
        return fail_code(p, WDM_ERR_ARG, "wdm_set_mass_reference: m_ref must be finite and > 0");
    if (p->kspace) return fail_code(p, WDM_ERR_STATE, "wdm_set_mass_reference: psi is in k-space");
    return guarded(p, [&] {
        WDM_LAUNCH((k_ledger_set), dim3(1), dim3(64), p->stream, p->ledger_buf.get<double>(2), m_ref, 1);
        hip_check(hipGetLastError(), "ledger set");
        hip_check(hipStreamSynchronize(p->stream), "ledger set sync");
        p->ledger_stale = false;
    });
}

int wdm_set_psi(wdm_plan* p, const void* host, int dtype) {
    if (!p || !host) return fail_code(p, WDM_ERR_ARG, "wdm_set_psi: NULL argument");
    if (dtype != WDM_C128 && dtype != WDM_C64) return fail_code(p, WDM_ERR_ARG, "wdm_set_psi: bad dtype");
    return guarded(p, [&] { DISPATCH(p, set_psi_dev, p, host, dtype, true); });
}

int wdm_get_psi(wdm_plan* p, void* host, int dtype) {
    if (!p || !host) return fail_code(p, WDM_ERR_ARG, "wdm_get_psi: NULL argument");
    if (dtype != WDM_C128 && dtype != WDM_C64) return fail_code(p, WDM_ERR_ARG, "wdm_get_psi: bad dtype");
    return guarded(p, [&] { DISPATCH(p, get_psi_dev, p, host, dtype, true); });
}

int wdm_set_psi_device(wdm_plan* p, const void* dev, int dtype) {
    if (!p || !dev) return fail_code(p, WDM_ERR_ARG, "wdm_set_psi_device: NULL argument");
    if (dtype != WDM_C128 && dtype != WDM_C64) return fail_code(p, WDM_ERR_ARG, "wdm_set_psi_device: bad dtype");
    return guarded(p, [&] { DISPATCH(p, set_psi_dev, p, dev, dtype, false); });
}

int wdm_get_psi_device(wdm_plan* p, void* dev, int dtype) {
    if (!p || !dev) return fail_code(p, WDM_ERR_ARG, "wdm_get_psi_device: NULL argument");
    if (dtype != WDM_C128 && dtype != WDM_C64) return fail_code(p, WDM_ERR_ARG, "wdm_get_psi_device: bad dtype");
    return guarded(p, [&] { DISPATCH(p, get_psi_dev, p, dev, dtype, false); });
}

int wdm_set_border(wdm_plan* p, const double* border, int rank) {
    if (p) ++p->gen;
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (!border) {
        p->border_on = false;
        return WDM_OK;
    }
    if (rank != 1) return fail_code(p, WDM_ERR_UNSUPPORTED, "wdm_set_border: only rank 1 (length nx) is supported");
    if (p->ledger)
        return fail_code(p, WDM_ERR_UNSUPPORTED, "wdm_set_border: the mass projection assumes no absorber");
    return guarded(p, [&] {
        hip_check(hipStreamSynchronize(p->stream), "border sync");
        upload(p->border, std::vector<double>(border, border + p->nx));
        p->border_on = true;
    });
}

int wdm_set_static_potential(wdm_plan* p, const double* V) {
    if (p) ++p->gen;
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (!V) {
        p->vflags &= ~V_STATIC;
        return WDM_OK;
    }
    return guarded(p, [&] {
        hip_check(hipStreamSynchronize(p->stream), "V sync");
        double* d = p->vstatic.get<double>(p->ncell);
        hip_check(hipMemcpy(d, V, sizeof(double) * p->ncell, hipMemcpyHostToDevice), "V upload");
        p->vflags |= V_STATIC;
    });
}

int wdm_set_potential_model(wdm_plan* p, int kind, const double* params, int nparams) {
    if (p) ++p->gen;
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (!params) return fail_code(p, WDM_ERR_ARG, "wdm_set_potential_model: params is NULL");
    switch (kind) {
        case WDM_V_HERNQUIST:
            if (nparams != 5) return fail_code(p, WDM_ERR_ARG, "Hernquist takes 5 parameters (M, a, cx, cy, cz)");
            std::copy(params, params + 5, p->hq);
            p->vflags |= V_HERNQUIST;
            return WDM_OK;
        case WDM_V_QUADRUPOLE:
            if (nparams != 13) return fail_code(p, WDM_ERR_ARG, "quadrupole takes 13 parameters");
            std::copy(params, params + 13, p->quad);
            p->vflags |= V_QUADRUPOLE;
            return WDM_OK;
        case WDM_V_KAPPA:
            if (nparams != 1) return fail_code(p, WDM_ERR_ARG, "kappa takes 1 parameter");
            p->kappa = params[0];
            p->vflags |= V_KAPPA;
            return WDM_OK;
        default:
            return fail_code(p, WDM_ERR_ARG, "unknown potential model " + std::to_string(kind));
    }
}

int wdm_clear_potential_models(wdm_plan* p) {
    if (p) ++p->gen;
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    p->vflags &= V_STATIC;
    tidal_release(p);
    return WDM_OK;
}

int wdm_poisson(wdm_plan* p, const double* rho, double* phi_out, double scale) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    return guarded(p, [&] { DISPATCH(p, poisson_api, p, rho, phi_out, scale); });
}

int wdm_get_phi(wdm_plan* p, double* host_phi) {
    if (!p || !host_phi) return fail_code(p, WDM_ERR_ARG, "wdm_get_phi: NULL argument");
    if (p->kspace) return fail_code(p, WDM_ERR_STATE, "wdm_get_phi: psi is in k-space");
    return guarded(p, [&] {
        double* d = p->scratch_real.get<double>(p->ncell);
        DISPATCH(p, phi_real, p, d);
        hip_check(hipMemcpyAsync(host_phi, d, sizeof(double) * p->ncell, hipMemcpyDeviceToHost, p->stream), "phi");
        hip_check(hipStreamSynchronize(p->stream), "phi sync");
    });
}

int wdm_kick(wdm_plan* p, double dt_kick, double t) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (!std::isfinite(dt_kick) || !std::isfinite(t)) return fail_code(p, WDM_ERR_ARG, "wdm_kick: non-finite argument");
    return guarded(p, [&] {
        DISPATCH(p, do_kick, p, dt_kick, t);
        hip_check(hipStreamSynchronize(p->stream), "kick sync");
    });
}

int wdm_drift(wdm_plan* p, double dt) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (!std::isfinite(dt)) return fail_code(p, WDM_ERR_ARG, "wdm_drift: non-finite dt");
    return guarded(p, [&] {
        DISPATCH(p, do_drift, p, dt);
        hip_check(hipStreamSynchronize(p->stream), "drift sync");
    });
}

int wdm_step(wdm_plan* p, int nsteps, double dt, double dt_kick, double t0, int split) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (nsteps < 0 || !std::isfinite(dt) || !std::isfinite(dt_kick) || !std::isfinite(t0))
        return fail_code(p, WDM_ERR_ARG, "wdm_step: bad argument");
    if (split != WDM_SPLIT_REFERENCE && split != WDM_SPLIT_STRANG)
        return fail_code(p, WDM_ERR_ARG, "wdm_step: unknown split mode");
    return guarded(p, [&] {
        DISPATCH(p, do_step, p, nsteps, dt, dt_kick, t0, split);
        hip_check(hipStreamSynchronize(p->stream), "step sync");
        if (p->mass_valid) {
            double ms[2];
            finalize(p, int(p->npartials / 2), ms);
            if (!std::isfinite(ms[0]) || !std::isfinite(ms[1]))
                fail(WDM_ERR_NONFINITE, "wdm_step: psi became non-finite");
        }
    });
}

int wdm_reduce(wdm_plan* p, double* out, int nout) {
    if (!p || !out || nout <= 0) return fail_code(p, WDM_ERR_ARG, "wdm_reduce: bad argument");
    return guarded(p, [&] { DISPATCH(p, reduce_all, p, out, nout); });
}

int wdm_cfl(wdm_plan* p, double kappa, double* out3) {
    if (!p || !out3) return fail_code(p, WDM_ERR_ARG, "wdm_cfl: bad argument");
    return guarded(p, [&] {
        double r[WDM_RED_COUNT];
        DISPATCH(p, reduce_all, p, r, WDM_RED_COUNT);
        const double mh = 1.0 / p->hbar_over_m;
        out3[0] = 2.0 * p->dx[0] * p->dx[0] / M_PI / p->hbar_over_m;
        out3[1] = r[WDM_RED_PHIMAX] > 0 ? 2.0 * M_PI / (mh * r[WDM_RED_PHIMAX]) : INFINITY;
        out3[2] = (kappa != 0.0 && r[WDM_RED_PSIMAX] > 0) ? 2.0 * M_PI / std::fabs(kappa) / mh / r[WDM_RED_PSIMAX]
                                                           : INFINITY;
    });
}

int wdm_step_timed(wdm_plan* p, int nsteps, double dt, double dt_kick, double t0, double* pass_ms, int npass,
                   double* total_ms) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (nsteps < 0 || !std::isfinite(dt) || !std::isfinite(dt_kick) || !std::isfinite(t0))
        return fail_code(p, WDM_ERR_ARG, "wdm_step_timed: bad argument");
    if (p->kspace) return fail_code(p, WDM_ERR_STATE, "wdm_step_timed: psi is in k-space");
    return guarded(p, [&] {
        double acc[WDM_PASS_COUNT] = {0}, tot = 0.0;
        timed_steps(p, nsteps, dt, dt_kick, t0, acc, &tot);
        for (int i = 0; i < npass && i < WDM_PASS_COUNT; ++i) pass_ms[i] = acc[i];
        if (total_ms) *total_ms = tot;
    });
}

int wdm_step_timed_mask(wdm_plan* p, int nsteps, double dt, double dt_kick, double t0, unsigned pass_mask,
                        double* pass_ms, int npass, double* total_ms) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    if (nsteps < 0 || !std::isfinite(dt) || !std::isfinite(dt_kick) || !std::isfinite(t0))
        return fail_code(p, WDM_ERR_ARG, "wdm_step_timed_mask: bad argument");
    if (p->kspace) return fail_code(p, WDM_ERR_STATE, "wdm_step_timed_mask: psi is in k-space");
    return guarded(p, [&] {
        double acc[WDM_PASS_COUNT] = {0}, tot = 0.0;
        timed_steps(p, nsteps, dt, dt_kick, t0, acc, &tot, pass_mask);
        for (int i = 0; i < npass && i < WDM_PASS_COUNT; ++i) pass_ms[i] = acc[i];
        if (total_ms) *total_ms = tot;
    });
}

int wdm_ic_solitons(wdm_plan* p, const double* params, int nsol) {
    if (!p || !params || nsol <= 0) return fail_code(p, WDM_ERR_ARG, "wdm_ic_solitons: bad argument");
    return guarded(p, [&] { DISPATCH(p, ic_solitons, p, params, nsol); });
}

int wdm_synchronize(wdm_plan* p) {
    if (!p) return fail_code(nullptr, WDM_ERR_ARG, "plan is NULL");
    return guarded(p, [&] { hip_check(hipStreamSynchronize(p->stream), "synchronize"); });
}

}  // extern "C"

#include "wdm_adjoint.inl"
#include "wdm_ensemble.inl"
#include "wdm_diag.inl"
#include "wdm_ic.inl"
#include "wdm_noalloc.inl"
#include "wdm_tidal.inl"
