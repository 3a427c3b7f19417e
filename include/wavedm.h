/*
 * wavedm.h -- C ABI of libwavedm_hip.so, the MI355X (gfx950) backend for the
 * split-step Fourier (SSFM) kick-drift step and FFT Poisson gravity of
 * WaveDM.jl.
 *
 * Plain C types only.  All arrays are column-major / x-fastest (Julia layout:
 * psi[i + nx*(j + ny*k)]), complex values interleaved (re, im), exactly what
 * Julia's pointer(A) hands over for Array{ComplexF64,3} / Array{ComplexF32,3}.
 *
 * Every function returns WDM_OK (0) or a negative WDM_ERR_* code; the message
 * is available from wdm_last_error(plan) (or wdm_last_error(NULL) for errors
 * raised before a plan exists).  No exceptions cross the ABI.
 *
 * Threading: a plan owns one HIP stream; a plan is not thread-safe, distinct
 * plans are independent.  Calls return after the GPU work they issued has
 * completed.
 *
 * Reference interfaces replaced (paths relative to the WaveDM.jl tree):
 *   wdm_plan_create      setup_grid / setup_coordinates   src/physics/poisson.jl:8-28
 *                        + the loop's buffer set-up        src/simulation_main.jl:212-214,414-417
 *   wdm_set_psi          DeviceArray(psi) upload           src/simulation_main.jl:417
 *   wdm_get_psi          collect(device_psi)               src/simulation_main.jl:551
 *   wdm_set_border       setup_absorption_boundary         src/solver/fft_operators.jl:22-27
 *   wdm_set_static_potential / wdm_set_potential_model
 *                        the V(x,y,z,psi) callback         src/solver/KDK.jl:64
 *   wdm_poisson          parallel_poisson / compute_gravitational_potential
 *                                                          src/parallel_poisson.jl:36-68,
 *                                                          src/solver/gravity.jl:7-35
 *   wdm_get_phi          Phi_all returned by the kick      src/solver/KDK.jl:54-55
 *   wdm_kick             apply_kick_step!                  src/solver/KDK.jl:37-77
 *   wdm_drift            apply_drift_step! (+ setup_fft_operators, fft_operators.jl:8-15)
 *                                                          src/solver/KDK.jl:88-101
 *   wdm_step             the loop body, n iterations       src/simulation_main.jl:536-551
 *   wdm_cfl              compute_timestep limits           src/physics/poisson.jl:35-50
 *   wdm_reduce           mass / energy diagnostics         src/simulation_main.jl:552-558,
 *                                                          src/visualization/runtime.jl:176-179
 *   wdm_diagnostics      the loop's per-step diagnostics   src/simulation_main.jl:550-558,641-656,
 *                        (Prop.csv columns)                src/visualization/runtime.jl:210-238
 *   wdm_acceleration     grad_central(-dx, Phi_all)        src/simulation_main.jl:652
 *   wdm_density_*        time-averaged density buffer      src/simulation_main.jl:657-660
 *   wdm_ic_velocity      solve_vector_equation + sqrt(rho) exp(i theta)
 *                                                          src/physics/schrodinger.jl:8-30,
 *                                                          src/ic/generation.jl:185-187,330-336
 */
#ifndef WAVEDM_H
#define WAVEDM_H

#ifdef __cplusplus
extern "C" {
#endif

#define WDM_ABI_VERSION 1

/* status codes */
#define WDM_OK 0
#define WDM_ERR_ARG (-1)
#define WDM_ERR_OOM (-2)
#define WDM_ERR_HIP (-3)
#define WDM_ERR_NONFINITE (-4)
#define WDM_ERR_UNSUPPORTED (-5)
#define WDM_ERR_STATE (-6)

/* plan precision */
#define WDM_F64 0 /* complex128 FFTs and phases (the plain form) */
#define WDM_F32 1 /* complex64 FFTs; FP64 phase reduction, FP64 reductions */

/* host/device array element types */
#define WDM_C128 0
#define WDM_C64 1

/* plan flags */
#define WDM_GRID_CONSISTENT 1u /* Delta = 2*Xmax/N instead of the reference 2*Xmax/(N-1) */

/* split modes for wdm_step */
#define WDM_SPLIT_REFERENCE 0 /* K(dt_kick) D(dt) per iteration (src/simulation_main.jl:536-547) */
#define WDM_SPLIT_STRANG 1    /* K(dt/2) D(dt) K(dt/2) with Phi recomputed after the drift */

/* potential models (wdm_set_potential_model kind) */
#define WDM_V_HERNQUIST 2  /* params: M, a, cx, cy, cz          V = -G M / (|x-c| + a) */
#define WDM_V_QUADRUPOLE 4 /* params: Txx,Tyy,Tzz,Txy,Txz,Tyz, nx,ny,nz, period, cx,cy,cz:
                              V = 1/2 (x-c)^T R(t) T R(t)^T (x-c), R(t) rotation about n
                              by 2 pi t / period */
#define WDM_V_KAPPA 8      /* params: kappa                      V = kappa |psi|^2 */
#define WDM_V_TIDAL 16     /* set by wdm_set_tidal_model (not wdm_set_potential_model) */

/* wdm_reduce output slots */
#define WDM_RED_MASS 0    /* sum |psi|^2 dV                                    */
#define WDM_RED_KINETIC 1 /* (hbar/m)^2/2 sum_k k^2 |psi_k|^2 dV / N           */
#define WDM_RED_WSELF 2   /* 1/2 sum |psi|^2 Phi dV                            */
#define WDM_RED_VEXT 3    /* sum |psi|^2 V dV (V at the last kick time)        */
#define WDM_RED_PSIMAX 4  /* max |psi|                                         */
#define WDM_RED_PHIMAX 5  /* max |Phi|                                         */
#define WDM_RED_COUNT 6

typedef struct wdm_plan wdm_plan;

int wdm_version(void);
const char* wdm_last_error(const wdm_plan* plan);

/* Grid nx*ny*nz on [-xmax, xmax] x ...; periodic box length 2*xmax per axis. */
int wdm_plan_create(wdm_plan** plan, int nx, int ny, int nz, double xmax, double ymax, double zmax,
                    int precision, unsigned flags);
int wdm_plan_destroy(wdm_plan* plan);

/* hbar/m (drift and kick phases) and 4 pi G (Poisson source); the reference's
   code units are hbar_over_m = 1, four_pi_G = 4 pi. */
int wdm_set_coeffs(wdm_plan* plan, double hbar_over_m, double four_pi_G);
int wdm_set_self_gravity(wdm_plan* plan, int on);

/* Mass projection (no counterpart in the reference; opt-in, off by default):
   after every step psi is scaled so that sum |psi|^2 keeps the value it had
   when the projection was switched on or psi was last set, removing the
   random walk of FP32 rounding in the norm.  The scale f of a step is
   computed on the device from the inverse x pass's mass partials and applied
   by the next step (its z pass multiplies the FP64 drift factor by f, its
   kick uses Phi f^2); wdm_step and wdm_drift apply the last one to the state
   they return with unbiased stochastic rounding.  The deviation does not
   grow with the number of steps.  Needs self-gravity and no absorber; not
   differentiated by the adjoint. */
int wdm_set_mass_ledger(wdm_plan* plan, int on);
/* The projection's state, for snapshots (no counterpart in the reference,
   which keeps no exact-restart state).  wdm_get_mass_ledger: *on = 1 if the
   projection is on; *m_ref = the reference mass (sum |psi|^2, no cell volume)
   it holds the state to, or NaN while it is not seeded (it is seeded from psi
   by the next kick or step after wdm_set_mass_ledger or a new psi).
   wdm_set_mass_reference: seeds it with m_ref (> 0) instead, so that a run
   restored from a snapshot continues as the saved run would have. */
int wdm_get_mass_ledger(wdm_plan* plan, int* on, double* m_ref);
int wdm_set_mass_reference(wdm_plan* plan, double m_ref);

int wdm_set_psi(wdm_plan* plan, const void* host_psi, int dtype);
int wdm_get_psi(wdm_plan* plan, void* host_psi, int dtype);
/* device-pointer variants (memory from hipMalloc / a torch tensor on this GPU) */
int wdm_set_psi_device(wdm_plan* plan, const void* dev_psi, int dtype);
int wdm_get_psi_device(wdm_plan* plan, void* dev_psi, int dtype);

/* rank 1: length-nx FP64 vector applied along x after each drift (the
   reference's broadcast, src/simulation_main.jl:415).  NULL clears. */
int wdm_set_border(wdm_plan* plan, const double* border, int rank);
/* full-grid FP64 static potential added to Phi in every kick; NULL clears. */
int wdm_set_static_potential(wdm_plan* plan, const double* V);
int wdm_set_potential_model(wdm_plan* plan, int kind, const double* params, int nparams);
int wdm_clear_potential_models(wdm_plan* plan);
/* Milky Way tidal potential added to V in every kick (opt-in; replaces the
   MW_tidal_interpolate branch of compute_tidal_potential,
   src/physics/tidal_fields.jl:64-146, whose result the reference drops, :117).
   table: tnx*tny*tnz FP64 potential (x fastest) on the uniform grid
   lo[a] + i (hi[a] - lo[a]) / (tn_a - 1); traj_t[ntraj] non-decreasing
   kick times and traj_xyz[3*ntraj] the offsets (df_traj rows over uL): the
   kick at time t uses the last row with traj_t <= t (the first before the
   table starts).  Per kick: V_tidal(x) = interp(x + o) minus its value at the
   density peak, minus the ramps of its central differences there, plus the
   rho-weighted mean of those ramps (cancel_field_gradient_at_center!,
   :126-146).  Coordinates outside the table are clamped to it.  The table and
   trajectory are copied; table == NULL clears; wdm_clear_potential_models
   clears it too.  Not differentiated (wdm_adjoint_forward refuses it) and
   time-dependent (ensemble graph mode refuses it). */
int wdm_set_tidal_model(wdm_plan* plan, const double* table, int tnx, int tny, int tnz, const double* lo,
                        const double* hi, const double* traj_t, const double* traj_xyz, int ntraj);

/* phi_out = scale * lap^-1 (rho - <rho>) on the plan's periodic lattice
   (k = 2 pi m / (2 xmax), k = 0 -> 0).  rho == NULL: rho = |psi|^2 of the
   plan's psi.  phi_out == NULL: keep the result on the device for the next
   kick (scale must then be 4 pi G, pass 0 to use the plan's coefficient). */
int wdm_poisson(wdm_plan* plan, const double* rho, double* phi_out, double scale);
/* real-space Phi (4 pi G lap^-1 (|psi|^2 - mean)) of the current psi, FP64 */
int wdm_get_phi(wdm_plan* plan, double* host_phi);

/* apply_kick_step!: psi <- FFT[exp(-i (m/hbar) dt_kick (Phi + V(t))) psi];
   afterwards the plan holds the spectrum (get_psi returns it). */
int wdm_kick(wdm_plan* plan, double dt_kick, double t);
/* apply_drift_step!: psi <- border .* IFFT[exp(-i (hbar/m) k^2 dt/2) psi_k] */
int wdm_drift(wdm_plan* plan, double dt);
/* nsteps iterations of the fused step; kick i happens at t0 + i*dt. */
int wdm_step(wdm_plan* plan, int nsteps, double dt, double dt_kick, double t0, int split_mode);

/* out[0..min(nout,WDM_RED_COUNT)) = WDM_RED_* diagnostics of the current psi */
int wdm_reduce(wdm_plan* plan, double* out, int nout);
/* out3 = {2 dx^2/(pi hbar/m), 2 pi/((m/hbar) max|Phi|), 2 pi/((m/hbar)|kappa| max|psi|)} */
int wdm_cfl(wdm_plan* plan, double kappa, double* out3);

int wdm_synchronize(wdm_plan* plan);

/* wdm_step with HIP events around every kernel of the step on the plan's
   stream: pass_ms[i] (i < min(npass, WDM_PASS_COUNT)) receives the summed
   kernel time of pass i over the nsteps steps, in milliseconds; total_ms the
   event time from the first to the last launch.  Reference split only. */
#define WDM_PASS_XKICK 0    /* x: Phi C2R + kick + forward FFT_x          */
#define WDM_PASS_YFWD 1     /* y: FFT_y                                   */
#define WDM_PASS_ZDRIFT 2   /* z: FFT_z * drift * IFFT_z                  */
#define WDM_PASS_YINV 3     /* y: IFFT_y                                  */
#define WDM_PASS_XINV 4     /* x: IFFT_x + absorber + density R2C_x       */
#define WDM_PASS_PYFWD 5    /* Poisson y forward (half spectrum)          */
#define WDM_PASS_PZ 6       /* Poisson z: FFT_z * -4 pi G/k^2 * IFFT_z    */
#define WDM_PASS_PYINV 7    /* Poisson y inverse                          */
#define WDM_PASS_COUNT 8
int wdm_step_timed(wdm_plan* plan, int nsteps, double dt, double dt_kick, double t0, double* pass_ms,
                   int npass, double* total_ms);
/* The same with events only on the passes whose bit (1u << WDM_PASS_*) is set
   in pass_mask: each event pair costs a few microseconds between kernels, so a
   timed region that needs one kernel's duration sets only that bit. */
int wdm_step_timed_mask(wdm_plan* plan, int nsteps, double dt, double dt_kick, double t0, unsigned pass_mask,
                        double* pass_ms, int npass, double* total_ms);

/* Initial condition on the device (FP64 evaluation, stored in the plan
   precision): psi = sum_i sqrt(rho_c,i / (1 + 0.091 (r_i/r_c,i)^2)^8)
   * exp(i (phase_i + k_i . x)), r_i the minimum-image distance to the
   centre in the periodic box.  params: nsol rows of
   {x0, y0, z0, r_c, rho_c, phase, kx, ky, kz}.  Replaces evaluating the IC
   function on the host grid (src/physics/schrodinger.jl:37-46). */
int wdm_ic_solitons(wdm_plan* plan, const double* params, int nsol);

/* Initial condition from a density and a velocity field, all host arrays of
   nx*ny*nz FP64 values, x-fastest: theta = the reference's
   solve_vector_equation (src/physics/schrodinger.jl:8-30; three cumulative
   trapezoid line integrals, each from the first cell of its axis, with the
   plan's dx), psi = sqrt(rho) exp(i theta / (hbar/m)) (src/ic/generation.jl:
   185-187).  remove_bulk = 1 first subtracts the rho-weighted mean velocity
   (src/ic/generation.jl:330-336). */
int wdm_ic_velocity(wdm_plan* plan, const double* rho, const double* vx, const double* vy, const double* vz,
                    int remove_bulk);

/* Adjoint SSFM (reference split).  No counterpart in the reference, which
   imports Zygote (src/WaveDM.jl:42) but defines no rule for the step; this is
   what a ChainRulesCore rrule around wdm_step calls.

   wdm_adjoint_forward: nsteps fused steps from the current psi (as wdm_step
   with split WDM_SPLIT_REFERENCE), keeping a checkpoint of psi every
   ckpt_every steps on the device.
   wdm_adjoint_backward: given grad_psi = dL/dRe(psi_N) + i dL/dIm(psi_N) on
   the host (dtype WDM_C128/WDM_C64), returns grad_psi0 (same packing and
   dtype; NULL to skip) and grad_params[0..nparams):
     0..4  dL/dM, dL/da, dL/dcx, dL/dcy, dL/dcz of the Hernquist model
     5     dL/dsigma for dt -> sigma dt, dt_kick -> sigma dt_kick at sigma = 1
           (only with WDM_ADJ_DT_SCALE, else NaN; needs a time-independent V)
     6..11 dL/dTxx, dL/dTyy, dL/dTzz, dL/dTxy, dL/dTxz, dL/dTyz of the
           rotating quadrupole's T0 (the tidal tensor; 0 without the model)
   Segments of ckpt_every steps are recomputed during the backward sweep.
   Self-gravity, static V, Hernquist, quadrupole(t) and kappa|psi|^2 are
   differentiated; psi is left at psi_N. */
#define WDM_ADJ_DT_SCALE 1u
#define WDM_ADJ_NPARAM 12
int wdm_adjoint_forward(wdm_plan* plan, int nsteps, double dt, double dt_kick, double t0, int ckpt_every);
int wdm_adjoint_backward(wdm_plan* plan, const void* grad_psi, int dtype, void* grad_psi0, double* grad_params,
                         int nparams, unsigned flags);

/* Autotuner.  No counterpart in the reference.  Each axis length has up to
   WDM_MAX_VARIANTS compiled pass-table variants (0: default factorisation and
   workgroup size; 1: other workgroup size; 2: alternative factorisation,
   lengths 64, 128, 192, 256).  wdm_autotune times trial_steps reference-split
   steps of dt per variant on a copy of the plan's psi (the state is kept) and
   selects the fastest per axis; variants_out[3] / ms_out[3] (may be NULL)
   receive the choice and its ms per step for the axis's passes.
   wdm_select_variants sets the three variants directly. */
#define WDM_MAX_VARIANTS 3
int wdm_autotune(wdm_plan* plan, int trial_steps, double dt, int* variants_out, double* ms_out);
int wdm_select_variants(wdm_plan* plan, const int* variants);

/* Per-step diagnostics of the reference loop, evaluated on the device from the
   plan's current psi (src/simulation_main.jl:550-558,641-656; the virial terms
   and momenta of src/visualization/runtime.jl:210-238; the Prop.csv columns of
   src/io/save.jl:41-72).  rho = |psi|^2, dV = dx dy dz, Phi_all = self-gravity
   Phi of the current psi + the external potential (static, Hernquist,
   quadrupole at the last kick time; not the kappa term).  Sums marked "int"
   run over the interior cells 1..n-2 of every axis (0-based).
   out[0..min(nout, WDM_DIAG_COUNT)); asking for fewer than
   WDM_DIAG_PE_ABS + 1 slots skips computing Phi. */
#define WDM_DIAG_MASS 0      /* TotalMass = sum rho dV                                  */
#define WDM_DIAG_RHOMAX 1    /* max rho                                                 */
#define WDM_DIAG_PEAK_IX 2   /* 0-based cell of the first maximum (x-fastest order)     */
#define WDM_DIAG_PEAK_IY 3
#define WDM_DIAG_PEAK_IZ 4
#define WDM_DIAG_R 5         /* std(|x| on plane z = nz/2-1, weights rho), uncorrected  */
#define WDM_DIAG_R1 6        /* R1..R9: rho-weighted quantiles 0.1..0.9 of |x - x_peak| */
#define WDM_DIAG_MOMX 15     /* sum Im(conj psi d psi/dx) dV (central differences,      */
#define WDM_DIAG_MOMY 16     /*   periodic neighbours on the faces), likewise y, z      */
#define WDM_DIAG_MOMZ 17
#define WDM_DIAG_PE_ABS 18   /* -1/2 sum_int rho Phi_all dV                             */
#define WDM_DIAG_KE 19       /* 1/2 sum_int |Im(conj psi grad psi)|^2 / rho dV          */
#define WDM_DIAG_QE 20       /* 1/2 sum_int |grad |psi||^2 dV                           */
#define WDM_DIAG_VIRIAL 21   /* 2 KE + 2 QE - PE_abs                                    */
#define WDM_DIAG_COUNT 22
/* The radii come from a weighted histogram of |x - x_peak| with WDM_DIAG_BINS
   bins over [0, distance to the farthest corner], linear inside the bin that
   holds the quantile; the reference sorts every cell (StatsBase quantile). */
#define WDM_DIAG_BINS 4096
int wdm_diagnostics(wdm_plan* plan, double* out, int nout);
/* -grad Phi_all by central differences (grad_central(-dx, -dy, -dz, Phi_all),
   src/simulation_main.jl:652; periodic neighbours on the faces) into three
   host arrays of nx*ny*nz doubles, x-fastest. */
int wdm_acceleration(wdm_plan* plan, double* host_ax, double* host_ay, double* host_az);

/* Time-averaged density on the device: the loop's buffer_psi2 += rho and
   averaged_psi2 = buffer_psi2 / average_N (src/simulation_main.jl:657-660,
   src/io/save.jl:79-81).  wdm_density_accumulate adds |psi|^2 of the current
   psi to an FP64 device sum; wdm_density_average writes sum / count to a host
   array of nx*ny*nz doubles (host_out may be NULL to read only *count);
   wdm_density_reset clears the sum.  wdm_density_sum reads the FP64 sum
   itself and wdm_density_restore replaces it (count > 0) -- the exact state a
   snapshot needs to continue the average (the reference restarts it). */
int wdm_density_accumulate(wdm_plan* plan);
int wdm_density_average(wdm_plan* plan, double* host_out, long long* count);
int wdm_density_reset(wdm_plan* plan);
int wdm_density_sum(wdm_plan* plan, double* host_sum, long long* count);
int wdm_density_restore(wdm_plan* plan, const double* host_sum, long long count);

/* The acceleration field of an initial condition (src/ic/generation.jl:246-291,
   periodic gravity): Phi = 4 pi G lap^-1 (rho_halo - <rho_halo>) on the
   plan's lattice (compute_acceleration_field), a = -grad Phi by central
   differences (periodic on the faces), and a_CDM = |a| on the interior cells,
   0 on the faces and at the 1-based cell (div(N,2), div(N,2), div(N,2))
   (generate_velocity_field).  rho_halo and every output are host arrays of
   nx*ny*nz doubles, x fastest; outputs may be NULL.  The plan's psi is kept;
   its Poisson buffer is overwritten (the next kick recomputes Phi). */
int wdm_ic_acceleration(wdm_plan* plan, const double* rho_halo, double* phi, double* ax, double* ay, double* az,
                        double* a_cdm);

/* Allocation-free Poisson solve.  No counterpart in the reference: it is for
   callers that cannot allocate device memory (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  The variant of
   wdm_poisson(plan, rho, phi, scale) on caller-owned device buffers, with no
   plan: phi_dev = scale * lap^-1 (rho_dev - <rho_dev>) on the periodic
   lattice of period 2*{x,y,z}max (k = 0 -> 0), both arrays nx*ny*nz FP64,
   x fastest; the arithmetic in `precision`, with the same kernels and tables
   as a fresh plan's wdm_poisson (equal results).  scale must be nonzero.
   Its work memory (half spectrum + tables) is a file-scope device array of
   WDM_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a grid
   that needs more is refused (WDM_ERR_UNSUPPORTED).  The call allocates
   nothing, is asynchronous on `stream` (a hipStream_t, NULL = the default
   stream) and leaves the result ready when the stream is; calls that overlap
   on one GPU share the array and race.  Errors: wdm_last_error(NULL). */
#define WDM_NOALLOC_SCRATCH_BYTES (64u << 20)
int wdm_poisson_noalloc(int nx, int ny, int nz, double xmax, double ymax, double zmax, int precision,
                        const double* rho_dev, double* phi_dev, double scale, void* stream);

/* Batched ensemble of independent plans (uneven grids; no counterpart in the
   reference, which steps one box at a time).  The plans stay owned by the
   caller and must outlive the ensemble.  wdm_ensemble_step runs nsteps
   reference-split steps of every member (member i with dt[i], dt_kick[i],
   kick k at t0 + k dt[i]).  Members of one shape, precision, pass table,
   gravity, absorber and potential kind (without the mass projection or a
   time-dependent potential) form a batch: each pass is one launch over all
   of them; the others step on their own streams.  Every member's result is
   the same, bit for bit, as stepping it alone.  use_graph = 1: one step of
   all members is captured once into a HIP graph and launched once per step;
   needs time-independent potentials.  use_graph = 0: the same launches issued
   eagerly.  wdm_ensemble_autotune tunes each
   distinct (shape, precision) once and gives the choice to every member of
   that shape.  Errors: wdm_last_error(NULL). */
typedef struct wdm_ensemble wdm_ensemble;
int wdm_ensemble_create(wdm_ensemble** ens, wdm_plan* const* plans, int nplans);
int wdm_ensemble_destroy(wdm_ensemble* ens);
int wdm_ensemble_step(wdm_ensemble* ens, int nsteps, const double* dt, const double* dt_kick, double t0,
                      int use_graph);
int wdm_ensemble_autotune(wdm_ensemble* ens, int trial_steps, const double* dt);
/* How the last wdm_ensemble_step ran: *nbatches groups of >= 2 members that
   share one launch per pass, *nsolo members stepped on their own streams. */
int wdm_ensemble_layout(wdm_ensemble* ens, int* nbatches, int* nsolo);

#ifdef __cplusplus
}
#endif
#endif /* WAVEDM_H */
