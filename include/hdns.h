/*
 * hdns -- MI355X-native DNS time step of GBGHomogeneousTurbulence.jl behind a plain C ABI.
 *
 * One context owns all device memory of one run (one grid, one precision, one HIP stream).
 * Host arrays cross the boundary only in the reference's own layouts:
 *   spectral vector  Complex{T}[N/2+1, N, N, 3] column-major  == C order [3][N][N][N/2+1] complex,
 *   spectral scalar  Complex{T}[N/2+1, N, N]                  == C order [N][N][N/2+1] complex,
 *   physical vector  T[N, N, N, 3]                            == C order [3][N][N][N] real,
 *   tracer positions T[M, 3] (a PointSet's matrix)            == C order [3][M] real,
 * always as FP64 (complex interleaved re, im) with the unnormalised-forward convention of the
 * reference (docs/src/validation.md:73-74).  A FP32-bulk context rounds to FP32 on upload.
 *
 * Every function returns an hdns_status; hdns_last_error(ctx) names the failure.
 * Calls on one context are not re-entrant; all synchronisation happens inside calls that
 * return host data.  Reference file:line citations are relative to the reference repository.
 */
#ifndef HDNS_H
#define HDNS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  HDNS_OK = 0,
  HDNS_ERR_INVALID_ARGUMENT = 1, /* the reference's ArgumentError class */
  HDNS_ERR_STATE = 2,            /* the reference's ErrorException for schedule/epoch errors */
  HDNS_ERR_NONFINITE = 3,        /* blow-up guard (src/SimulationRun.jl:1216-1224) */
  HDNS_ERR_DEVICE = 4,           /* a HIP runtime failure */
  HDNS_ERR_UNSUPPORTED = 5       /* a configuration this build does not implement */
} hdns_status;

enum { HDNS_FP32_BULK = 0, HDNS_FP64 = 1 };
/* IFLowStorageRK3, IFLowStorageRK4 (src/Propagator.jl:264-320), IFSSPRK3 (:596-691) and IFHeun (:420-544) */
enum { HDNS_IF_LOW_STORAGE_RK3 = 3, HDNS_IF_LOW_STORAGE_RK4 = 4, HDNS_IF_SSP_RK3 = 5, HDNS_IF_HEUN = 6 };
/* TwoThirds, HouLi and PhaseShift dealiasing (src/SimulationDomain.jl:44-195) */
enum { HDNS_TWO_THIRDS = 0, HDNS_HOU_LI = 1, HDNS_PHASE_SHIFT = 2 };
/* NoForcing, ConstantPower, Linear, VarianceController, VarianceRescaling and StochasticOU
 * (src/Forcing.jl:464, :714-719, :600-642, :817-871, :566-597, :1120-1413) */
enum { HDNS_NO_FORCING = 0, HDNS_CONSTANT_POWER = 1, HDNS_LINEAR = 2, HDNS_VARIANCE_CONTROLLER = 3,
       HDNS_VARIANCE_RESCALING = 4, HDNS_STOCHASTIC_OU = 5 };
/* VarianceRescaling's `over`: hold the band's kinetic energy, or the whole field's */
enum { HDNS_OVER_BAND = 0, HDNS_OVER_TOTAL = 1 };
enum { HDNS_CATMULL_ROM = 0 };
/* the particle schemes (src/ParticlePropagator.jl:48-253): Heun = ABM(1), ABM(k) k = 1..7, AB(k) k = 1..8 */
enum { HDNS_HEUN = 0, HDNS_ABM = 1, HDNS_AB = 2 };
enum { HDNS_FIELD_VELOCITY = 0, HDNS_FIELD_SCALAR = 1, HDNS_FIELD_VORTICITY = 2 };

/* Everything a run fixes at construction: FixedDomain (src/SimulationDomain.jl:295-328),
 * FlowConditions (src/Generator.jl:81-147), the velocity forcings (src/Forcing.jl:566-1117),
 * the scheme (src/Propagator.jl:264-320, :456-691), and the tracer set (src/Particles.jl:135-176). */
typedef struct {
  int32_t device;            /* HIP device ordinal */
  int32_t grid_points;       /* N (the box is N^3): a power of two in 16..1024, or 48, 96, 192 */
  double side_length;        /* L */
  int32_t precision;         /* HDNS_FP32_BULK or HDNS_FP64 */
  int32_t scheme;            /* HDNS_IF_LOW_STORAGE_RK3 / _RK4, HDNS_IF_SSP_RK3, HDNS_IF_HEUN */
  int32_t dealiasing;        /* HDNS_TWO_THIRDS, HDNS_HOU_LI or HDNS_PHASE_SHIFT (N in 16..256 but 96 and 192) */
  int32_t has_scalar;        /* evolve the passive scalar */
  double viscosity;          /* nu > 0 */
  double schmidt;            /* Sc > 0 */
  int32_t velocity_forcing;  /* HDNS_NO_FORCING, _CONSTANT_POWER, _LINEAR, _VARIANCE_CONTROLLER, _VARIANCE_RESCALING */
  double injection_rate;     /* ConstantPower epsilon, or the Linear rate A (src/Forcing.jl:600-642) */
  int32_t lowest_shell;      /* >= 1 */
  int32_t highest_shell;     /* >= lowest_shell; -1 means Inf (every mode but the mean) */
  double mean_gradient;      /* G of an imposed mean scalar gradient along z (extension) */
  int64_t tracer_capacity;   /* 0: no tracers */
  int32_t interpolation;     /* HDNS_CATMULL_ROM */
  int32_t particle_scheme;   /* HDNS_HEUN, HDNS_ABM or HDNS_AB, with particle_steps k */
  int32_t store_physical;    /* bit 0: keep physical u, bit 1: keep physical phi, bit 2: keep physical w */
  int32_t forcing_over;      /* VarianceRescaling: HDNS_OVER_BAND or HDNS_OVER_TOTAL */
  double forcing_target;     /* VarianceController k_inf > 0, or VarianceRescaling's target >= 0 */
  double forcing_gain;       /* VarianceController G >= 0 (0: the marginal law) */
  double forcing_time;       /* VarianceController turnover time t_l > 0 */
  int32_t particle_steps;    /* k of ABM(k) (1..7) or AB(k) (1..8); ignored for HDNS_HEUN */
  int32_t unwrapped;         /* 1: also track the unwrapped positions (src/Particles.jl:131-138); needs tracers */
  /* the scalar forcing (src/Generator.jl:652-653): the velocity forcings' strategies with the
   * variance factor 1 and the diffusivity nu/Sc (src/Forcing.jl:456); needs has_scalar */
  int32_t scalar_forcing;        /* HDNS_NO_FORCING, _CONSTANT_POWER, _LINEAR, _VARIANCE_CONTROLLER, _VARIANCE_RESCALING */
  int32_t scalar_lowest_shell;   /* >= 1 */
  int32_t scalar_highest_shell;  /* >= scalar_lowest_shell; -1 means Inf */
  int32_t scalar_over;           /* VarianceRescaling: HDNS_OVER_BAND or HDNS_OVER_TOTAL */
  double scalar_rate;            /* ConstantPower's injection rate, or the Linear rate */
  double scalar_target;          /* VarianceController's or VarianceRescaling's target */
  double scalar_gain;            /* VarianceController G >= 0 */
  double scalar_time;            /* VarianceController turnover time > 0 */
  /* HouLi(order, strength) (src/SimulationDomain.jl:131-138, :365-376): the filter
   * exp(-strength (|k|/k_Ny)^order) on the nonlinear terms; strength NaN is the reference's default
   * -log(eps(T)) of the element type */
  int32_t filter_order;          /* >= 1 */
  double filter_strength;
} hdns_config;

typedef struct hdns_ctx hdns_ctx;

/* Build the frame, the registers and the device state.  Replaces the constructors
 * FixedDomain (src/SimulationDomain.jl:295-328), allocate_buffers (src/Propagator.jl:709-739),
 * ResolvedGenerator (src/Generator.jl:288-345) and ObservablesCache (src/Cache.jl:150-173)
 * for the device-resident part of a Simulation (src/SimulationRun.jl:750-931). */
hdns_status hdns_ctx_create(const hdns_config *config, hdns_ctx **out);
hdns_status hdns_ctx_destroy(hdns_ctx *ctx);
/* A batch: nbox boxes of one size that differ in viscosity (viscosities[nbox]; NULL: the
 * config's for all), stored back to back and advanced by the same launches (every kernel's grid
 * carries a box dimension).  hdns_step, hdns_step_graph, hdns_advect, hdns_deposit_tracers and
 * hdns_rebuild_traversal act on every box; the per-box entry points (state upload/download,
 * reductions, max speed, spectra, physical fields, tracers) address the box chosen by
 * hdns_select_box (box 0 at creation).  Every box carries the same number of tracers.  No
 * counterpart in the reference, which advances one Simulation at a time; it is the building block
 * of the ensemble below. */
hdns_status hdns_ctx_create_batch(const hdns_config *config, int32_t nbox, const double *viscosities,
                                  hdns_ctx **out);
hdns_status hdns_select_box(hdns_ctx *ctx, int32_t box);
/* Message of the last failure on this context (or of the last failed create when ctx is NULL). */
const char *hdns_last_error(const hdns_ctx *ctx);
/* Device bytes the context holds. */
hdns_status hdns_memory_bytes(const hdns_ctx *ctx, int64_t *bytes);

/* Adopt a state (the door Simulation(domain; state = ResolvedFields(u_hat, phi_hat)),
 * src/SimulationRun.jl:769-788).  The state must carry nothing outside the 2/3 ball
 * (truncate_to_band!, src/SimulationDomain.jl:520-537); content there is an
 * HDNS_ERR_INVALID_ARGUMENT.  phi_hat may be NULL for a context without a scalar.
 * Ends with the field deposit of the adopted state (src/SimulationRun.jl:923). */
hdns_status hdns_state_upload(hdns_ctx *ctx, const double *u_hat, const double *phi_hat);
/* Read the state back in the reference layout (zeros outside the retained modes). */
hdns_status hdns_state_download(hdns_ctx *ctx, double *u_hat, double *phi_hat);

/* One step!(scheme, state, generator, domain, buffers, cache, dt) followed by the field
 * deposit! (src/Propagator.jl:219,304-419,808-842): every stage, the state Leray and reality
 * projections, and the recovery of u, w, phi and max|u| at the new state. */
hdns_status hdns_step(hdns_ctx *ctx, double dt);
/* The same step as hdns_step, launched as one HIP graph captured from its kernel sequence on the
 * first call (and again after a schedule change), so a step costs two launches: a one-thread kernel
 * that writes this step's dt and substeps into device memory, which the captured kernels read,
 * and the graph.  One capture therefore serves every dt of an adaptive run.  The step holds no host
 * readback and no memset node, so the replay is bitwise the eager step. */
hdns_status hdns_step_graph(hdns_ctx *ctx, double dt);
/* The field deposit! of the current state on its own (src/Propagator.jl:808-842). */
hdns_status hdns_deposit(hdns_ctx *ctx);
/* max|u| of the last deposit (src/Observables.jl:304-305, src/Cache.jl:243). */
hdns_status hdns_max_speed(hdns_ctx *ctx, double *out);
/* out[0..4] = k, eps, <phi^2>, eps_phi, band energy of the last deposit's state, in FP64 from
 * compensated device reductions (src/Observables.jl:84,155-158,345-400; src/Forcing.jl:1109). */
hdns_status hdns_reductions(hdns_ctx *ctx, double *out);
/* The distance between the states of two contexts on one grid and device (their selected boxes;
 * the element types may differ): out[0] = rms |u_a - u_b|, out[1] = rms |u_b|, out[2] =
 * rms |phi_a - phi_b|, out[3] = rms |phi_b| (0 without a scalar), over the grid points, in FP64
 * (compensated sums).  No reference counterpart: it is the FP32-bulk error report against an FP64
 * run of the same flow, which the reference's tests form on host arrays between two runs
 * (test/box_scaling.jl:18,36).  Waits for b's stream and runs on a's. */
hdns_status hdns_state_distance(hdns_ctx *a, hdns_ctx *b, double *out);
/* The velocity-derivative skewness of the selected box's current state,
 * S = C * 3 <det S> / <S:S>^(3/2) with C = (4/7) sqrt(15/2) and S = sym grad u
 * (src/Observables.jl:406,465-477): the strain determinant on the device, summed in FP64;
 * <S:S> = (1/2) <|w|^2> from the deposit's reductions; 0 for a strain-free field.  Reuses the
 * pencil buffers and redoes the deposit. */
hdns_status hdns_derivative_skewness(hdns_ctx *ctx, double *out);
/* Shell spectrum E(kappa_n), n = 0..nshells-1 (src/Observables.jl:502-514); which = velocity or
 * scalar.  nshells must be >= floor(sqrt(3)*N/2)+2.  The FP64 shell sums run in a fixed order, so
 * two calls on one state agree to the last bit. */
hdns_status hdns_spectrum(hdns_ctx *ctx, int32_t which, double *out, int32_t nshells);
/* hdns_spectrum without any allocation: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the same kernels and the same output, bit for bit, with the partial sums in a
 * file-scope scratch array of 272 KiB per device (at most 2048 shells, i.e. N <= 2362; a larger
 * grid is an HDNS_ERR_INVALID_ARGUMENT).  Every call on one device shares that device's array:
 * calls that overlap on one GPU, from any context or thread, race. */
hdns_status hdns_spectrum_noalloc(hdns_ctx *ctx, int32_t which, double *out, int32_t nshells);
/* A deposited physical field in the reference layout (velocity_physical / scalar_physical /
 * vorticity_physical, src/Cache.jl:203-240); the context must keep it (store_physical). */
hdns_status hdns_field_download(hdns_ctx *ctx, int32_t which, double *out);

/* Tracers (src/Particles.jl:135-176): set M <= tracer_capacity positions, layout [3][M]. */
hdns_status hdns_tracers_upload(hdns_ctx *ctx, int64_t m, const double *positions);
hdns_status hdns_tracers_download(hdns_ctx *ctx, double *positions);
/* unwrapped_positions(tracers) (src/Particles.jl:317-323): the unwrapped record, [3][M], which
 * hdns_tracers_upload starts equal to the positions and every push advances by the same
 * displacement terms without the wrap (src/ParticlePropagator.jl:299,313).  Fails on a context
 * built without unwrapped = 1. */
hdns_status hdns_tracers_unwrapped_download(hdns_ctx *ctx, double *positions);
/* Set the unwrapped record of the current M tracers (read_tracers, ext/GBGHomogeneousTurbulenceHDF5Ext.jl:236-241),
 * after hdns_tracers_upload has set the positions. */
hdns_status hdns_tracers_unwrapped_upload(hdns_ctx *ctx, int64_t m, const double *positions);
/* rebuild_traversal! (src/Particles.jl:268-271): the tracers sorted by 8^3-cell tile, and within a
 * tile by index (a stable sort), so the order depends on the positions alone. */
hdns_status hdns_rebuild_traversal(hdns_ctx *ctx);
/* advect!(scheme, ...) (src/ParticlePropagator.jl:287-372): the Adams predictor over the trail's
 * newest J samples with weights from their actual times, and for ABM (or while the trail holds
 * fewer samples than the scheme integrates over) the gather at X* and the corrector; wrap.  Needs
 * the deposited velocity one step ahead of the positions and the newest sample at the positions. */
hdns_status hdns_advect(hdns_ctx *ctx, double dt);
/* Particle deposit! (src/ParticlePropagator.jl:378-386): sample u (and phi) at the positions.
 * After an advect the sample is new, at the newest sample's time plus that step's dt (the first
 * sample after an upload is at t = 0); otherwise it replaces the newest sample. */
hdns_status hdns_deposit_tracers(hdns_ctx *ctx);
/* hdns_advect(dt) then hdns_deposit_tracers() -- an iteration's advect! and particle deposit!
 * (src/SimulationRun.jl:1203-1207) -- as one launch: the sample is taken at the new positions while
 * the push holds them, so its gathers find the field lines the corrector just read.  Results, trail
 * and times are bit for bit those of the two calls. */
hdns_status hdns_advect_deposit(hdns_ctx *ctx, double dt);
/* The run's velocity trail (src/Particles.jl:404-526): *filled samples, their times newest first
 * (times[capacity], capacity = k + 1 for ABM(k), k for AB(k), 2 for Heun). */
hdns_status hdns_tracer_trail(hdns_ctx *ctx, int32_t *filled, int32_t *capacity, double *times);
/* Restore a stored trail (read_simulation, ext/GBGHomogeneousTurbulenceHDF5Ext.jl:535-544): the n
 * earlier samples [n][3][M] newest first with their times, n <= capacity - 1; the next
 * hdns_deposit_tracers measures the newest sample again at the positions, at time t_next. */
hdns_status hdns_tracer_trail_restore(hdns_ctx *ctx, int32_t n, const double *times, const double *samples,
                                      double t_next);
/* The velocity sample `offset` steps back from the newest (offset <= 0), [3][M]: trail[offset]. */
hdns_status hdns_tracer_trail_sample(hdns_ctx *ctx, int32_t offset, double *velocity);
/* kinetic_energy_and_dissipation_below_scale (src/Observables.jl:195-264) of the selected box's
 * current state: for every side l = n L / N (n = 1..N), scales[n-1] = l, the kinetic energy and the
 * dissipation below l, k(l) = (k_all - sum w W_l^2 |u|^2) / 2 and eps(l) = nu (sum w |k|^2 |u|^2 -
 * sum w |k|^2 W_l^2 |u|^2), W_l(k) = prod_a sinc(k_a l / 2), FP64 sums.  An occasional diagnostic:
 * O(N^4) work. */
hdns_status hdns_below_scale(hdns_ctx *ctx, double *scales, double *kinetic_energy, double *dissipation);
/* The additive band force of a StochasticOU forcing for the next steps (src/Forcing.jl:1404-1405,
 * add_forcing_source!): field 0 (velocity, [3][2r+1][2r+1][r+1]) or 1 (scalar, [2r+1][2r+1][r+1])
 * of the selected box, complex, in the unnormalised spectral convention of the run, with r the
 * forcing's highest shell: entry [c][mz + r][my + r][mx] is mode (mx, my, mz).  Modes outside the
 * band are ignored.  The force is held over a step and added at every stage; the host advances the
 * Ornstein-Uhlenbeck process once per step and uploads it before the step. */
hdns_status hdns_band_source_upload(hdns_ctx *ctx, int32_t field, const double *source);
/* time_derivative!(out, trail; order) (src/Particles.jl:597-631): d^m v/dt^m at the newest
 * sample's time of the polynomial through every sample held, Fornberg weights from the times;
 * order 1 is the Lagrangian acceleration.  out [3][M]. */
hdns_status hdns_tracer_time_derivative(hdns_ctx *ctx, int32_t order, double *out);
/* The newest velocity sample [3][M] and, when the run carries a scalar, phi at the tracers [M]. */
hdns_status hdns_tracer_samples(hdns_ctx *ctx, double *velocity, double *scalar);
/* Canonical cell key (k*N + j)*N + i of every tracer, computed on the device. */
hdns_status hdns_tracer_cells(hdns_ctx *ctx, int64_t *keys);
/* The traversal permutation built by the last rebuild (0-based), for inspection. */
hdns_status hdns_traversal_order(hdns_ctx *ctx, int32_t *order);

/* The frame's integer tables as the device evaluates them: the signed wavenumber index of every
 * bin of the three axes (lengths N/2+1, N, N) and the 2/3 mask over the reference layout
 * [N][N][N/2+1] (src/SimulationDomain.jl:342-376). */
hdns_status hdns_frame_tables(hdns_ctx *ctx, int32_t *mx, int32_t *my, int32_t *mz, uint8_t *mask);

/* The low-storage coefficients a, b, d (d: the substep increments) of a scheme rounded to the
 * precision's element type, exactly as a step uses them (src/Propagator.jl:264-274, :304-320).
 * Arrays hold at least 5 entries; *stages receives 3 or 5.  Needs no device. */
hdns_status hdns_scheme_coefficients(int32_t scheme, int32_t precision, double *a, double *b, double *d,
                                     int32_t *stages);
/* The Heun push weights for a step h: out[0] = b1 (predictor), out[1] = beta1 (on the predicted
 * sample), out[2] = beta2 - b1 (correction on the previous sample), from the Adams weights by
 * 4-point Gauss-Legendre (src/ParticlePropagator.jl:261-315).  Needs no device. */
hdns_status hdns_heun_weights(double h, double *out3);

/* Autotuner: time every launch schedule of each kernel family on this context's own buffers,
 * keep the fastest of each, and redo the deposit.  The families and their schedules:
 *   y passes: 256/512 threads (512/1024 from N = 512 on);
 *   z forward: (0) split scalar and velocity kernels with LDS-staged products, (1) one fused
 *              kernel, (2) schedule 0 with the scalar half's flux tiles streamed through LDS;
 *   z inverse: (0) full or (1) half-width tiles, (2) full tiles written back 16 bytes per lane
 *              through LDS, (3) FP32 transforms on column pairs;
 *   x pass: 256/512 threads, or with/without prefetch;
 *   y radix plan: the default plan, or one led by a radix-16 pass.
 * The timed sequence leaves the state bitwise unchanged.  No counterpart in the reference (whose
 * field library has no launch shapes); it tunes the kernels behind hdns_step.  On an ensemble's
 * member context it tunes that box's batch (every box of its size, as the ensemble step launches
 * them); the ensemble's next graph step captures the new schedule.
 * ms[HDNS_SCHED_FAMILIES * HDNS_SCHED_VARIANTS] (optional) receives the average ms of each
 * family's schedules, row by row in the family order (-1 where a family has fewer);
 * chosen[HDNS_SCHED_FAMILIES] (optional) the picks. */
#define HDNS_SCHED_FAMILIES 5
#define HDNS_SCHED_VARIANTS 4
hdns_status hdns_autotune(hdns_ctx *ctx, int32_t reps, double *ms, int32_t *chosen);
/* Set (set != NULL) and/or read (get != NULL) the schedule: five entries, in the order y passes
 * (0-1), z forward (0-2), z inverse (0-3), x pass (0-1), y radix plan (0-1). */
hdns_status hdns_schedule(hdns_ctx *ctx, const int32_t *set, int32_t *get);

/* The launch of the tracer gathers (hdns_advect, hdns_deposit_tracers, hdns_advect_deposit): 1 puts
 * them on 128-thread blocks at most four to a CU, so fewer 4^3 stencils are in flight per XCD and
 * more of their field lines stay in L2; 0 on 256-thread blocks at full occupancy; -1 (the default)
 * chooses 1 for FP64 from 2^16 tracers per launch on, else 0; -2 leaves the setting.  *get (optional) receives the
 * launch in force.  The results do not depend on it.  No counterpart in the reference. */
hdns_status hdns_tracer_schedule(hdns_ctx *ctx, int32_t set, int32_t *get);

/* Per-kernel timing with HIP events on the context's stream, for the roofline report.  Kinds:
 * 0 y forward, 1 z forward + stage update, 2 reality projection, 3 z inverse + reductions,
 * 4 reduction finalize, 5 y inverse, 6 x pass (products), 7 other.  read: ms[8] accumulated
 * milliseconds and launches[8] counts since the last reset (synchronises the stream). */
hdns_status hdns_profile_enable(hdns_ctx *ctx, int32_t on);
hdns_status hdns_profile_read(hdns_ctx *ctx, double *ms, int64_t *launches, int32_t reset);

/* The adjoint of the step (no counterpart in the reference, which has no derivative: the north-star
 * rrule of step!).  hdns_step_vjp applies the transpose of one step's derivative at the context's
 * current state: lam_*_out (the adjoint of the step's output, full half-spectrum layout like
 * hdns_state_download) becomes lam_*_in (that of its input); the state is unchanged.  Adjoints are
 * gradients under the real inner product of the half spectrum with the Parseval weights (1 on
 * kx = 0, 2 inside), which equals the full-spectrum product.  hdns_adjoint_gradient runs `steps`
 * steps from the current state, keeping a checkpoint every checkpoint_every steps, and returns
 * J = k + <phi^2> at the end and its gradient with respect to the initial u_hat (and phi_hat), by a
 * reverse sweep that recomputes each segment from its checkpoint; the context is left at the
 * initial state. */
hdns_status hdns_step_vjp(hdns_ctx *ctx, double dt, const double *lam_u_out, const double *lam_phi_out,
                          double *lam_u_in, double *lam_phi_in);
hdns_status hdns_adjoint_gradient(hdns_ctx *ctx, double dt, int32_t steps, int32_t checkpoint_every,
                                  double *objective, double *grad_u, double *grad_phi);
/* The same gradient under binomial checkpointing (Griewank's revolve): `snapshots` state slots, the
 * initial state's included, each step run forward at most r times, r the smallest with
 * C(snapshots + r, r) >= steps, and r steps - C(snapshots + r, r - 1) forward steps besides the
 * adjoint steps' own (the minimum).  The result equals hdns_adjoint_gradient's bit for bit.  hdns_revolve_schedule writes the
 * schedule's actions as (op, step, slot) triples (op: 0 advance to step, 1 store state `step` in
 * `slot`, 2 restore it, 3 first turn, 4 adjoint of step `step`) for inspection; with actions NULL it
 * only sets *count.  It is host arithmetic and needs no device. */
hdns_status hdns_adjoint_gradient_revolve(hdns_ctx *ctx, double dt, int32_t steps, int32_t snapshots,
                                          double *objective, double *grad_u, double *grad_phi);
hdns_status hdns_revolve_schedule(int32_t steps, int32_t snapshots, int32_t *actions, int64_t capacity,
                                  int64_t *count);

/* Ensembles of boxes of uneven sizes and viscosities (no counterpart in the reference).  The
 * configs may differ in grid_points and viscosity only.  Boxes of one size form a batch context;
 * all batches share one stream, and hdns_ensemble_step_graph advances every box with one graph
 * launch (captured once; each step's dt is written before the launch).  hdns_ensemble_member returns box `box`'s batch context with that
 * box selected, for the per-box entry points (the ensemble owns it: do not destroy it). */
typedef struct hdns_ensemble hdns_ensemble;
hdns_status hdns_ensemble_create(const hdns_config *configs, int32_t nbox, hdns_ensemble **out);
hdns_status hdns_ensemble_destroy(hdns_ensemble *ens);
const char *hdns_ensemble_last_error(const hdns_ensemble *ens);
hdns_status hdns_ensemble_member(hdns_ensemble *ens, int32_t box, hdns_ctx **ctx);
hdns_status hdns_ensemble_step(hdns_ensemble *ens, double dt);
hdns_status hdns_ensemble_step_graph(hdns_ensemble *ens, double dt);
/* The tracer half of an iteration for every box: traversal rebuild, Heun push, sample. */
hdns_status hdns_ensemble_advance_tracers(hdns_ensemble *ens, double dt);
hdns_status hdns_ensemble_synchronize(hdns_ensemble *ens);

/* Synchronise the context's stream (for timing from the host). */
hdns_status hdns_synchronize(hdns_ctx *ctx);
/* The HIP stream the context launches on (a hipStream_t), for event timing on that stream. */
hdns_status hdns_stream(hdns_ctx *ctx, void **stream);

#ifdef __cplusplus
}
#endif

#endif /* HDNS_H */
