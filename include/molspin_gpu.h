/*
 * molspin_gpu.h — C ABI of the MI355X (gfx950) Hilbert-space propagation backend.
 *
 * The boundary replaces the per-orientation time loop of MolSpin's HSGeneral task
 * (RunSection/General/HS/TaskHSGeneral.cpp:470-576) together with the propagator it
 * drives (RunSection/General/HS/HSPropagator.cpp:51-140) and the observable
 * collector it samples (RunSection/General/HS/HSObservableCollector.cpp:276-326).
 * A caller describes a spin system structurally (spins, interactions, states,
 * transitions, relaxation operators — the SpinAPI object model), uploads it once,
 * and then propagates batches of "items" (field orientations or field strengths)
 * on the device.  The library owns the whole time loop: observable sampling at
 * every grid point, the trapezoid and the optional finite-time correction.
 *
 * Conventions
 *   - Plain C: pointers and sizes only.  No entry point throws; every entry point
 *     returns an msg_status code and leaves a message retrievable with
 *     msg_last_error() (mirrors the bool + std::string& error convention of
 *     HSPropagator::Step, HSPropagator.h:40-41).
 *   - Complex arrays are interleaved (re, im) doubles.  State vectors exchanged
 *     with the caller use MolSpin's basis order: spins in declaration order, the
 *     first spin most significant (SpinSpace::CreateOperator kron order,
 *     SpinSpace_operators.cpp:84-102), and within a spin |+S> first
 *     (SpinSpace_states.cpp:206-209).  Column-major D x M, like arma::cx_mat.
 *   - Units are MolSpin's: fields and tensors as written in the input (T, or mT with
 *     prefactor=0.001), rates in 1/ns, times in ns.  The common prefactor
 *     87.9410005 rad ns^-1 T^-1 is applied where an interaction requests it
 *     (SpinSpace_hamiltonians.cpp:746-748).
 *   - One context per host thread at a time; a context owns one HIP stream.
 */
#ifndef MOLSPIN_GPU_H
#define MOLSPIN_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum msg_status {
    MSG_OK = 0,
    MSG_INVALID_ARG = 1,      /* malformed description or parameters            */
    MSG_UNSUPPORTED = 2,      /* valid for MolSpin but outside the device paths  */
    MSG_OOM_ADMISSION = 3,    /* the request does not fit the device budget      */
    MSG_DEVICE_ERROR = 4,     /* HIP runtime failure                             */
    MSG_NUMERIC_BREAKDOWN = 5 /* non-finite state or Krylov breakdown            */
} msg_status;

/* ------------------------------------------------------------------------- */
/* Structural system description (SpinAPI object model)                      */
/* ------------------------------------------------------------------------- */

/* SpinAPI::Spin (SpinAPI/Spin.cpp:29-53).  `tensor` is Tensor::LabFrame() of the
 * spin tensor (default isotropic 2.0, Spin.cpp:29), row-major.  Quantisation axes
 * are the laboratory axes. */
typedef struct msg_spin_desc {
    int32_t multiplicity;      /* 2S+1 */
    int32_t is_electron;       /* type=electron */
    int32_t tensor_isotropic;  /* IsIsotropic(tensor) (Tensor.cpp:834-840) */
    int32_t reserved;
    double  tensor[9];
} msg_spin_desc;

enum {
    MSG_INTERACTION_ZEEMAN = 1,     /* onespin/singlespin/zeeman      (Interaction.cpp:91)  */
    MSG_INTERACTION_DOUBLESPIN = 2, /* twospin/doublespin/hyperfine/dipole (Interaction.cpp:106) */
    MSG_INTERACTION_EXCHANGE = 3    /* exchange                        (Interaction.cpp:116) */
};

/* SpinAPI::Interaction, the part InteractionOperator reads
 * (SpinSpace_hamiltonians.cpp:438-777).  Spin groups index msg_system_desc.spins
 * through msg_system_desc.index_pool. */
typedef struct msg_interaction_desc {
    int32_t type;
    int32_t group1_offset, group1_count;
    int32_t group2_offset, group2_count;
    int32_t tensor_present;    /* coupling tensor given (else dot product)            */
    int32_t tensor_isotropic;  /* IsIsotropic(coupling tensor)                         */
    int32_t common_prefactor;  /* commonprefactor (default true, Interaction.cpp:25)  */
    int32_t ignore_tensors;    /* ignoretensors   (default false)                      */
    int32_t field_swept;       /* Zeeman field replaced by the batch item's field      */
    double  prefactor;         /* prefactor (default 1)                                */
    double  field[3];          /* Zeeman field as written                              */
    double  tensor[9];         /* Tensor::LabFrame() of the coupling tensor, row-major */
    /* Time dependence of a Zeeman field (fieldtype / timedependence, Interaction.cpp:
     * 500-560, 1240-1262, 1900-1933): MSG_TD_*.  Used by dynamics=dynamic runs only;
     * a static run takes `field` as written. */
    int32_t time_dependence;
    int32_t td_perpendicular;  /* perpendicularoscillations (circular; default false)  */
    double  td_frequency;      /* frequency, rad/ns (default 1)                        */
    double  td_phase;          /* phase, rad (default 0)                               */
    double  td_axis[3];        /* axis (circular; default 0 0 1, used as given)        */
} msg_interaction_desc;

enum {
    MSG_TD_NONE = 0,
    MSG_TD_LINEAR = 1,         /* linearpolarized: B(t) = field cos(w t + phase)            */
    MSG_TD_CIRCULAR = 2        /* circularpolarized: B(t) = R(w t + phase) b with the Rodrigues
                                  matrix R = 1 + sin(x) W + 2 sin^2(x/2) W^2 of the axis, b =
                                  field or, with perpendicularoscillations, field - (axis.field) axis */
};

/* SpinAPI::State restricted to the spins it fixes: a (not necessarily normalised)
 * amplitude vector over `spins` in the listed order, first spin most significant.
 * Products of independently specified spins are given as their kron product. */
typedef struct msg_state_desc {
    int32_t spins_offset, spins_count; /* into index_pool */
    int32_t amp_offset;                /* into amp_pool, in complex elements */
    int32_t reserved;
} msg_state_desc;

/* SpinAPI::Transition: a sink from `source_state` at `rate` (1/ns). */
typedef struct msg_transition_desc {
    int32_t source_state;
    int32_t reserved;
    double  rate;
} msg_transition_desc;

/* relaxationdephasing Operator on two spin-1/2 spins (SpinAPI/Operator.cpp:133,250;
 * SpinSpace_relaxation.cpp:1144-1199). */
typedef struct msg_relaxation_desc {
    int32_t spin1, spin2;
    double  rate;
} msg_relaxation_desc;

typedef struct msg_system_desc {
    int32_t n_spins;
    const msg_spin_desc *spins;
    int32_t n_interactions;
    const msg_interaction_desc *interactions;
    int32_t n_states;
    const msg_state_desc *states;
    int32_t n_transitions;
    const msg_transition_desc *transitions;
    int32_t n_initial;                 /* initialstate list (Properties initialstate=)  */
    const int32_t *initial_states;
    const double *initial_weights;     /* NULL: equal weights (HSStatePreparation.cpp:96-103) */
    int32_t n_relaxations;
    const msg_relaxation_desc *relaxations;
    int32_t index_pool_len;
    const int32_t *index_pool;
    int32_t amp_pool_len;              /* complex elements */
    const double *amp_pool;            /* interleaved re, im */
} msg_system_desc;

/* ------------------------------------------------------------------------- */
/* Run parameters (HSExecutionPlan, HSExecutionPlan.h:33-117)               */
/* ------------------------------------------------------------------------- */

enum {
    MSG_CALC_YIELDS = 0,          /* calculation=yields        */
    MSG_CALC_TIMEEVOLUTION = 1    /* calculation=timeevolution */
};
enum {
    MSG_OBS_TRANSITION_YIELDS = 0, /* rate * Tr[P_source rho]  (HSObservableCollector.cpp:276-288) */
    MSG_OBS_STATE_POPULATIONS = 1  /* Tr[P_state rho] for every State object                       */
};
enum {
    MSG_METHOD_AUTO = 0,          /* device chooses: dense U for D<=512, matrix-free Krylov above */
    MSG_METHOD_KRYLOV = 1,        /* propagationmethod=krylov, fixed krylovsize: Lanczos for K = kappa 1,
                                     Arnoldi on -iH - K for unequal rates (D <= 16384)            */
    MSG_METHOD_EXPONENTIAL = 2,   /* propagationmethod=normal: exact exp(dt G) applied as a GEMM  */
    MSG_METHOD_RK4 = 3,           /* propagationmethod=rk4 (HSPropagator.cpp:80-90): for a static G
                                     one RK4 step is the degree-4 Taylor polynomial of dt G, applied
                                     as a GEMM (D <= 512)                                          */
    MSG_METHOD_UNCOUPLED = 4      /* exact propagation of uncoupled radicals (no electron-electron
                                     coupling, symmetric recombination, no relaxation, direct
                                     sampling): H = H_A (x) 1 + 1 (x) H_B, the observables follow from
                                     per-radical electron-spin correlation functions computed from
                                     the radicals' eigenpairs (TaskStaticHSDirectYieldsSymmUncoupled
                                     .cpp:150-470 restated exactly); radicals of dimension <= 512 */
};
enum {
    MSG_SAMPLING_DIRECT = 0,      /* factors of rho0 (HSStatePreparation.cpp:77-165)              */
    MSG_SAMPLING_GIVEN = 1        /* caller-supplied columns (trace samples drawn on the host)    */
};

enum {
    MSG_TASK_HSGENERAL = 0,            /* type=HSGeneral (TaskHSGeneral.cpp:267-652)               */
    MSG_TASK_LEGACY_DIRECT_YIELDS = 1, /* type=statichs-direct-yields (TaskStaticHSDirectYields.cpp) */
    MSG_TASK_LEGACY_STOCH_YIELDS = 2,  /* type=statichs-stoch-yields (TaskStaticHSStochYields.cpp)   */
    MSG_TASK_SYMMETRIC_DECAY = 3       /* type=statichs-symmetricdecay: msg_symmetric_decay, not
                                          msg_propagate (TaskStaticHSSymmetricDecay.cpp:33-171)     */
};

typedef struct msg_run_params {
    int32_t  calculation;
    int32_t  observables;
    int32_t  method;
    int32_t  krylov_size;         /* krylovsize (default 16) */
    int32_t  sampling;
    int32_t  yield_corrections;   /* yieldcorrections (default false) */
    double   total_time;          /* totaltime (ns) */
    double   time_step;           /* timestep  (ns) */
    uint64_t seed;                /* key of the counter-based relaxation RNG */
    int32_t  task;                /* MSG_TASK_*: the legacy tasks sample k dt for k < ceil(T/dt)
                                     (endpoint excluded, TaskStaticHSDirectYields.cpp:347) and treat
                                     equal rates as symmetric recombination with the decay
                                     e^{-k t} applied analytically (:257-292) */
    int32_t  dynamics;            /* dynamics=dynamic (HSExecutionPlan.cpp:96-101): time-dependent
                                     Zeeman fields at their time-dependent value; each interval
                                     uses H at its midpoint (TaskHSGeneral.cpp:527-537), or the RK4
                                     stages at t, t+dt/2, t+dt with method RK4 (:493-519) */
    double   krylov_tol;          /* krylovtol of statichs-direct-yields (TaskStaticHSDirectYields.cpp:
                                     373-410, 569-652): a column keeps its Krylov basis while
                                     h_{m+1,m} |c_{m-1}| <= krylov_tol; 0 builds a fresh basis every
                                     step (the reference's default 1e-16 in effect) */
} msg_run_params;

/* propagationmethod as written in the plan (HSExecutionPlan.h:29) */
enum {
    MSG_PROP_EXPONENTIAL = 0,     /* normal / exp / exponential / matrixexponential */
    MSG_PROP_AUTOEXPM = 1,        /* autoexpm                                       */
    MSG_PROP_KRYLOV = 2,          /* krylov                                         */
    MSG_PROP_RK4 = 3              /* rk4 / explicit                                 */
};

/* A resolved Task property list (ResolveExecutionPlan, HSExecutionPlan.cpp:89-615). */
typedef struct msg_plan {
    msg_run_params params;        /* run parameters for msg_propagate / msg_symmetric_decay */
    int32_t montecarlo_samples;   /* montecarlosamples (default 1): columns for sampling=stochastic */
    int32_t sampling_method;      /* samplingmethod: MSG_TRACE_SUZ (default) or MSG_TRACE_COHERENT */
    int32_t auto_seed;            /* autoseed (default true) */
    int32_t device_backend;       /* backend=mi355x (default cpu) */
    int32_t time_infinity;        /* method=timeinf (calculation=yields): msg_time_infinity */
    int32_t propagation;          /* MSG_PROP_* */
} msg_plan;

/*
 * Resolves the property list of an MSD Task block, "key=value; key=value; ...",
 * with ResolveExecutionPlan's keys, aliases, defaults and error messages.  Keys
 * selecting physics off this path (powder grids, secular
 * approximation, spin-polarisation/CIDSP output, pulse sequences) return
 * MSG_UNSUPPORTED naming the key; invalid values return MSG_INVALID_ARG with the
 * reference's message.  `error` (error_len bytes, may be NULL) receives the message.
 * Host-only: needs no device.
 */
msg_status msg_resolve_plan(const char *task_properties, msg_plan *plan, char *error, int32_t error_len);

/* ------------------------------------------------------------------------- */
/* MSD front-end (host only)                                                 */
/* ------------------------------------------------------------------------- */

/*
 * Compiles the text of an MSD input file: one SpinSystem (Spin, Interaction, State,
 * Transition, Operator relaxationdephasing, Properties initialstate), the Settings
 * (steps; RotateVector, FibonacciSphere and LogSpace actions on one Zeeman field or
 * prefactor) and the Run tasks.  It replaces MSDParser::Parse + RunSection set-up
 * (MSDParser/MSDParser.cpp:50-270, main.cpp:600-680) for this path: the RunSection
 * step loop (tasks at step i, then Action::Step(i + 1), grid actions installing point
 * zero before step 1) is replayed on the host, so the swept field of every step is
 * known and the whole sweep runs as one msg_propagate batch, one item per step.
 * Objects off this path return MSG_UNSUPPORTED naming them; malformed input
 * MSG_INVALID_ARG.  Host-only: needs no device.
 */
typedef struct msg_msd msg_msd;
msg_status msg_msd_compile(const char *text, msg_msd **out, char *error, int32_t error_len);
/* The compiled system, for msg_system_upload; valid until msg_msd_destroy.  A swept
 * Zeeman interaction has field_swept set and its prefactor folded into the fields. */
msg_status msg_msd_system(const msg_msd *msd, const msg_system_desc **desc);
/* Calculation steps; *swept = 1 when an action drives a Zeeman field, and then
 * `fields` (n_steps x 3, may be NULL) receives each step's field (prefactor x field). */
msg_status msg_msd_steps(const msg_msd *msd, int32_t *n_steps, int32_t *swept, double *fields);
msg_status msg_msd_task_count(const msg_msd *msd, int32_t *n);
/* Task `index` in file order: its name, its plan (msg_resolve_plan of its properties)
 * and the resolver's status and message, which this call returns. */
msg_status msg_msd_task(const msg_msd *msd, int32_t index, msg_plan *plan, char *name, int32_t name_len,
                        char *why, int32_t why_len);
msg_status msg_msd_destroy(msg_msd *msd);

/* ------------------------------------------------------------------------- */
/* Context and system handles                                               */
/* ------------------------------------------------------------------------- */

typedef struct msg_ctx msg_ctx;
typedef struct msg_system msg_system;

/* Creates a context on HIP device `device` with its own stream. */
msg_status msg_ctx_create(int device, msg_ctx **out);
msg_status msg_ctx_destroy(msg_ctx *ctx);
/* Message of the last failed call on `ctx` (empty string if none).  Valid until
 * the next call on the context.  With ctx == NULL: the last msg_ctx_create error. */
const char *msg_last_error(const msg_ctx *ctx);

/* Validates and resolves a system description (SpinSpace construction,
 * SpinSpace.cpp:56-65; InteractionOperator resolution) and keeps its
 * field-independent operator factors resident on the device. */
msg_status msg_system_upload(msg_ctx *ctx, const msg_system_desc *desc, msg_system **out);
msg_status msg_system_release(msg_ctx *ctx, msg_system *sys);
/* Hilbert dimension of an uploaded system. */
msg_status msg_system_dimension(const msg_system *sys, int64_t *dim);

/* Device memory (bytes) one propagate call over `n_items` items with `n_columns`
 * columns each would hold at peak; used for admission by HBM footprint. */
msg_status msg_query_footprint(const msg_system *sys, const msg_run_params *params,
                               int32_t n_items, int32_t n_columns, int64_t *bytes);

enum {
    MSG_TRACE_COHERENT = 0,   /* samplingmethod=coherent: spin-coherent product states */
    MSG_TRACE_SUZ = 1         /* samplingmethod=suz: fixed State (x) random SU(Z) vector */
};

/* Trace samples of State `state` (SpinSpace::BuildTraceSamples,
 * SpinSpace_states.cpp:376-568) drawn on the host from a counter-based stream
 * (Philox4x32-10 keyed by `seed`): coherent - every spin in
 * exp(-i phi Sz) exp(-i theta Sy)|+S>, theta = acos(1-2u), phi = 2 pi u, then
 * projected on the State's support and normalised; suz - the State's vector (x)
 * a normalised complex Gaussian vector over the open spins.  Columns are unit
 * norm and in MolSpin order (D x m complex); scale by 1/sqrt(m) before
 * propagating (HSStatePreparation.cpp:226).  Host-only: needs no device. */
msg_status msg_trace_samples(const msg_system *sys, int32_t state, int32_t method, uint64_t seed,
                             int32_t m, double *columns);

/* Number of columns DIRECT sampling produces for `sys` (the rank of rho0). */
msg_status msg_direct_columns(const msg_system *sys, int32_t *n_columns);

/* ------------------------------------------------------------------------- */
/* Propagation (replaces TaskHSGeneral.cpp:470-576 for a batch of items)    */
/* ------------------------------------------------------------------------- */

/* Number of observables written per sample for `params` (transition count or
 * State count). */
msg_status msg_observable_count(const msg_system *sys, const msg_run_params *params,
                                int32_t *n_obs);
/* Number of time samples of the free-evolution grid (BuildFreeEvolutionTimes,
 * TaskHSGeneral.cpp:59-96): 0, dt, 2dt, ..., T with an exact endpoint.  (The
 * legacy tasks' grid has ceil(T/dt) samples k dt.) */
msg_status msg_time_grid(double total_time, double time_step, int64_t *n_samples,
                         double *times /* may be NULL; else n_samples entries */);

/*
 * Propagates `n_items` items of one system.
 *   fields      n_items x 3 field vectors for the field_swept Zeeman interactions
 *               (NULL: every item uses the fields as described).
 *   columns     SAMPLING_GIVEN only: D x n_columns complex (MolSpin order), shared
 *               by all items; scaled as given (HSStatePreparation.cpp:226 scales
 *               trace samples by 1/sqrt(M) before they arrive here).
 *   n_columns   SAMPLING_GIVEN only.
 *   yields      n_items x n_obs (calculation=yields): the integrated observables
 *               of each item, after the optional finite-time correction.
 *   trace       n_items x n_samples x n_obs (calculation=timeevolution), or NULL.
 * Blocks until the results are on the host.
 */
msg_status msg_propagate(msg_ctx *ctx, const msg_system *sys, const msg_run_params *params,
                         int32_t n_items, const double *fields,
                         const double *columns, int32_t n_columns,
                         double *yields, double *trace);

/*
 * msg_propagate without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Every device buffer of the run
 * (operator factors, grids, Krylov bases, partial sums, results) comes from a
 * library-owned scratch of 64 MiB per GPU, held until the process ends; runs that
 * do not fit return MSG_OOM_ADMISSION, and the number of resident columns shrinks to
 * what fits beside the rest.  Matrix-free path only (symmetric recombination with
 * 256 <= D <= 262144, or unequal rates up to D = 16384); the dense path returns
 * MSG_UNSUPPORTED.  Calls on one GPU share the scratch: calls that overlap on one
 * GPU (other host threads, other contexts) race.  Arguments and results as
 * msg_propagate; the same results.
 */
msg_status msg_propagate_noalloc(msg_ctx *ctx, const msg_system *sys, const msg_run_params *params,
                                 int32_t n_items, const double *fields,
                                 const double *columns, int32_t n_columns,
                                 double *yields, double *trace);

/*
 * One radical pair (system) of an uneven batch: what msg_propagate takes for one
 * system, plus where its results go.
 */
typedef struct msg_uneven_entry {
    const msg_system *system;
    const msg_run_params *params;  /* static Hamiltonian, symmetric recombination on the
                                      matrix-free path (any recombination at D <= 512) */
    int32_t n_items;
    const double *fields;          /* n_items x 3, or NULL */
    const double *columns;         /* SAMPLING_GIVEN: D x n_columns complex, else NULL */
    int32_t n_columns;
    double *yields;                /* n_items x n_obs */
    double *trace;                 /* n_items x n_samples x n_obs, or NULL */
} msg_uneven_entry;

/*
 * Propagates several systems of different Hilbert dimension in one persistent
 * launch: every (entry, item, column) of every entry goes into one work queue,
 * largest dimension first, and each workgroup of the launch takes whatever comes
 * next (matrix-free Lanczos bodies for 256 <= D <= 16384, dense MFMA bodies for
 * D <= 512 under method AUTO / EXPONENTIAL / RK4).  Replaces the loop over
 * systems and sweep points that rebuilds and reruns the task per point
 * (RunSection/RunSection.cpp:96-140, main.cpp:634-686) around the per-system
 * time loop (TaskHSGeneral.cpp:470-576).  An entry's results do not depend on the
 * other entries of the batch.  Blocks until every entry's results are on the host.
 */
msg_status msg_propagate_uneven(msg_ctx *ctx, int32_t n_entries, const msg_uneven_entry *entries);

/*
 * Infinite-time yields of every State object under symmetric decay at rate
 * `rate` (TaskStaticHSSymmetricDecay::RunLocal, TaskStaticHSSymmetricDecay.cpp:33-171):
 *   Phi_P = sum_{g,h} Re(<g|P|h><h|rho0|g>) k^2 / (k^2 + (lambda_g - lambda_h)^2)
 * over the eigenpairs of H (FP64 on the device), rho0 = sum of the initial-state
 * projectors / trace.  rate = 0 takes the reference's rule: the first Transition's
 * rate, else 1e-3 /ns.  fields: n_items x 3 (NULL: as described).
 * yields: n_items x n_states.  States on the radical-pair electrons, D <= 4096.
 */
msg_status msg_symmetric_decay(msg_ctx *ctx, const msg_system *sys, int32_t n_items, const double *fields,
                               double rate, double *yields);

/*
 * Time-infinity yields of HSGeneral (calculation=yields; method=timeinf): replaces
 * HSPropagator::SolveTimeInfinity + HSObservableCollector::EvaluateDensity
 * (TaskHSGeneral.cpp:453-467, HSPropagator.cpp:259-271, SpinSpace_operators.cpp:142-201,
 * HSObservableCollector.cpp:290-301).  Per item,
 *   X = int_0^inf rho(t) dt,   G X + X G^dagger = -rho0,   G = -iH - K,
 *   y_c = scale_c Re Tr[P_c X]  (scale = rate for transition yields, 1 for States),
 * with rho0 = rho_e (x) 1/Z (SAMPLING_DIRECT) or columns columns^dagger (SAMPLING_GIVEN,
 * n_columns <= D, MolSpin basis order, column-major).  FP64 on the device:
 *   - K = kappa 1 and D <= 4096 (method AUTO / KRYLOV): the eigenpairs of H give
 *     X_gh = rho0_gh / (2 kappa + i (lambda_g - lambda_h)) in the eigenbasis;
 *   - otherwise D <= 2048 (and always with method EXPONENTIAL): X by doubling,
 *     X <- X + U X U^dagger, U <- U^2 from U = e^{G tau} and the Taylor-series
 *     integral over [0, tau], until the added integral over [T, 2T] is below 1e-10
 *     of X (Frobenius).
 * Systems with relaxation operators are MSG_UNSUPPORTED (the reference adds a
 * Liouville-space relaxation term there).  A part of the space that never decays
 * adds nothing while rho(t) never reaches it, as the minimum-norm solution the
 * reference's singular solve falls back to; once populated it makes the integral
 * diverge: MSG_NUMERIC_BREAKDOWN (the doubling route reports it when the integral has
 * not converged by T = tau 2^36, with tau = 1/(4 ||G||_1)).  total_time, time_step, krylov_size and yield_corrections are not used
 * (HSObservableCollector.cpp:366 applies no correction at time infinity).
 * fields: n_items x 3 (NULL: as described).  yields: n_items x n_obs.
 */
msg_status msg_time_infinity(msg_ctx *ctx, const msg_system *sys, const msg_run_params *params,
                             int32_t n_items, const double *fields,
                             const double *columns, int32_t n_columns, double *yields);

/* Launch shape chosen by msg_autotune. */
typedef struct msg_launch_shape {
    int32_t path;                 /* 1 matrix-free Lanczos, 2 dense MFMA */
    int32_t groups_per_thread;    /* Lanczos: electron-pair groups per thread (1, 2, 4), -1 streaming with the
                                     Krylov vector in LDS, -2 streaming with LDS tiles, 0 plain streaming,
                                     -3 each column split over several workgroups */
    int32_t threads;              /* workgroup size */
    int32_t slots;                /* resident workgroups (persistent launch); columns in flight for -3 */
    double kernel_ms;             /* device time of the chosen shape on the sample */
} msg_launch_shape;

/*
 * Autotuner (no counterpart in the reference: the reference's CPU loop has no
 * launch shapes).  Times the candidate launch shapes of this system's path --
 * workgroup shape and Krylov-vector placement (LDS, LDS tiles, streamed, or a
 * column split over several workgroups), and the number of resident workgroups --
 * on a sample of `params` (n_items items
 * of `fields`/`columns` as for msg_propagate, `sample_steps` time steps) and
 * keeps the fastest for this context and Hilbert dimension (a candidate tried later
 * must be more than 1.5 % faster than the best earlier one to replace it): later
 * msg_propagate, msg_step and server launches on `ctx` use it.  Results do not
 * depend on the shape beyond FP32 rounding.  `best` may be NULL.
 */
msg_status msg_autotune(msg_ctx *ctx, const msg_system *sys, const msg_run_params *params, int32_t n_items,
                        const double *fields, const double *columns, int32_t n_columns, int32_t sample_steps,
                        msg_launch_shape *best);

/*
 * Per-step compatibility entry with the meaning of HSPropagator::StepStochastic
 * (HSPropagator.cpp:131-140): advances `factors` (D x n_columns complex, MolSpin
 * order, in place) by one interval `dt` under G = -iH - K of the system at
 * `field` (NULL: as described), with the stochastic relaxation half-steps drawn
 * from the counter-based RNG at (seed, item 0, step_index).
 */
msg_status msg_step(msg_ctx *ctx, const msg_system *sys, const msg_run_params *params,
                    const double *field, double dt, int64_t step_index,
                    double *factors, int32_t n_columns);

/* Device time (ms, HIP events on the context stream) and launch count of the
 * propagation kernels of the last msg_propagate call on `ctx`.  Diagnostic
 * counterpart of MolSpin's per-step wall clock (main.cpp:521-522,651-685). */
msg_status msg_last_kernel_time(const msg_ctx *ctx, double *ms, int32_t *launches);

/* ------------------------------------------------------------------------- */
/* Serving: continuous batching of yield requests                           */
/* ------------------------------------------------------------------------- */

typedef struct msg_server msg_server;

/* A server keeps up to `max_slots` requests resident and admits new ones as
 * slots and the `hbm_budget_bytes` budget free up. */
msg_status msg_server_create(msg_ctx *ctx, int32_t max_slots, int64_t hbm_budget_bytes,
                             msg_server **out);
msg_status msg_server_destroy(msg_server *srv);
/* Queues a yield request: a system, its run parameters and `n_fields` field
 * vectors.  Returns a request id.  The system must stay uploaded until done. */
msg_status msg_server_submit(msg_server *srv, const msg_system *sys, const msg_run_params *params,
                             int32_t n_fields, const double *fields, int64_t *request_id);
/* Collects finished requests (without blocking), then admits queued requests
 * while a slot and the HBM budget allow, shortest first (state-steps x D; arrival
 * order among equals).  A request with method MSG_METHOD_UNCOUPLED is computed
 * within the call (at most one per call).  max_steps <= 0: queued requests that
 * the uneven launch takes (msg_propagate_uneven: static Hamiltonian, symmetric
 * recombination above D = 512) and that are small enough to fit about 50 ms of
 * device time together share one persistent launch on one stream, smallest first;
 * every other admitted request
 * is propagated over its whole grid on its own stream, in persistent launches of
 * a few fields each (about 50 ms of device time: a large request does not hold the
 * device while small ones wait), so requests overlap on the device.
 * max_steps > 0: resident requests advance
 * in chunks of max_steps intervals per call, their states kept on the device.
 * Sets *n_active to the number of requests still queued or running. */
msg_status msg_server_pump(msg_server *srv, int32_t max_steps, int32_t *n_active);
/* Copies the yields (n_fields x n_obs) of a finished request and forgets it.
 * Returns MSG_INVALID_ARG if the request is unknown, *done = 0 if unfinished. */
msg_status msg_server_collect(msg_server *srv, int64_t request_id, double *yields, int32_t *done);
/* Counts since the server was created: shared launches and the requests they held.
 * (No counterpart in the reference: it reports how the server batched the stream.) */
msg_status msg_server_stats(const msg_server *srv, int64_t *shared_launches, int64_t *shared_requests);

#ifdef __cplusplus
}
#endif
#endif /* MOLSPIN_GPU_H */
