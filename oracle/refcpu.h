/*
 * refcpu — FP64 CPU restatement of MolSpin's Hilbert-space propagation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle the HIP backend is checked
 * against and the CPU baseline bench.py times.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product (molspin_amd/) never
 * links or calls it.
 *
 * Every function restates a reference routine and cites it.  The reference itself
 * cannot be compiled here (it requires Armadillo, which is absent), so this code is
 * pinned by the reference's golden values, closed forms and invariants (tests/).
 *
 * Functions whose result no check pins bit-for-bit:
 *   - ref_coherent_samples / ref_suz_samples: the sample streams use this project's
 *     SplitMix64/Philox streams, not mt19937 (RNG-stream parity unpinned; the
 *     samples are pinned statistically and fed identically to both sides).
 */
#ifndef REFCPU_H
#define REFCPU_H
#include <stdint.h>
#include "../include/molspin_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Hilbert dimension (product of multiplicities). */
int64_t ref_dimension(const msg_system_desc *sys);

/* Dense Hamiltonian in rad/ns, column-major D x D complex (interleaved), at field
 * override `field` (NULL: as described). SpinSpace::Hamiltonian
 * (SpinSpace_hamiltonians.cpp:1456-1524) over InteractionOperator (:438-777). */
int ref_hamiltonian_dense(const msg_system_desc *sys, const double *field, double *out);

/* Dense projector onto State `state` (SpinSpace::GetState, SpinSpace_states.cpp:626-700). */
int ref_state_projector_dense(const msg_system_desc *sys, int state, double *out);

/* Haberkorn reaction operator K = sum_c k_c/2 P_c, dense
 * (SpinSpace::ReactionOperatorHilbertRotated, SpinSpace_transitions.cpp:201-229). */
int ref_reaction_dense(const msg_system_desc *sys, double *out);

/* Direct-sampling factors B with B B^dagger = rho0 (HSStatePreparation.cpp:77-165,
 * SpinSpace::FactorizeDensityMatrix SpinSpace_states.cpp:570-622).  Writes
 * *n_columns; `out` must hold D*D complex. */
int ref_direct_factors(const msg_system_desc *sys, double *out, int32_t *n_columns);

/* The HSGeneral free-evolution run (TaskHSGeneral.cpp:267-652) for a batch of
 * items, FP64.  method: MSG_METHOD_KRYLOV (Arnoldi, SpinSpace_operators.cpp:1257-1408),
 * MSG_METHOD_EXPONENTIAL (dense expmat every step, HSPropagator.cpp:92-100),
 * 3 = RK4 (HSPropagator.cpp:80-90).  Same argument meaning as msg_propagate.
 * `threads` > 1 runs items / column blocks with OpenMP. */
int ref_run_hsgeneral(const msg_system_desc *sys, const msg_run_params *params,
                      int32_t n_items, const double *fields,
                      const double *columns, int32_t n_columns,
                      double *yields, double *trace, int threads);

/* Same loop, but only columns [col_begin, col_end) of the factors and only the
 * first `max_steps` intervals; returns the per-column observables summed over
 * the selected columns (no trapezoid).  Used for bounded CPU timing samples and
 * per-column parity checks.  values: (max_steps+1) x n_obs. */
int ref_run_columns(const msg_system_desc *sys, const msg_run_params *params,
                    const double *field, const double *columns, int32_t n_columns,
                    int32_t col_begin, int32_t col_end, int64_t max_steps,
                    double *values, int threads);

/* Infinite-time symmetric-decay yields for every State object
 * (TaskStaticHSSymmetricDecay.cpp:114-158): Phi_P = sum_{g,h} Re<g|P|h><h|rho0|g> k^2/(k^2+w_gh^2). */
int ref_symmetric_decay(const msg_system_desc *sys, const double *field, double k, double *yields);

/* Time-infinity yields of HSGeneral (method=timeinf, TaskHSGeneral.cpp:453-467):
 * X = int_0^inf rho(t) dt from the D^2 x D^2 solve of SpinSpace::SolveHilbertTimeIntegral
 * (SpinSpace_operators.cpp:142-201) with rho0 = factors factors^dagger, then
 * EvaluateDensity (HSObservableCollector.cpp:290-301): scale_c Re Tr[P_c X].  No
 * relaxation; D <= 64.  yields: n_obs values. */
int ref_time_infinity(const msg_system_desc *sys, const msg_run_params *params, const double *field,
                      const double *columns, int32_t n_columns, double *yields);

/* The same solve for a given generator: X with G X + X G^dagger = -rho0 (column-major
 * n x n complex, interleaved). */
int ref_lyapunov_integral(int32_t n, const double *g, const double *rho0, double *out);

/* Legacy StaticHS-Direct-Yields, symmetric Krylov/Lanczos branch with basis reuse
 * (TaskStaticHSDirectYields.cpp:190-292,569-652,753-807).  yields: one per transition. */
int ref_legacy_direct_yields(const msg_system_desc *sys, const double *field,
                             double total_time, double dt, int krylov_size, double krylov_tol,
                             int correction, double *yields);

/* Legacy StaticHS-Stoch-Yields, symmetric branch with the dense propagator
 * (TaskStaticHSStochYields.cpp:257-286,684-768); columns are the sampled
 * trajectories (D x M).  yields: one per transition. */
int ref_legacy_stoch_yields(const msg_system_desc *sys, const double *field,
                            double total_time, double dt, const double *columns, int32_t m,
                            int correction, double *yields);

/* Spin-coherent trace samples |S-part> (x) nuclear coherent states, drawn from
 * (seed, index) as BuildTraceSamples' coherent route (SpinSpace_states.cpp:519-563):
 * theta = acos(1-2u), phi = 2 pi u per spin, projected on the support of State
 * `state` and normalised.  Not scaled by 1/sqrt(M).  out: D x m complex. */
int ref_coherent_samples(const msg_system_desc *sys, int state, uint64_t seed, int32_t m, double *out);

/* SU(Z) trace samples (SpinSpace_states.cpp:452-486): the State's normalised vector
 * (x) a normalised complex Gaussian vector over the open spins, Gaussians by
 * Box-Muller from Philox at counter (sample, component, 1, 0x53555A00).  out: D x m. */
int ref_suz_samples(const msg_system_desc *sys, int state, uint64_t seed, int32_t m, double *out);

/* Philox4x32-10 (Salmon et al., SC'11). */
void ref_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Fibonacci-sphere point `index` of `n` scaled by `magnitude`
 * (RunSection/Utility.cpp:44-79, ActionFibonacciSphere.cpp:50-64). */
void ref_fibonacci_point(int32_t n, int32_t index, double magnitude, double out[3]);

/* arma::logspace(a, b, n) (ActionLogSpace.cpp:11-18). */
void ref_logspace(double a, double b, int32_t n, double *out);

/* Trapezoid of values over times (HSObservableCollector.cpp:304-326). */
double ref_trapezoid(const double *times, const double *values, int64_t n);

/* Message of the last failed call on this thread. */
const char *ref_last_error(void);

/* BuildFreeEvolutionTimes (TaskHSGeneral.cpp:59-96); times may be NULL. */
int ref_time_grid(double total_time, double time_step, int64_t *n_samples, double *times);

/* Hermitian eigen-decomposition (complex Jacobi), column-major in/out. */
int ref_eig_hermitian(int32_t n, const double *a, double *evals, double *evecs);

/* exp(A) of a dense complex matrix (Pade 13 + scaling and squaring, Higham 2005). */
int ref_expm(int32_t n, const double *a, double *out);

#ifdef __cplusplus
}
#endif
#endif
