// Host-side resolution of a structural spin-system description into the
// device operator formats.  Restates the operator semantics of SpinAPI
// (SpinSpace::InteractionOperator, SpinSpace_hamiltonians.cpp:438-777;
// Spin::Tx/Ty/Tz, Spin.cpp:360-408; SpinSpace::GetState,
// SpinSpace_states.cpp:626-700; reaction operators,
// SpinSpace_transitions.cpp:201-229) without building Hilbert-space matrices
// for the matrix-free path.
#pragma once
#include <array>
#include <complex>
#include <string>
#include <vector>

#include "../../include/molspin_gpu.h"
#include "msg_internal.hpp"

namespace msg {

using cd = std::complex<double>;
using Mat4 = std::array<cd, 16>;  // 4x4 electron-pair matrix, row-major, c = 2a + b

struct Term1 {       // H += c . S_s, c = Mz B (swept) or fixed c
    int s;
    bool swept;
    double c[3];
    double Mz[9];
    // time-dependent Zeeman field (MSG_TD_*): B(t) from the field as written
    int td = 0;
    bool td_perp = false;
    double td_freq = 1.0, td_phase = 0.0, td_axis[3] = {0.0, 0.0, 1.0}, field[3] = {0.0, 0.0, 0.0};
};

// Time point of a dynamics=dynamic evaluation (SpinSpace::StaticHamiltonian +
// DynamicHamiltonian at time t, SpinSpace_hamiltonians.cpp:1590-1740): time-dependent
// Zeeman terms take their field at `t`, or are left out (`skip`: the static part).
// Without a time point (dynamics=static) they keep the field as written.
struct TimePoint {
    double t = 0.0;
    bool skip = false;
};

// Field of a time-dependent Zeeman term at time t (Interaction.cpp:1900-1933).
void td_field(const Term1 &t, double time, double B[3]);

struct Term2 {       // H += sum_ab M_ab S_{s1,a} S_{s2,b}
    int s1, s2;
    double M[9];
};

struct Resolved {
    int nspins = 0;
    std::vector<int> mult;
    int64_t D = 0;
    int e1 = -1, e2 = -1;         // the radical-pair electrons (declaration order)
    std::vector<int> nuc;         // nuclei in device order (nucleus 0 least significant)
    std::vector<int64_t> perm;    // MolSpin index -> device index
    std::vector<Term1> t1;
    std::vector<Term2> t2;
    double shift = 0.0;
    bool pair_ok = false;         // two spin-1/2 electrons, electron-only states
    bool fast_ok = false;         // + matrix-free operator (nuclei spin-1/2 but at most one,
                                  //   nucleus-nucleus terms between spin-1/2 nuclei only)
    std::vector<std::array<int, 2>> nn_pairs;  // nucleus-nucleus couplings: device nucleus
                                               // indices (k1 < k2), one entry per pair
    bool has_ee = false;          // electron-electron couplings present
    bool has_td = false;          // time-dependent Zeeman terms present
    bool td_nuclear = false;      // ... acting on a nucleus
    std::string why_not_fast;
    // observables
    std::vector<Mat4> P_states;        // every State object (electron-pair projector)
    std::vector<Mat4> P_trans;         // transition source projectors
    std::vector<double> rates;         // transition rates
    double kmax = 0.0;
    Mat4 K{};                          // Haberkorn K on the electron pair
    bool symmetric = false;            // K = kappa * 1
    double kappa = 0.0;
    // relaxation
    int mix_mult = 0;  // multiplicity of the one nucleus with spin > 1/2 on the matrix-free path (0: none)
    bool relax = false;
    double relax_rate = 0.0;               // sum of the channels' rates
    std::vector<double> relax_rates;       // relaxationdephasing channels on the electron pair
    // direct factors: rho0 = rho_e (x) 1/Z; columns evec_i (x) e_nu / sqrt(Z)
    std::vector<std::array<cd, 4>> evec;
    int64_t zdim = 0;
    std::vector<int> initial;          // initial State indices as listed
    double first_rate = -1.0;          // rate of the first Transition object (< 0: none)
    // State objects as described (spins in listed order, amplitudes), for trace sampling
    std::vector<std::vector<int>> state_spins;
    std::vector<std::vector<cd>> state_amps;
};

// Returns an empty string on success, else the reason.
std::string resolve_system(const msg_system_desc *d, Resolved &r);

// Per-item coefficient block of the matrix-free operator (layout in msg_internal.hpp).
void fast_coefficients(const Resolved &r, const double *field, std::vector<float> &out, int &op_flags,
                       int nuc_flags[kMaxNuclei], const TimePoint *tp = nullptr);

// One-spin coefficient c of term t (H += c . S) at the item's field and time point;
// false when the term is left out (time-dependent term of a static part).
bool term1_value(const Term1 &t, const double *field, const TimePoint *tp, double v[3]);

// S_axis of spin `spin` (x 1 on the others), dense D x D in device order, row-major.
void spin_operator_dense(const Resolved &r, int spin, int axis, std::vector<cd> &H);

// Dense D x D Hamiltonian (rad/ns) in device order, row-major complex.
void dense_hamiltonian(const Resolved &r, const double *field, std::vector<cd> &H, const TimePoint *tp = nullptr);

// Split of a radical pair without electron-electron couplings into the two
// uncoupled radicals (TaskStaticHSDirectYieldsSymmUncoupled.cpp:100-150): A = e1
// plus every spin coupled to it, B = e2 plus every spin coupled to it; spins coupled
// to neither go to A.  Within each list the electron comes first.  Returns an empty
// string on success, else why the Hamiltonian does not factor as H_A (x) 1 + 1 (x) H_B.
std::string uncoupled_split(const Resolved &r, std::vector<int> &A, std::vector<int> &B);

// Dense Hamiltonian (rad/ns, row-major) of the spins in `spins` (first spin most
// significant): the one-spin terms of those spins and the two-spin terms among them.
void subsystem_hamiltonian(const Resolved &r, const std::vector<int> &spins, const double *field,
                           std::vector<cd> &H);

// exp(A) for a dense complex matrix (row-major), Pade(13) with scaling and
// squaring (Higham 2005).
void expm_dense(int n, const std::vector<cd> &A, std::vector<cd> &E);

// Trace samples of State `state` (SpinSpace::BuildTraceSamples): method 0
// coherent, 1 SU(Z); unit-norm columns in MolSpin order, D x m interleaved.
std::string trace_samples(const Resolved &r, int state, int method, uint64_t seed, int m, double *out);

// BuildFreeEvolutionTimes (TaskHSGeneral.cpp:59-96).
std::string free_times(double total, double dt, std::vector<double> &t);

// Legacy StaticHS grid: k dt for k < ceil(T/dt) (TaskStaticHSDirectYields.cpp:347).
std::string legacy_times(double total, double dt, std::vector<double> &t);

// Flip ranges of one relaxation half-step of length h for relaxationdephasing
// channels of rates `rates` on the electron pair (ApplyStochasticRelaxationHilbert,
// SpinSpace_relaxation.cpp:1211-1254); a column flips when the half-step's uniform
// word x (Philox) lies in [t[0], t[1]) or [t[2], t[3]).  One channel: the parity of
// its events, p = -expm1(-2 lambda h) / 2 with lambda = rate / 2, as x < floor(p 2^32).
// Several channels: the event count of the superposed Poisson process of mean
// mu = sum lambda_j h, drawn by inversion (count >= n exactly when x >= ceil(F(n-1)
// 2^32), F the Poisson CDF); every channel applies the same U = P_S - P_T with U^2 = 1,
// so the column flips when the count is odd (counts above 3 have probability below
// mu^4 / 24 and are not distinguished).
void relax_thresholds(const std::vector<double> &rates, double h, uint32_t t[4]);

}  // namespace msg
