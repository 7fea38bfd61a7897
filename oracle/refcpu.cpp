// refcpu — FP64 CPU restatement of MolSpin's Hilbert-space propagation path.
//
// TEST INFRASTRUCTURE ONLY (the oracle): loaded by tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg, never by the product under molspin_amd/.
//
// Each routine restates the reference routine it cites; the numerics are plain
// FP64 (std::complex<double>), sparse CSR for operators, dense column-major
// matrices for small exponentials and eigen-decompositions.
//
// Parity pins (tests/test_oracle.py): reference golden values of
// Tests/tests_TaskStaticHSSymmetricDecay.cpp:64-69,150-174; closed forms of
// Tests/tests_HSGeneral.cpp (trapezoid/grid/correction); Breit-Rabi and
// eigen-decomposition closed forms for the time loop; Philox known-answer
// vectors; Fibonacci/LogSpace vectors of Tests/tests_actions.cpp:266-335.
// Time-infinity yields (ref_time_infinity) are pinned by the closed forms of
// Tests/tests_HSGeneral.cpp:1410-1431,1704-1730, by scipy's Bartels-Stewart Lyapunov
// solver and by the eigen closed form for K = kappa 1.
// Parity unpinned: the SplitMix64/Philox sample streams of ref_coherent_samples
// (checked statistically only).
#include "refcpu.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;
const cd I1(0.0, 1.0);
const double kCommonPrefactor = 8.79410005e+1;  // SpinSpace_hamiltonians.cpp:746-748

// ---------------------------------------------------------------------------
// Dense column-major complex matrix
// ---------------------------------------------------------------------------
struct Mat {
    int r = 0, c = 0;
    std::vector<cd> a;
    Mat() = default;
    Mat(int rr, int cc) : r(rr), c(cc), a((size_t)rr * cc, cd(0.0)) {}
    cd &operator()(int i, int j) { return a[(size_t)j * r + i]; }
    const cd &operator()(int i, int j) const { return a[(size_t)j * r + i]; }
    static Mat eye(int n) { Mat m(n, n); for (int i = 0; i < n; ++i) m(i, i) = 1.0; return m; }
};

Mat matmul(const Mat &A, const Mat &B) {
    Mat C(A.r, B.c);
    for (int j = 0; j < B.c; ++j)
        for (int k = 0; k < A.c; ++k) {
            const cd b = B(k, j);
            if (b == 0.0) continue;
            const cd *ac = &A.a[(size_t)k * A.r];
            cd *cc = &C.a[(size_t)j * C.r];
            for (int i = 0; i < A.r; ++i) cc[i] += ac[i] * b;
        }
    return C;
}
Mat add(const Mat &A, const Mat &B, cd sb = 1.0) {
    Mat C = A;
    for (size_t i = 0; i < C.a.size(); ++i) C.a[i] += sb * B.a[i];
    return C;
}
Mat scale(const Mat &A, cd s) { Mat C = A; for (auto &x : C.a) x *= s; return C; }
double norm1(const Mat &A) {
    double m = 0.0;
    for (int j = 0; j < A.c; ++j) {
        double s = 0.0;
        for (int i = 0; i < A.r; ++i) s += std::abs(A(i, j));
        m = std::max(m, s);
    }
    return m;
}
// Solve A X = B by LU with partial pivoting.
Mat solve(Mat A, Mat B) {
    const int n = A.r;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::abs(A(k, k));
        for (int i = k + 1; i < n; ++i)
            if (std::abs(A(i, k)) > best) { best = std::abs(A(i, k)); p = i; }
        if (best == 0.0) throw std::runtime_error("singular matrix in Pade solve");
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A(k, j), A(p, j));
            for (int j = 0; j < B.c; ++j) std::swap(B(k, j), B(p, j));
        }
        for (int i = k + 1; i < n; ++i) {
            const cd f = A(i, k) / A(k, k);
            if (f == 0.0) continue;
            A(i, k) = f;
            for (int j = k + 1; j < n; ++j) A(i, j) -= f * A(k, j);
            for (int j = 0; j < B.c; ++j) B(i, j) -= f * B(k, j);
        }
    }
    for (int j = 0; j < B.c; ++j)
        for (int i = n - 1; i >= 0; --i) {
            cd s = B(i, j);
            for (int k = i + 1; k < n; ++k) s -= A(i, k) * B(k, j);
            B(i, j) = s / A(i, i);
        }
    return B;
}

// exp(A): Pade(13) with scaling and squaring (Higham, SIAM J. Matrix Anal. Appl.
// 26, 1179-1193 (2005)), the algorithm behind arma::expmat for this purpose
// (HSPropagator.cpp:96, SpinSpace_operators.cpp:1294).
Mat expm(const Mat &A) {
    const int n = A.r;
    static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                                 1187353796428800.0,  129060195264000.0,   10559470521600.0,
                                 670442572800.0,      33522128640.0,       1323241920.0,
                                 40840800.0,          960960.0,            16380.0,
                                 182.0,               1.0};
    const double theta13 = 5.371920351148152;
    const double nrm = norm1(A);
    int s = 0;
    if (nrm > theta13) s = std::max(0, (int)std::ceil(std::log2(nrm / theta13)));
    const Mat As = scale(A, std::ldexp(1.0, -s));
    const Mat Id = Mat::eye(n);
    const Mat A2 = matmul(As, As), A4 = matmul(A2, A2), A6 = matmul(A4, A2);
    Mat U1 = add(add(scale(A6, b[13]), scale(A4, b[11])), scale(A2, b[9]));
    U1 = matmul(A6, U1);
    U1 = add(U1, add(add(add(scale(A6, b[7]), scale(A4, b[5])), scale(A2, b[3])), scale(Id, b[1])));
    const Mat U = matmul(As, U1);
    Mat V1 = add(add(scale(A6, b[12]), scale(A4, b[10])), scale(A2, b[8]));
    V1 = matmul(A6, V1);
    const Mat V = add(V1, add(add(add(scale(A6, b[6]), scale(A4, b[4])), scale(A2, b[2])), scale(Id, b[0])));
    Mat R = solve(add(V, U, -1.0), add(V, U, 1.0));
    for (int i = 0; i < s; ++i) R = matmul(R, R);
    return R;
}

// Hermitian eigen-decomposition by cyclic complex Jacobi rotations.  Each
// rotation removes |a_pq| after a phase turns it real; eigenvalues ascending.
void eig_hermitian(const Mat &Ain, std::vector<double> &w, Mat &V) {
    const int n = Ain.r;
    Mat A = Ain;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {  // symmetrise
            const cd h = 0.5 * (A(i, j) + std::conj(A(j, i)));
            A(i, j) = h; A(j, i) = std::conj(h);
        }
    for (int i = 0; i < n; ++i) A(i, i) = A(i, i).real();
    V = Mat::eye(n);
    double total = 0.0;
    for (auto &x : A.a) total += std::norm(x);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                if (i != j) off += std::norm(A(i, j));
        if (off <= 1e-30 * std::max(total, 1e-300)) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double g = std::abs(A(p, q));
                if (g < 1e-300) continue;
                const cd ph = A(p, q) / g;  // e^{i phi}
                const double app = A(p, p).real(), aqq = A(q, q).real();
                const double theta = 0.5 * std::atan2(2.0 * g, aqq - app);
                const double c = std::cos(theta), s = std::sin(theta);
                // V_rot = diag(1, e^{-i phi}) * [[c, s], [-s, c]]
                const cd v00 = c, v01 = s, v10 = -s * std::conj(ph), v11 = c * std::conj(ph);
                for (int k = 0; k < n; ++k) {  // columns
                    const cd akp = A(k, p), akq = A(k, q);
                    A(k, p) = akp * v00 + akq * v10;
                    A(k, q) = akp * v01 + akq * v11;
                }
                for (int k = 0; k < n; ++k) {  // rows
                    const cd apk = A(p, k), aqk = A(q, k);
                    A(p, k) = std::conj(v00) * apk + std::conj(v10) * aqk;
                    A(q, k) = std::conj(v01) * apk + std::conj(v11) * aqk;
                }
                A(p, q) = 0.0; A(q, p) = 0.0;
                A(p, p) = A(p, p).real(); A(q, q) = A(q, q).real();
                for (int k = 0; k < n; ++k) {
                    const cd vkp = V(k, p), vkq = V(k, q);
                    V(k, p) = vkp * v00 + vkq * v10;
                    V(k, q) = vkp * v01 + vkq * v11;
                }
            }
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return A(x, x).real() < A(y, y).real(); });
    w.resize(n);
    Mat Vs(n, n);
    for (int j = 0; j < n; ++j) {
        w[j] = A(order[j], order[j]).real();
        for (int i = 0; i < n; ++i) Vs(i, j) = V(i, order[j]);
    }
    V = Vs;
}

// ---------------------------------------------------------------------------
// Sparse CSR complex operator
// ---------------------------------------------------------------------------
struct CSR {
    int64_t n = 0;
    std::vector<int64_t> ptr;
    std::vector<int32_t> col;
    std::vector<cd> val;
    void matvec(const cd *x, cd *y) const {
        for (int64_t r = 0; r < n; ++r) {
            cd s = 0.0;
            for (int64_t k = ptr[r]; k < ptr[r + 1]; ++k) s += val[k] * x[col[k]];
            y[r] = s;
        }
    }
};
struct Triplets {
    int64_t n;
    std::vector<std::map<int32_t, cd>> rows;
    explicit Triplets(int64_t nn) : n(nn), rows(nn) {}
    void add(int64_t r, int32_t c, cd v) { if (v != 0.0) rows[r][c] += v; }
    CSR csr() const {
        CSR m; m.n = n; m.ptr.assign(n + 1, 0);
        for (int64_t r = 0; r < n; ++r) {
            for (auto &kv : rows[r]) if (kv.second != 0.0) { m.col.push_back(kv.first); m.val.push_back(kv.second); }
            m.ptr[r + 1] = (int64_t)m.col.size();
        }
        return m;
    }
};
Mat dense_of(const CSR &m) {
    Mat d((int)m.n, (int)m.n);
    for (int64_t r = 0; r < m.n; ++r)
        for (int64_t k = m.ptr[r]; k < m.ptr[r + 1]; ++k) d((int)r, m.col[k]) += m.val[k];
    return d;
}

// ---------------------------------------------------------------------------
// Spin system
// ---------------------------------------------------------------------------
struct System {
    const msg_system_desc *d;
    int nspins;
    std::vector<int> mult;
    std::vector<int64_t> stride;  // first spin most significant (kron order)
    int64_t D = 1;
    explicit System(const msg_system_desc *desc) : d(desc), nspins(desc->n_spins) {
        if (nspins <= 0) throw std::runtime_error("system has no spins");
        mult.resize(nspins); stride.resize(nspins);
        for (int i = 0; i < nspins; ++i) {
            mult[i] = d->spins[i].multiplicity;
            if (mult[i] < 1) throw std::runtime_error("invalid multiplicity");
        }
        for (int i = nspins - 1; i >= 0; --i) { stride[i] = D; D *= mult[i]; }
    }
    int digit(int64_t idx, int s) const { return (int)((idx / stride[s]) % mult[s]); }
};

// Single-spin matrices, basis |+S> first (SpinSpace_states.cpp:206-209):
// axis 0..2 = Sx, Sy, Sz.
Mat spin_matrix(int mult, int axis) {
    const double S = 0.5 * (mult - 1);
    Mat sp(mult, mult);
    for (int i = 1; i < mult; ++i) {  // S+ |m_i> -> |m_{i-1}>, m_i = S - i
        const double m = S - i;
        sp(i - 1, i) = std::sqrt(S * (S + 1) - m * (m + 1));
    }
    Mat out(mult, mult);
    for (int i = 0; i < mult; ++i)
        for (int j = 0; j < mult; ++j) {
            const cd p = sp(i, j), mm = std::conj(sp(j, i));
            if (axis == 0) out(i, j) = 0.5 * (p + mm);
            else if (axis == 1) out(i, j) = (p - mm) / (2.0 * I1);
            else out(i, j) = (i == j) ? cd(S - i) : cd(0.0);
        }
    return out;
}

// Adds coef * (opA on spin a) (opB on spin b) to T (b < 0: single-spin term);
// this is the CreateOperator embedding (SpinSpace_operators.cpp:65-106) of the
// products InteractionOperator forms.
void add_product(Triplets &T, const System &S, int a, const Mat &opA, int b, const Mat *opB, cd coef) {
    if (coef == 0.0) return;
    for (int64_t r = 0; r < S.D; ++r) {
        const int ra = S.digit(r, a);
        for (int ca = 0; ca < S.mult[a]; ++ca) {
            const cd va = opA(ra, ca);
            if (va == 0.0) continue;
            const int64_t c1 = r + (int64_t)(ca - ra) * S.stride[a];
            if (b < 0) { T.add(r, (int32_t)c1, coef * va); continue; }
            const int rb = S.digit(r, b);
            if (a == b) {  // same spin: product opA*opB acts on one index
                for (int cb = 0; cb < S.mult[b]; ++cb) {
                    const cd vb = (*opB)(ca, cb);
                    if (vb == 0.0) continue;
                    T.add(r, (int32_t)(r + (int64_t)(cb - ra) * S.stride[a]), coef * va * vb);
                }
                continue;
            }
            for (int cb = 0; cb < S.mult[b]; ++cb) {
                const cd vb = (*opB)(rb, cb);
                if (vb == 0.0) continue;
                T.add(r, (int32_t)(c1 + (int64_t)(cb - rb) * S.stride[b]), coef * va * vb);
            }
        }
    }
}

// The 3x3 matrix g with T_a = sum_c g[a][c] S_c (Spin::Tx/Ty/Tz with laboratory
// quantisation axes, Spin.cpp:360-408): the spin tensor, or the identity when
// the interaction ignores tensors.
void spin_t_matrix(const msg_spin_desc &sp, bool ignore, double g[9]) {
    for (int i = 0; i < 9; ++i) g[i] = ignore ? (i % 4 == 0 ? 1.0 : 0.0) : sp.tensor[i];
}

// Field of a time-dependent Zeeman interaction at time t (Interaction.cpp:1900-1933):
// linear  B = v cos(w t + phase);
// circular B = R (v or v - (a.v) a), R = 1 + sin(x) W + 2 sin^2(x/2) W^2, x = w t + phase,
// W the cross-product matrix of the axis a as given.
void td_field(const msg_interaction_desc &in, double t, double B[3]) {
    const double *v = in.field;
    if (in.time_dependence == MSG_TD_LINEAR) {
        const double c = std::cos(in.td_frequency * t + in.td_phase);
        for (int a = 0; a < 3; ++a) B[a] = v[a] * c;
        return;
    }
    const double *ax = in.td_axis;
    const double x = t * in.td_frequency + in.td_phase;
    const double W[9] = {0.0, -ax[2], ax[1], ax[2], 0.0, -ax[0], -ax[1], ax[0], 0.0};
    double R[9];
    const double sx = std::sin(x), h = 2.0 * std::sin(0.5 * x) * std::sin(0.5 * x);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double w2 = 0.0;
            for (int k = 0; k < 3; ++k) w2 += W[3 * i + k] * W[3 * k + j];
            R[3 * i + j] = (i == j ? 1.0 : 0.0) + sx * W[3 * i + j] + h * w2;
        }
    double b[3] = {v[0], v[1], v[2]};
    if (in.td_perpendicular) {
        const double d = ax[0] * v[0] + ax[1] * v[1] + ax[2] * v[2];
        for (int a = 0; a < 3; ++a) b[a] -= d * ax[a];
    }
    for (int i = 0; i < 3; ++i) B[i] = R[3 * i] * b[0] + R[3 * i + 1] * b[1] + R[3 * i + 2] * b[2];
}

// H at a field override; `time` (dynamics=dynamic) puts time-dependent Zeeman fields at
// their value at that time (SpinSpace::StaticHamiltonian + DynamicHamiltonian,
// SpinSpace_hamiltonians.cpp:1590-1740); without it they keep the field as written.
CSR build_hamiltonian(const System &S, const double *field_override, const double *time = nullptr) {
    Triplets T(S.D);
    const msg_system_desc *d = S.d;
    Mat ops[3][8];  // cached single-spin matrices per multiplicity (<= 8)
    auto sop = [&](int mult, int axis) -> const Mat & {
        if (mult > 8) throw std::runtime_error("multiplicity above 8 not supported by the oracle");
        Mat &m = ops[axis][mult - 1];
        if (m.r == 0) m = spin_matrix(mult, axis);
        return m;
    };
    for (int it = 0; it < d->n_interactions; ++it) {
        const msg_interaction_desc &in = d->interactions[it];
        // Per-interaction scale: isotropic coupling value, prefactor, common prefactor
        // (SpinSpace_hamiltonians.cpp:739-748).
        double sc = in.prefactor;
        if (in.tensor_present && in.tensor_isotropic) sc *= in.tensor[0];
        if (in.common_prefactor) sc *= kCommonPrefactor;
        const int32_t *g1 = d->index_pool + in.group1_offset;
        const int32_t *g2 = d->index_pool + in.group2_offset;
        if (in.type == MSG_INTERACTION_ZEEMAN) {
            // S . (G B) (SpinSpace_hamiltonians.cpp:450-479); G = spin tensor unless ignoretensors.
            double Bt[3];
            const double *B = (in.field_swept && field_override) ? field_override : in.field;
            if (time && in.time_dependence != MSG_TD_NONE) {
                td_field(in, *time, Bt);
                B = Bt;
            }
            for (int k = 0; k < in.group1_count; ++k) {
                const int s = g1[k];
                double G[9];
                spin_t_matrix(d->spins[s], in.ignore_tensors != 0, G);
                for (int a = 0; a < 3; ++a) {
                    const double ca = G[3 * a] * B[0] + G[3 * a + 1] * B[1] + G[3 * a + 2] * B[2];
                    add_product(T, S, s, sop(S.mult[s], a), -1, nullptr, sc * ca);
                }
            }
        } else if (in.type == MSG_INTERACTION_DOUBLESPIN || in.type == MSG_INTERACTION_EXCHANGE) {
            // sum_ab A_ab T1a T2b (anisotropic) or T1.T2 (SpinSpace_hamiltonians.cpp:481-535);
            // exchange doubles it and adds I/2 per pair (:567-631).
            const bool ex = in.type == MSG_INTERACTION_EXCHANGE;
            const bool aniso = in.tensor_present && !in.tensor_isotropic;
            for (int i = 0; i < in.group1_count; ++i)
                for (int j = 0; j < in.group2_count; ++j) {
                    const int s1 = g1[i], s2 = g2[j];
                    double ga[9], gb[9];
                    spin_t_matrix(d->spins[s1], in.ignore_tensors != 0, ga);
                    spin_t_matrix(d->spins[s2], in.ignore_tensors != 0, gb);
                    // coefficient of S1_c S2_e: sum_ab A_ab ga[a][c] gb[b][e]
                    for (int c = 0; c < 3; ++c)
                        for (int e = 0; e < 3; ++e) {
                            double m = 0.0;
                            for (int a = 0; a < 3; ++a)
                                for (int b = 0; b < 3; ++b) {
                                    const double A = aniso ? in.tensor[3 * a + b] : (a == b ? 1.0 : 0.0);
                                    m += A * ga[3 * a + c] * gb[3 * b + e];
                                }
                            if (ex) m *= 2.0;
                            if (m != 0.0)
                                add_product(T, S, s1, sop(S.mult[s1], c), s2, &sop(S.mult[s2], e), sc * m);
                        }
                    if (ex)
                        for (int64_t r = 0; r < S.D; ++r) T.add(r, (int32_t)r, sc * 0.5);
                }
        } else {
            throw std::runtime_error("interaction type not supported by the oracle");
        }
    }
    return T.csr();
}

// Projector onto a State (SpinSpace::GetState, SpinSpace_states.cpp:626-700):
// |v><v|/<v|v> on the listed spins, identity on the others.
CSR state_projector(const System &S, int state) {
    const msg_system_desc *d = S.d;
    if (state < 0 || state >= d->n_states) throw std::runtime_error("state index out of range");
    const msg_state_desc &st = d->states[state];
    const int32_t *sp = d->index_pool + st.spins_offset;
    const int ns = st.spins_count;
    int64_t L = 1;
    for (int k = 0; k < ns; ++k) L *= S.mult[sp[k]];
    std::vector<cd> v(L);
    double nn = 0.0;
    for (int64_t i = 0; i < L; ++i) {
        v[i] = cd(d->amp_pool[2 * (st.amp_offset + i)], d->amp_pool[2 * (st.amp_offset + i) + 1]);
        nn += std::norm(v[i]);
    }
    if (!(nn > 0.0)) throw std::runtime_error("state with zero norm");
    Triplets T(S.D);
    for (int64_t r = 0; r < S.D; ++r) {
        int64_t sub = 0, base = r;
        for (int k = 0; k < ns; ++k) {
            const int dg = S.digit(r, sp[k]);
            sub = sub * S.mult[sp[k]] + dg;
            base -= (int64_t)dg * S.stride[sp[k]];
        }
        if (v[sub] == 0.0) continue;
        for (int64_t c = 0; c < L; ++c) {
            if (v[c] == 0.0) continue;
            int64_t rem = c, idx = base;
            for (int k = ns - 1; k >= 0; --k) {
                const int m = S.mult[sp[k]];
                idx += (rem % m) * S.stride[sp[k]];
                rem /= m;
            }
            T.add(r, (int32_t)idx, v[sub] * std::conj(v[c]) / nn);
        }
    }
    return T.csr();
}

// Direct factors of rho0 (HSStatePreparation.cpp:77-165 then
// FactorizeDensityMatrix, SpinSpace_states.cpp:570-622).  A single initial
// State gives rho0 = P/Tr P, whose factors are (v/|v|) (x) e_n / sqrt(Tr P)
// over the basis e_n of the spins the State leaves open; several States are
// factorised through a dense eigen-decomposition.
Mat direct_factors(const System &S) {
    const msg_system_desc *d = S.d;
    if (d->n_initial < 1) throw std::runtime_error("no initial state");
    if (d->n_initial == 1) {
        const msg_state_desc &st = d->states[d->initial_states[0]];
        const int32_t *sp = d->index_pool + st.spins_offset;
        const int ns = st.spins_count;
        std::vector<char> fixed(S.nspins, 0);
        int64_t L = 1;
        for (int k = 0; k < ns; ++k) { fixed[sp[k]] = 1; L *= S.mult[sp[k]]; }
        const int64_t Z = S.D / L;
        std::vector<cd> v(L);
        double nn = 0.0;
        for (int64_t i = 0; i < L; ++i) {
            v[i] = cd(d->amp_pool[2 * (st.amp_offset + i)], d->amp_pool[2 * (st.amp_offset + i) + 1]);
            nn += std::norm(v[i]);
        }
        const double sc = 1.0 / std::sqrt(nn * (double)Z);
        Mat B((int)S.D, (int)Z);
        std::vector<int> open;
        for (int s = 0; s < S.nspins; ++s) if (!fixed[s]) open.push_back(s);
        for (int64_t n = 0; n < Z; ++n) {
            int64_t rem = n, base = 0;
            for (int k = (int)open.size() - 1; k >= 0; --k) {
                const int m = S.mult[open[k]];
                base += (rem % m) * S.stride[open[k]];
                rem /= m;
            }
            for (int64_t c = 0; c < L; ++c) {
                if (v[c] == 0.0) continue;
                int64_t r2 = c, idx = base;
                for (int k = ns - 1; k >= 0; --k) {
                    const int m = S.mult[sp[k]];
                    idx += (r2 % m) * S.stride[sp[k]];
                    r2 /= m;
                }
                B((int)idx, (int)n) = v[c] * sc;
            }
        }
        return B;
    }
    std::vector<double> w(d->n_initial, 1.0 / d->n_initial);
    if (d->initial_weights) {
        double t = 0.0;
        for (int i = 0; i < d->n_initial; ++i) { w[i] = d->initial_weights[i]; t += w[i]; }
        for (auto &x : w) x /= t;
    }
    Mat rho((int)S.D, (int)S.D);
    for (int i = 0; i < d->n_initial; ++i) {
        const CSR P = state_projector(S, d->initial_states[i]);
        cd tr = 0.0;
        for (int64_t r = 0; r < P.n; ++r)
            for (int64_t k = P.ptr[r]; k < P.ptr[r + 1]; ++k) if (P.col[k] == r) tr += P.val[k];
        for (int64_t r = 0; r < P.n; ++r)
            for (int64_t k = P.ptr[r]; k < P.ptr[r + 1]; ++k) rho((int)r, P.col[k]) += w[i] * P.val[k] / tr;
    }
    std::vector<double> ev;
    Mat V;
    eig_hermitian(rho, ev, V);
    const double tol = std::max(1e-14, 1e-10 * std::abs(ev.back()));
    std::vector<int> keep;
    for (int j = 0; j < (int)ev.size(); ++j) if (ev[j] > tol) keep.push_back(j);
    Mat B((int)S.D, (int)keep.size());
    for (int k = 0; k < (int)keep.size(); ++k)
        for (int i = 0; i < (int)S.D; ++i) B(i, k) = std::sqrt(ev[keep[k]]) * V(i, keep[k]);
    return B;
}

// ---------------------------------------------------------------------------
// Time grid and integration
// ---------------------------------------------------------------------------
// BuildFreeEvolutionTimes (TaskHSGeneral.cpp:59-96).
std::vector<double> free_times(double T, double dt) {
    std::vector<double> t{0.0};
    if (T == 0.0) return t;
    const double ratio = T / dt;
    if (!std::isfinite(ratio)) throw std::runtime_error("too many time intervals");
    const double nearest = std::round(ratio);
    const double tol = 64.0 * 2.220446049250313e-16 * std::max(1.0, std::abs(ratio));
    const size_t n = (size_t)((nearest >= 1.0 && std::abs(ratio - nearest) <= tol) ? nearest : std::ceil(ratio));
    for (size_t s = 1; s <= n; ++s) {
        const double target = s == n ? T : std::min(T, (double)s * dt);
        if (!(target > t.back())) throw std::runtime_error("non-increasing time grid");
        t.push_back(target);
    }
    return t;
}

// ---------------------------------------------------------------------------
// Krylov (Arnoldi) step: SpinSpace::KrylovExpmGeneral + ArnoldiProcess
// (SpinSpace_operators.cpp:1257-1304, 1352-1408) with KrylovDimTol = 0.
// ---------------------------------------------------------------------------
void krylov_expm_general(const CSR &G, const cd *b, double dt, int m, cd *out) {
    const int64_t D = G.n;
    m = std::max(1, (int)std::min<int64_t>(m, D));
    double beta = 0.0;
    for (int64_t i = 0; i < D; ++i) beta += std::norm(b[i]);
    beta = std::sqrt(beta);
    if (beta == 0.0) { std::fill(out, out + D, cd(0.0)); return; }
    std::vector<cd> V((size_t)D * m), z(D);
    Mat H(m, m);
    for (int64_t i = 0; i < D; ++i) V[i] = b[i] / beta;
    int actual = 0;
    for (int j = 0; j < m; ++j) {
        G.matvec(&V[(size_t)j * D], z.data());
        for (int i = 0; i <= j; ++i) {  // modified Gram-Schmidt, h = cdot(v_i, z)
            const cd *vi = &V[(size_t)i * D];
            cd h = 0.0;
            for (int64_t k = 0; k < D; ++k) h += std::conj(vi[k]) * z[k];
            H(i, j) = h;
            for (int64_t k = 0; k < D; ++k) z[k] -= h * vi[k];
        }
        double zn = 0.0;
        for (int64_t k = 0; k < D; ++k) zn += std::norm(z[k]);
        zn = std::sqrt(zn);
        if (j == m - 1) { actual = m; break; }
        H(j + 1, j) = zn;
        if (zn < 1e-14) { actual = j + 1; break; }  // happy breakdown
        for (int64_t k = 0; k < D; ++k) V[(size_t)(j + 1) * D + k] = z[k] / zn;
        actual = j + 1;
    }
    Mat Hs(actual, actual);
    for (int j = 0; j < actual; ++j)
        for (int i = 0; i < actual; ++i) Hs(i, j) = H(i, j) * dt;
    const Mat E = expm(Hs);
    std::fill(out, out + D, cd(0.0));
    for (int j = 0; j < actual; ++j) {
        const cd w = beta * E(j, 0);
        const cd *vj = &V[(size_t)j * D];
        for (int64_t k = 0; k < D; ++k) out[k] += w * vj[k];
    }
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the relaxation flips.
// ---------------------------------------------------------------------------
inline void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
    const uint64_t p = (uint64_t)a * b;
    hi = (uint32_t)(p >> 32); lo = (uint32_t)p;
}
void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
    uint32_t c[4] = {c_in[0], c_in[1], c_in[2], c_in[3]};
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(0xD2511F53u, c[0], hi0, lo0);
        mulhilo(0xCD9E8D57u, c[2], hi1, lo1);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

// Flip threshold for one relaxation half-step: an ST-dephasing event of rate
// lambda = r/2 changes the parity with p = -expm1(-2 lambda h)/2
// (SpinSpace_relaxation.cpp:1211-1236), drawn as x < floor(p 2^32).
uint32_t flip_threshold(double rate, double h) {
    const double p = -0.5 * std::expm1(-2.0 * (0.5 * rate) * h);
    return (uint32_t)std::floor(p * 4294967296.0);
}
bool flip_draw(uint64_t seed, uint32_t item, uint32_t column, uint32_t step, uint32_t half, uint32_t thr) {
    const uint32_t ctr[4] = {column, step, half, item};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    philox(ctr, key, o);
    return o[0] < thr;
}

// Several relaxation channels (SpinSpace_relaxation.cpp:1237-1252): the event count of
// the superposed Poisson process of mean mu = sum_j lambda_j h, lambda_j = rate_j / 2, by
// inversion of the Philox word x (count > n exactly when x >= ceil(F(n) 2^32), F the
// Poisson CDF), then per event a channel drawn from the next words with probabilities
// lambda_j / sum lambda (cumulative search, word / 2^32); applies the events in order.
void relax_events(const std::vector<std::pair<double, CSR>> &ch, double h, uint64_t seed, uint32_t item,
                  uint32_t column, uint32_t step, uint32_t half, cd *b, std::vector<cd> &tmp) {
    const uint32_t ctr[4] = {column, step, half, item};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    philox(ctr, key, o);
    double tot = 0.0;
    for (const auto &c : ch) tot += 0.5 * c.first;
    const double mu = tot * h;
    // thresholds T_n = ceil(F(n) 2^32) held in 32 bits (2^32 - 1 at most): count 1 on
    // [T_0, T_1), 3 on [T_2, 2^32 - 1), 0 below T_0 and 2 otherwise (counts above 3 have
    // probability below mu^4 / 24; the word 2^32 - 1 counts as even)
    uint32_t T[3];
    double term = std::exp(-mu), F = term;
    for (int n = 0; n < 3; ++n) {
        T[n] = (uint32_t)std::min(4294967295.0, std::ceil(F * 4294967296.0));
        term *= mu / (n + 1);
        F += term;
    }
    const uint32_t x = o[0];
    const int count = x < T[0] ? 0 : x < T[1] ? 1 : (x >= T[2] && x < 0xFFFFFFFFu) ? 3 : 2;
    for (int e = 0; e < count; ++e) {
        const double v = (double)o[1 + e] / 4294967296.0 * tot;
        size_t j = 0;
        double cum = 0.5 * ch[0].first;
        while (j + 1 < ch.size() && v >= cum) cum += 0.5 * ch[++j].first;
        ch[j].second.matvec(b, tmp.data());
        std::copy(tmp.begin(), tmp.begin() + ch[j].second.n, b);
    }
}

// U = P_S - P_T for the relaxationdephasing pair (SpinSpace_relaxation.cpp:1144-1199).
CSR st_unitary(const System &S, int s1, int s2) {
    if (S.mult[s1] != 2 || S.mult[s2] != 2) throw std::runtime_error("relaxationdephasing needs spin-1/2 spins");
    Triplets T(S.D);
    Mat sx = spin_matrix(2, 0), sy = spin_matrix(2, 1), sz = spin_matrix(2, 2);
    // P_S = 1/4 - S1.S2, P_T = 1 - P_S, U = P_S - P_T = 2 P_S - 1 = -1/2 - 2 S1.S2
    add_product(T, S, s1, sx, s2, &sx, -2.0);
    add_product(T, S, s1, sy, s2, &sy, -2.0);
    add_product(T, S, s1, sz, s2, &sz, -2.0);
    for (int64_t r = 0; r < S.D; ++r) T.add(r, (int32_t)r, -0.5);
    return T.csr();
}

// ---------------------------------------------------------------------------
// HSGeneral free-evolution loop for one item over a block of columns.
// ---------------------------------------------------------------------------
struct ObsSet {
    std::vector<CSR> ops;
    std::vector<double> scale;  // rate for transition yields, 1 for populations
    double kmax = 0.0;
};
ObsSet observables(const System &S, int mode) {
    ObsSet o;
    const msg_system_desc *d = S.d;
    if (mode == MSG_OBS_TRANSITION_YIELDS) {
        for (int t = 0; t < d->n_transitions; ++t) {
            if (d->transitions[t].source_state < 0) continue;
            o.ops.push_back(state_projector(S, d->transitions[t].source_state));
            o.scale.push_back(d->transitions[t].rate);
            o.kmax = std::max(o.kmax, d->transitions[t].rate);
        }
    } else {
        for (int s = 0; s < d->n_states; ++s) { o.ops.push_back(state_projector(S, s)); o.scale.push_back(1.0); }
    }
    return o;
}

CSR reaction_operator(const System &S) {
    Triplets T(S.D);
    const msg_system_desc *d = S.d;
    for (int t = 0; t < d->n_transitions; ++t) {
        if (d->transitions[t].source_state < 0) continue;
        const CSR P = state_projector(S, d->transitions[t].source_state);
        for (int64_t r = 0; r < P.n; ++r)
            for (int64_t k = P.ptr[r]; k < P.ptr[r + 1]; ++k)
                T.add(r, P.col[k], 0.5 * d->transitions[t].rate * P.val[k]);
    }
    return T.csr();
}

struct ItemModel {
    CSR G;       // -iH - K
    Mat Gdense;  // for EXPONENTIAL
    bool has_relax = false;
    CSR U;
    double relax_rate = 0.0;
    std::vector<std::pair<double, CSR>> channels;  // (rate, U) of every relaxation channel
};

ItemModel make_model(const System &S, const double *field, int method, const double *time = nullptr) {
    ItemModel m;
    const CSR H = build_hamiltonian(S, field, time);
    const CSR K = reaction_operator(S);
    Triplets T(S.D);
    for (int64_t r = 0; r < H.n; ++r)
        for (int64_t k = H.ptr[r]; k < H.ptr[r + 1]; ++k) T.add(r, H.col[k], -I1 * H.val[k]);
    for (int64_t r = 0; r < K.n; ++r)
        for (int64_t k = K.ptr[r]; k < K.ptr[r + 1]; ++k) T.add(r, K.col[k], -K.val[k]);
    m.G = T.csr();
    if (method == MSG_METHOD_EXPONENTIAL) m.Gdense = dense_of(m.G);
    const msg_system_desc *d = S.d;
    // relaxationdephasing channels with a nonzero rate (PrepareStochasticRelaxationHilbert,
    // SpinSpace_relaxation.cpp:1171-1199)
    for (int i = 0; i < d->n_relaxations; ++i) {
        const auto &rl = d->relaxations[i];
        if (rl.rate == 0.0) continue;
        if (!(rl.rate > 0.0)) throw std::runtime_error("invalid stochastic relaxation rate");
        m.channels.push_back({rl.rate, st_unitary(S, rl.spin1, rl.spin2)});
    }
    if (!m.channels.empty()) {
        m.has_relax = true;
        m.relax_rate = m.channels[0].first;
        m.U = m.channels[0].second;
    }
    return m;
}

// Propagates columns [c0, c1) of B0 over the grid; adds sum-over-columns
// observables into vals ((nsteps+1) x nobs).
// dynamic (dynamics=dynamic, TaskHSGeneral.cpp:487-537): each interval uses the model at
// its midpoint, or with RK4 the stages at t, t+dt/2, t+dt; `field` is the item's field.
void run_block(const System &S, const ItemModel &M, const ObsSet &O, const std::vector<double> &times,
               int64_t nsamples, const Mat &B0, int c0, int c1, int method, int krylov, uint64_t seed,
               uint32_t item, std::vector<double> &vals, bool dynamic = false, const double *field = nullptr) {
    const int64_t D = S.D;
    const int nobs = (int)O.ops.size();
    const int ncol = c1 - c0;
    std::vector<cd> B((size_t)D * ncol), tmp(D), k1(D), k2(D), k3(D), k4(D), y(D);
    for (int c = 0; c < ncol; ++c)
        std::copy(&B0.a[(size_t)(c0 + c) * D], &B0.a[(size_t)(c0 + c + 1) * D], &B[(size_t)c * D]);
    Mat U;
    double Udt = -1.0;
    for (int64_t s = 0; s < nsamples; ++s) {
        // HSObservableCollector::Evaluate (HSObservableCollector.cpp:276-288)
        for (int o = 0; o < nobs; ++o) {
            double acc = 0.0;
            for (int c = 0; c < ncol; ++c) {
                O.ops[o].matvec(&B[(size_t)c * D], tmp.data());
                const cd *b = &B[(size_t)c * D];
                for (int64_t k = 0; k < D; ++k) acc += (std::conj(b[k]) * tmp[k]).real();
            }
            vals[(size_t)s * nobs + o] += O.scale[o] * acc;
        }
        if (s + 1 >= nsamples) break;
        const double dt = times[s + 1] - times[s];
        const uint32_t thr = M.has_relax ? flip_threshold(M.relax_rate, 0.5 * dt) : 0u;
        ItemModel Mt[3];  // dynamic: midpoint (or RK4 start, midpoint, end)
        const ItemModel *Gs[3] = {&M, &M, &M};
        if (dynamic) {
            const double tm = times[s] + 0.5 * dt;
            if (method == MSG_METHOD_RK4) {
                const double ts[3] = {times[s], tm, times[s + 1]};
                for (int k = 0; k < 3; ++k) { Mt[k] = make_model(S, field, method, &ts[k]); Gs[k] = &Mt[k]; }
            } else {
                Mt[1] = make_model(S, field, method, &tm);
                Gs[0] = Gs[1] = Gs[2] = &Mt[1];
                Udt = -1.0;  // a new propagator every interval
            }
        }
        const ItemModel &Mm = *Gs[1];
        for (int c = 0; c < ncol; ++c) {
            cd *b = &B[(size_t)c * D];
            // StepStochastic: R(dt/2) Step(dt) R(dt/2) (HSPropagator.cpp:131-140)
            if (M.channels.size() > 1) {
                relax_events(M.channels, 0.5 * dt, seed, item, (uint32_t)(c0 + c), (uint32_t)s, 0u, b, tmp);
            } else if (M.has_relax && flip_draw(seed, item, (uint32_t)(c0 + c), (uint32_t)s, 0u, thr)) {
                M.U.matvec(b, tmp.data()); std::copy(tmp.begin(), tmp.end(), b);
            }
            if (method == MSG_METHOD_KRYLOV) {
                krylov_expm_general(Mm.G, b, dt, krylov, y.data());
                std::copy(y.begin(), y.end(), b);
            } else if (method == MSG_METHOD_EXPONENTIAL) {
                if (dt != Udt) {
                    U = expm(scale(Mm.Gdense, dt));
                    Udt = dt;
                }
                for (int64_t i = 0; i < D; ++i) {
                    cd acc = 0.0;
                    for (int64_t k = 0; k < D; ++k) acc += U((int)i, (int)k) * b[k];
                    y[i] = acc;
                }
                std::copy(y.begin(), y.end(), b);
            } else {  // RK4 (HSPropagator.cpp:80-90; time-dependent stages :104-129)
                Gs[0]->G.matvec(b, k1.data());
                for (int64_t i = 0; i < D; ++i) tmp[i] = b[i] + 0.5 * dt * k1[i];
                Gs[1]->G.matvec(tmp.data(), k2.data());
                for (int64_t i = 0; i < D; ++i) tmp[i] = b[i] + 0.5 * dt * k2[i];
                Gs[1]->G.matvec(tmp.data(), k3.data());
                for (int64_t i = 0; i < D; ++i) tmp[i] = b[i] + dt * k3[i];
                Gs[2]->G.matvec(tmp.data(), k4.data());
                for (int64_t i = 0; i < D; ++i) b[i] += (dt / 6.0) * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
            }
            if (M.channels.size() > 1) {
                relax_events(M.channels, 0.5 * dt, seed, item, (uint32_t)(c0 + c), (uint32_t)s, 1u, b, tmp);
            } else if (M.has_relax && flip_draw(seed, item, (uint32_t)(c0 + c), (uint32_t)s, 1u, thr)) {
                M.U.matvec(b, tmp.data()); std::copy(tmp.begin(), tmp.end(), b);
            }
        }
    }
}

// Minimum-norm least-squares solution of L x = b (row-major n x n), through the
// eigenpairs of L^dagger L: x = W diag(1/lambda) W^dagger L^dagger b over the
// eigenvalues above 1e-12 lambda_max (singular values above 1e-6 sigma_max; the
// eigenvalues of the null space come out at the rounding level, ~n eps lambda_max).
std::vector<cd> min_norm_solve(const std::vector<cd> &L, const std::vector<cd> &b, int n) {
    Mat A(n, n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            cd acc = 0.0;
            for (int k = 0; k < n; ++k) acc += std::conj(L[(size_t)i * n + k]) * L[(size_t)j * n + k];
            A(i, j) = acc;
        }
    std::vector<double> lam;
    Mat W;
    eig_hermitian(A, lam, W);
    double lmax = 0.0;
    for (double l : lam) lmax = std::max(lmax, std::fabs(l));
    std::vector<cd> c(n, cd(0.0)), x(n, cd(0.0));  // c = L^dagger b
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) c[i] += std::conj(L[(size_t)i * n + k]) * b[k];
    for (int m = 0; m < n; ++m) {
        if (!(lam[m] > 1e-12 * lmax)) continue;
        cd proj = 0.0;
        for (int i = 0; i < n; ++i) proj += std::conj(W(i, m)) * c[i];
        proj /= lam[m];
        for (int i = 0; i < n; ++i) x[i] += W(i, m) * proj;
    }
    return x;
}

// X = int_0^inf e^{Gt} rho0 e^{G^dagger t} dt from G X + X G^dagger = -rho0, solved as
// the D^2 x D^2 linear system of SpinSpace::SolveHilbertTimeIntegral
// (SpinSpace_operators.cpp:142-201): Liouvillian kron(G, 1) + kron(1, conj G) over the
// row-major vectorisation X_ij -> x[i D + j], right-hand side -vec(rho0), LU with partial
// pivoting (the reference's arma::solve).  Column-oriented elimination with OpenMP over
// the trailing columns.  A singular system (a part of the space that never decays,
// e.g. triplets without a sink or mixing; detected as a pivot below 1e-13 of the largest
// entry) takes the minimum-norm least-squares solution, arma::solve's fallback for
// singular systems (D <= 32 here): the never-decaying part then stays empty.
Mat lyapunov_integral(const Mat &G, const Mat &rho0) {
    const int D = G.r;
    const int n = D * D;
    if (n > 4096) throw std::runtime_error("time-integral oracle limited to D <= 64");
    std::vector<cd> L((size_t)n * n, cd(0.0)), x(n);
    auto at = [&](int i, int j) -> cd & { return L[(size_t)j * n + i]; };  // column-major
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            const cd g = G(i, j), gc = std::conj(G(i, j));
            for (int k = 0; k < D; ++k) {
                at(i * D + k, j * D + k) += g;   // kron(G, 1)
                at(k * D + i, k * D + j) += gc;  // kron(1, conj G)
            }
        }
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) x[(size_t)i * D + j] = -rho0(i, j);
    const std::vector<cd> L0 = n <= 1024 ? L : std::vector<cd>(), b0 = x;
    auto to_mat = [&](const std::vector<cd> &v) {
        Mat X(D, D);
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) X(i, j) = v[(size_t)i * D + j];
        return X;
    };
    double amax = 0.0;
    for (const cd &v : L) amax = std::max(amax, std::abs(v));
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::abs(at(k, k));
        for (int i = k + 1; i < n; ++i)
            if (std::abs(at(i, k)) > best) { best = std::abs(at(i, k)); p = i; }
        if (best <= 1e-13 * amax) {  // singular up to rounding
            if (L0.empty()) throw std::runtime_error("failed to solve the static Hilbert time-integral equation");
            std::vector<cd> Lt((size_t)n * n);  // row-major copy for min_norm_solve
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) Lt[(size_t)i * n + j] = L0[(size_t)j * n + i];
            return to_mat(min_norm_solve(Lt, b0, n));
        }
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(at(k, j), at(p, j));
            std::swap(x[k], x[p]);
        }
        const cd inv = 1.0 / at(k, k);
        for (int i = k + 1; i < n; ++i) at(i, k) *= inv;
        const cd *lk = &L[(size_t)k * n];
#pragma omp parallel for schedule(static)
        for (int j = k + 1; j < n; ++j) {
            cd *cj = &L[(size_t)j * n];
            const cd akj = cj[k];
            if (akj == 0.0) continue;
            for (int i = k + 1; i < n; ++i) cj[i] -= lk[i] * akj;
        }
        const cd xk = x[k];
        if (xk != 0.0)
            for (int i = k + 1; i < n; ++i) x[i] -= lk[i] * xk;
    }
    for (int i = n - 1; i >= 0; --i) {
        cd s = x[i];
        for (int k = i + 1; k < n; ++k) s -= at(i, k) * x[k];
        x[i] = s / at(i, i);
    }
    return to_mat(x);
}

Mat load_columns(int64_t D, const double *cols, int m) {
    Mat B((int)D, m);
    for (size_t i = 0; i < B.a.size(); ++i) B.a[i] = cd(cols[2 * i], cols[2 * i + 1]);
    return B;
}

void store(const Mat &M, double *out) {
    for (size_t i = 0; i < M.a.size(); ++i) { out[2 * i] = M.a[i].real(); out[2 * i + 1] = M.a[i].imag(); }
}

thread_local std::string g_error;

template <class F>
int guarded(F f) {
    try { f(); return 0; }
    catch (const std::exception &e) { g_error = e.what(); return 1; }
}

}  // namespace

extern "C" {

const char *ref_last_error(void) { return g_error.c_str(); }

int64_t ref_dimension(const msg_system_desc *sys) {
    int64_t D = -1;
    guarded([&] { D = System(sys).D; });
    return D;
}

int ref_hamiltonian_dense(const msg_system_desc *sys, const double *field, double *out) {
    return guarded([&] { System S(sys); store(dense_of(build_hamiltonian(S, field)), out); });
}

int ref_state_projector_dense(const msg_system_desc *sys, int state, double *out) {
    return guarded([&] { System S(sys); store(dense_of(state_projector(S, state)), out); });
}

int ref_reaction_dense(const msg_system_desc *sys, double *out) {
    return guarded([&] { System S(sys); store(dense_of(reaction_operator(S)), out); });
}

int ref_direct_factors(const msg_system_desc *sys, double *out, int32_t *n_columns) {
    return guarded([&] {
        System S(sys);
        const Mat B = direct_factors(S);
        store(B, out);
        *n_columns = B.c;
    });
}

int ref_run_hsgeneral(const msg_system_desc *sys, const msg_run_params *p, int32_t n_items,
                      const double *fields, const double *columns, int32_t n_columns,
                      double *yields, double *trace, int threads) {
    return guarded([&] {
        System S(sys);
        const std::vector<double> times = free_times(p->total_time, p->time_step);
        const int64_t ns = (int64_t)times.size();
        const ObsSet O = observables(S, p->observables);
        const int nobs = (int)O.ops.size();
        if (nobs == 0) throw std::runtime_error("no observables");
        const Mat B0 = p->sampling == MSG_SAMPLING_GIVEN ? load_columns(S.D, columns, n_columns) : direct_factors(S);
        const int method = p->method == MSG_METHOD_AUTO ? MSG_METHOD_KRYLOV : p->method;
        const int kry = p->krylov_size > 0 ? p->krylov_size : 16;
        const int nthreads = std::max(1, threads);
        std::vector<std::vector<double>> item_vals(n_items, std::vector<double>((size_t)ns * nobs, 0.0));
        std::vector<ItemModel> models(n_items);
        for (int it = 0; it < n_items; ++it) models[it] = make_model(S, fields ? fields + 3 * it : nullptr, method);
        const int ncol = B0.c;
        const int blocks = std::max(1, std::min(ncol, nthreads));
        std::string err;
#pragma omp parallel for collapse(2) schedule(dynamic) num_threads(nthreads)
        for (int it = 0; it < n_items; ++it)
            for (int blk = 0; blk < blocks; ++blk) {
                const int c0 = (int)((int64_t)ncol * blk / blocks), c1 = (int)((int64_t)ncol * (blk + 1) / blocks);
                if (c1 <= c0) continue;
                std::vector<double> v((size_t)ns * nobs, 0.0);
                try {
                    run_block(S, models[it], O, times, ns, B0, c0, c1, method, kry, p->seed, (uint32_t)it, v,
                              p->dynamics != 0, fields ? fields + 3 * it : nullptr);
                } catch (const std::exception &e) {
#pragma omp critical
                    err = e.what();
                }
#pragma omp critical
                for (size_t k = 0; k < v.size(); ++k) item_vals[it][k] += v[k];
            }
        if (!err.empty()) throw std::runtime_error(err);
        for (int it = 0; it < n_items; ++it) {
            const std::vector<double> &v = item_vals[it];
            if (trace) std::copy(v.begin(), v.end(), trace + (size_t)it * ns * nobs);
            if (yields) {
                for (int o = 0; o < nobs; ++o) {
                    // IntegrateFiniteTime (HSObservableCollector.cpp:304-326)
                    double acc = 0.0;
                    for (int64_t s = 1; s < ns; ++s)
                        acc += 0.5 * (times[s] - times[s - 1]) * (v[(size_t)(s - 1) * nobs + o] + v[(size_t)s * nobs + o]);
                    if (p->yield_corrections) {  // ApplyFiniteTimeYieldCorrection (:361-392)
                        const double den = 1.0 - std::exp(-p->total_time * O.kmax);
                        if (!(den > 0.0)) throw std::runtime_error("yieldcorrections needs a positive rate and time");
                        acc /= den;
                    }
                    yields[(size_t)it * nobs + o] = acc;
                }
            }
        }
    });
}

int ref_run_columns(const msg_system_desc *sys, const msg_run_params *p, const double *field,
                    const double *columns, int32_t n_columns, int32_t col_begin, int32_t col_end,
                    int64_t max_steps, double *values, int threads) {
    return guarded([&] {
        System S(sys);
        std::vector<double> times = free_times(p->total_time, p->time_step);
        const int64_t ns = std::min<int64_t>((int64_t)times.size(), max_steps + 1);
        const ObsSet O = observables(S, p->observables);
        const int nobs = (int)O.ops.size();
        const Mat B0 = p->sampling == MSG_SAMPLING_GIVEN ? load_columns(S.D, columns, n_columns) : direct_factors(S);
        if (col_begin < 0 || col_end > B0.c || col_end <= col_begin) throw std::runtime_error("bad column range");
        const int method = p->method == MSG_METHOD_AUTO ? MSG_METHOD_KRYLOV : p->method;
        const int kry = p->krylov_size > 0 ? p->krylov_size : 16;
        const ItemModel M = make_model(S, field, method);
        const int nthreads = std::max(1, threads);
        const int ncol = col_end - col_begin;
        const int blocks = std::max(1, std::min(ncol, nthreads));
        std::vector<double> total((size_t)ns * nobs, 0.0);
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
        for (int blk = 0; blk < blocks; ++blk) {
            const int c0 = col_begin + (int)((int64_t)ncol * blk / blocks);
            const int c1 = col_begin + (int)((int64_t)ncol * (blk + 1) / blocks);
            std::vector<double> v((size_t)ns * nobs, 0.0);
            if (c1 > c0) run_block(S, M, O, times, ns, B0, c0, c1, method, kry, p->seed, 0u, v, p->dynamics != 0, field);
#pragma omp critical
            for (size_t k = 0; k < v.size(); ++k) total[k] += v[k];
        }
        std::copy(total.begin(), total.end(), values);
    });
}

int ref_symmetric_decay(const msg_system_desc *sys, const double *field, double k, double *yields) {
    return guarded([&] {
        System S(sys);
        const Mat H = dense_of(build_hamiltonian(S, field));
        // rho0 = sum P_init / Tr (TaskStaticHSSymmetricDecay.cpp:98-112)
        Mat rho((int)S.D, (int)S.D);
        for (int i = 0; i < sys->n_initial; ++i) rho = add(rho, dense_of(state_projector(S, sys->initial_states[i])));
        cd tr = 0.0;
        for (int i = 0; i < rho.r; ++i) tr += rho(i, i);
        rho = scale(rho, 1.0 / tr);
        std::vector<double> lam;
        Mat V;
        eig_hermitian(H, lam, V);
        Mat Vh((int)S.D, (int)S.D);
        for (int i = 0; i < V.r; ++i)
            for (int j = 0; j < V.c; ++j) Vh(j, i) = std::conj(V(i, j));
        const Mat rhoE = matmul(Vh, matmul(rho, V));
        for (int st = 0; st < sys->n_states; ++st) {
            const Mat PE = matmul(Vh, matmul(dense_of(state_projector(S, st)), V));
            double acc = 0.0;
            for (int h = 0; h < V.c; ++h)
                for (int g = 0; g < V.c; ++g) {
                    const double w = lam[g] - lam[h];
                    acc += (PE(g, h) * rhoE(h, g)).real() / (1.0 + w * w / (k * k));
                }
            yields[st] = acc;
        }
    });
}

int ref_time_infinity(const msg_system_desc *sys, const msg_run_params *p, const double *field,
                      const double *columns, int32_t n_columns, double *yields) {
    return guarded([&] {
        System S(sys);
        if (sys->n_relaxations > 0) throw std::runtime_error("the time-integral oracle covers systems without relaxation");
        const ObsSet O = observables(S, p->observables);
        // rho0 = factors factors^dagger (TaskHSGeneral.cpp:455-456)
        const Mat B = p->sampling == MSG_SAMPLING_GIVEN ? load_columns(S.D, columns, n_columns) : direct_factors(S);
        Mat Bh(B.c, B.r);
        for (int i = 0; i < B.r; ++i)
            for (int j = 0; j < B.c; ++j) Bh(j, i) = std::conj(B(i, j));
        const Mat rho0 = matmul(B, Bh);
        // G = -iH - K (SpinSpace_operators.cpp:165-166)
        const Mat G = add(scale(dense_of(build_hamiltonian(S, field)), -I1), dense_of(reaction_operator(S)), -1.0);
        const Mat X = lyapunov_integral(G, rho0);
        // EvaluateDensity (HSObservableCollector.cpp:290-301): rate * Re Tr[P X]
        for (size_t o = 0; o < O.ops.size(); ++o) {
            const Mat P = dense_of(O.ops[o]);
            cd tr = 0.0;
            for (int i = 0; i < P.r; ++i)
                for (int k = 0; k < P.c; ++k) tr += P(i, k) * X(k, i);
            yields[o] = O.scale[o] * tr.real();
        }
    });
}

int ref_lyapunov_integral(int32_t n, const double *g, const double *rho0, double *out) {
    return guarded([&] {
        Mat G(n, n), R(n, n);
        for (size_t i = 0; i < G.a.size(); ++i) {
            G.a[i] = cd(g[2 * i], g[2 * i + 1]);
            R.a[i] = cd(rho0[2 * i], rho0[2 * i + 1]);
        }
        store(lyapunov_integral(G, R), out);
    });
}

int ref_legacy_direct_yields(const msg_system_desc *sys, const double *field, double total_time, double dt,
                             int krylov_size, double krylov_tol, int correction, double *yields) {
    return guarded([&] {
        System S(sys);
        const CSR H = build_hamiltonian(S, field);
        const ObsSet O = observables(S, MSG_OBS_TRANSITION_YIELDS);
        const int nobs = (int)O.ops.size();
        double kmin = 1e300;
        for (double r : O.scale) kmin = std::min(kmin, r);
        for (double r : O.scale)
            if (r != O.scale[0]) throw std::runtime_error("legacy oracle covers the symmetric branch only");
        // columns |psi_e> (x) e_n, unit norm (TaskStaticHSDirectYields.cpp:190-198)
        Mat B = direct_factors(S);
        const int Z = B.c;
        for (auto &x : B.a) x *= std::sqrt((double)Z);
        const int num_steps = (int)std::ceil(total_time / dt);  // endpoint excluded (:347)
        const int m = krylov_size > 0 ? krylov_size : 16;
        const double tol = krylov_tol > 0 ? krylov_tol : 1e-16;
        const int64_t D = S.D;
        std::vector<double> ev((size_t)num_steps * nobs, 0.0), time(num_steps);
        for (int k = 0; k < num_steps; ++k) time[k] = k * dt;
        std::vector<cd> tmp(D);
        for (int col = 0; col < Z; ++col) {
            std::vector<cd> psi(&B.a[(size_t)col * D], &B.a[(size_t)(col + 1) * D]);
            auto observe = [&](int k) {
                for (int o = 0; o < nobs; ++o) {
                    O.ops[o].matvec(psi.data(), tmp.data());
                    cd acc = 0.0;
                    for (int64_t i = 0; i < D; ++i) acc += std::conj(psi[i]) * tmp[i];
                    ev[(size_t)k * nobs + o] += std::exp(-kmin * time[k]) * std::abs(acc);
                }
            };
            // Lanczos with basis reuse (:569-652, LanczosProcess SpinSpace_operators.cpp:1411-1471)
            std::vector<cd> V;
            Mat T;
            double hlast = 0.0;
            std::vector<cd> cx;
            auto lanczos = [&]() {
                double nrm = 0.0;
                for (auto &x : psi) nrm += std::norm(x);
                nrm = std::sqrt(nrm);
                V.assign((size_t)D * m, cd(0.0));
                T = Mat(m, m);
                for (int64_t i = 0; i < D; ++i) V[i] = psi[i] / nrm;
                std::vector<cd> z(D);
                for (int j = 0; j < m; ++j) {
                    H.matvec(&V[(size_t)j * D], z.data());
                    for (int i = std::max(0, j - 1); i <= j; ++i) {
                        cd h = 0.0;
                        const cd *vi = &V[(size_t)i * D];
                        for (int64_t q = 0; q < D; ++q) h += std::conj(vi[q]) * z[q];
                        T(i, j) = h;
                        for (int64_t q = 0; q < D; ++q) z[q] -= h * vi[q];
                    }
                    double zn = 0.0;
                    for (auto &x : z) zn += std::norm(x);
                    zn = std::sqrt(zn);
                    if (j == m - 1) { hlast = zn; break; }
                    T(j + 1, j) = zn;
                    hlast = zn;
                    if (zn < 1e-14) { hlast = 0.0; break; }
                    for (int64_t q = 0; q < D; ++q) V[(size_t)(j + 1) * D + q] = z[q] / zn;
                }
                return nrm;
            };
            auto apply = [&](double nrm) {
                std::fill(psi.begin(), psi.end(), cd(0.0));
                for (int j = 0; j < m; ++j)
                    for (int64_t q = 0; q < D; ++q) psi[q] += nrm * cx[j] * V[(size_t)j * D + q];
            };
            observe(0);
            double nrm = lanczos();
            Mat Ed = expm(scale(T, -I1 * dt));
            cx.assign(m, 0.0);
            for (int j = 0; j < m; ++j) cx[j] = Ed(j, 0);
            apply(nrm);
            for (int k = 1; k < num_steps; ++k) {
                observe(k);
                if (hlast * std::abs(cx[m - 1]) > tol) {
                    nrm = lanczos();
                    Ed = expm(scale(T, -I1 * dt));
                    for (int j = 0; j < m; ++j) cx[j] = Ed(j, 0);
                } else {
                    std::vector<cd> nx(m, 0.0);
                    for (int i = 0; i < m; ++i)
                        for (int j = 0; j < m; ++j) nx[i] += Ed(i, j) * cx[j];
                    cx = nx;
                    double n2 = 0.0;
                    for (auto &x : psi) n2 += std::norm(x);
                    nrm = std::sqrt(n2);
                }
                apply(nrm);
            }
        }
        for (auto &x : ev) x /= Z;
        for (int o = 0; o < nobs; ++o) {
            double acc = 0.0;  // arma::trapz
            for (int k = 1; k < num_steps; ++k)
                acc += 0.5 * (time[k] - time[k - 1]) * (ev[(size_t)(k - 1) * nobs + o] + ev[(size_t)k * nobs + o]);
            acc *= O.scale[o];
            if (correction) acc /= (1.0 - std::exp(-total_time * O.kmax));
            yields[o] = acc;
        }
    });
}

int ref_legacy_stoch_yields(const msg_system_desc *sys, const double *field, double total_time, double dt,
                            const double *columns, int32_t mcols, int correction, double *yields) {
    return guarded([&] {
        System S(sys);
        const CSR H = build_hamiltonian(S, field);
        const ObsSet O = observables(S, MSG_OBS_TRANSITION_YIELDS);
        const int nobs = (int)O.ops.size();
        double kmin = 1e300;
        for (double r : O.scale) kmin = std::min(kmin, r);
        for (double r : O.scale)
            if (r != O.scale[0]) throw std::runtime_error("legacy oracle covers the symmetric branch only");
        Mat B = load_columns(S.D, columns, mcols);
        const Mat U = expm(scale(dense_of(H), -I1 * dt));  // exp(-iH dt) (TaskStaticHSStochYields.cpp:690-694)
        const int num_steps = (int)std::ceil(total_time / dt);
        const int64_t D = S.D;
        std::vector<double> ev((size_t)num_steps * nobs, 0.0), time(num_steps);
        std::vector<cd> tmp(D);
        for (int k = 0; k < num_steps; ++k) {
            time[k] = k * dt;
            for (int o = 0; o < nobs; ++o) {
                cd acc = 0.0;
                for (int c = 0; c < mcols; ++c) {
                    O.ops[o].matvec(&B.a[(size_t)c * D], tmp.data());
                    for (int64_t i = 0; i < D; ++i) acc += std::conj(B.a[(size_t)c * D + i]) * tmp[i];
                }
                ev[(size_t)k * nobs + o] = std::exp(-kmin * time[k]) * std::abs(acc) / mcols;
            }
            B = matmul(U, B);
        }
        for (int o = 0; o < nobs; ++o) {
            double acc = 0.0;
            for (int k = 1; k < num_steps; ++k)
                acc += 0.5 * (time[k] - time[k - 1]) * (ev[(size_t)(k - 1) * nobs + o] + ev[(size_t)k * nobs + o]);
            acc *= O.scale[o];
            if (correction) acc /= (1.0 - std::exp(-total_time * O.kmax));
            yields[o] = acc;
        }
    });
}

int ref_coherent_samples(const msg_system_desc *sys, int state, uint64_t seed, int32_t m, double *out) {
    return guarded([&] {
        System S(sys);
        const CSR P = state_projector(S, state);
        std::vector<cd> cand(S.D), proj(S.D);
        for (int j = 0; j < m; ++j) {
            // Per spin: theta = acos(1-2u), phi = 2 pi u (SpinSpace_states.cpp:535-547), u from
            // Philox at counter (sample, spin, 0, 0x434f4852) under key seed.
            std::vector<cd> vec{1.0};
            for (int s = 0; s < S.nspins; ++s) {
                const uint32_t ctr[4] = {(uint32_t)j, (uint32_t)s, 0u, 0x434f4852u};
                const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
                uint32_t o[4];
                philox(ctr, key, o);
                const double u1 = ((double)(((uint64_t)o[0] << 21) ^ (o[1] >> 11)) + 0.5) * std::ldexp(1.0, -53);
                const double u2 = ((double)(((uint64_t)o[2] << 21) ^ (o[3] >> 11)) + 0.5) * std::ldexp(1.0, -53);
                const double theta = std::acos(1.0 - 2.0 * u1), phi = 2.0 * M_PI * u2;
                const int mult = S.mult[s];
                // exp(-i phi Sz) exp(-i theta Sy) |+S>
                Mat Sy = spin_matrix(mult, 1), Sz = spin_matrix(mult, 2);
                Mat Ry = expm(scale(Sy, -I1 * theta)), Rz = expm(scale(Sz, -I1 * phi));
                std::vector<cd> loc(mult);
                for (int i = 0; i < mult; ++i) loc[i] = Ry(i, 0);
                std::vector<cd> loc2(mult, 0.0);
                for (int i = 0; i < mult; ++i)
                    for (int k = 0; k < mult; ++k) loc2[i] += Rz(i, k) * loc[k];
                std::vector<cd> nv(vec.size() * mult);
                for (size_t a = 0; a < vec.size(); ++a)
                    for (int b = 0; b < mult; ++b) nv[a * mult + b] = vec[a] * loc2[b];
                vec.swap(nv);
            }
            P.matvec(vec.data(), proj.data());
            double nn = 0.0;
            for (auto &x : proj) nn += std::norm(x);
            nn = std::sqrt(nn);
            if (!(nn > 64 * 2.220446049250313e-16)) throw std::runtime_error("zero projected coherent sample");
            for (int64_t i = 0; i < S.D; ++i) {
                out[2 * ((size_t)j * S.D + i)] = proj[i].real() / nn;
                out[2 * ((size_t)j * S.D + i) + 1] = proj[i].imag() / nn;
            }
        }
    });
}

int ref_suz_samples(const msg_system_desc *sys, int state, uint64_t seed, int32_t m, double *out) {
    return guarded([&] {
        System S(sys);
        const msg_state_desc &st = sys->states[state];
        const int32_t *sp = sys->index_pool + st.spins_offset;
        const int ns = st.spins_count;
        std::vector<char> fixed(S.nspins, 0);
        int64_t L = 1;
        for (int k = 0; k < ns; ++k) { fixed[sp[k]] = 1; L *= S.mult[sp[k]]; }
        std::vector<cd> v(L);
        double vn = 0.0;
        for (int64_t i = 0; i < L; ++i) {
            v[i] = cd(sys->amp_pool[2 * (st.amp_offset + i)], sys->amp_pool[2 * (st.amp_offset + i) + 1]);
            vn += std::norm(v[i]);
        }
        std::vector<int> open;
        for (int s = 0; s < S.nspins; ++s) if (!fixed[s]) open.push_back(s);
        const int64_t Z = S.D / L;
        std::vector<cd> chi(Z);
        for (int j = 0; j < m; ++j) {
            double cn = 0.0;
            for (int64_t k = 0; k < Z; ++k) {
                const uint32_t ctr[4] = {(uint32_t)j, (uint32_t)k, 1u, 0x53555A00u};
                const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
                uint32_t o[4];
                philox(ctr, key, o);
                const double u1 = ((double)(((uint64_t)o[0] << 21) ^ (o[1] >> 11)) + 0.5) * std::ldexp(1.0, -53);
                const double u2 = ((double)(((uint64_t)o[2] << 21) ^ (o[3] >> 11)) + 0.5) * std::ldexp(1.0, -53);
                const double rr = std::sqrt(-2.0 * std::log(u1));
                chi[k] = cd(rr * std::cos(2.0 * M_PI * u2), rr * std::sin(2.0 * M_PI * u2));
                cn += std::norm(chi[k]);
            }
            const double sc = 1.0 / std::sqrt(cn * vn);
            for (int64_t i = 0; i < S.D; ++i) { out[2 * ((size_t)j * S.D + i)] = 0.0; out[2 * ((size_t)j * S.D + i) + 1] = 0.0; }
            for (int64_t k = 0; k < Z; ++k) {
                // open-spin digits of k, first open spin most significant
                int64_t rem = k, base = 0;
                for (int q = (int)open.size() - 1; q >= 0; --q) {
                    const int mm = S.mult[open[q]];
                    base += (rem % mm) * S.stride[open[q]];
                    rem /= mm;
                }
                for (int64_t c = 0; c < L; ++c) {
                    int64_t r2 = c, idx = base;
                    for (int q = ns - 1; q >= 0; --q) {
                        const int mm = S.mult[sp[q]];
                        idx += (r2 % mm) * S.stride[sp[q]];
                        r2 /= mm;
                    }
                    const cd val = v[c] * chi[k] * sc;
                    out[2 * ((size_t)j * S.D + idx)] = val.real();
                    out[2 * ((size_t)j * S.D + idx) + 1] = val.imag();
                }
            }
        }
    });
}

void ref_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }

void ref_fibonacci_point(int32_t n, int32_t index, double magnitude, double out[3]) {
    // CalculateFibPoints + RetrievePoint: y and theta pass through float.
    const double phi = M_PI * (3.0 - std::sqrt(5.0));
    const double yd0 = 1.0 - ((double)index / (double)(n - 1)) * 2;
    const double th0 = phi * (double)index;
    const float y = (float)yd0, theta = (float)th0;
    const double r = std::sqrt(1.0 - (y * y));
    out[0] = magnitude * (std::cos(theta) * r);
    out[1] = magnitude * (double)y;
    out[2] = magnitude * (std::sin(theta) * r);
}

void ref_logspace(double a, double b, int32_t n, double *out) {
    for (int i = 0; i < n; ++i) {
        const double e = n == 1 ? b : a + (b - a) * (double)i / (double)(n - 1);
        out[i] = std::pow(10.0, e);
    }
}

double ref_trapezoid(const double *times, const double *values, int64_t n) {
    double acc = 0.0;
    for (int64_t i = 1; i < n; ++i) acc += 0.5 * (times[i] - times[i - 1]) * (values[i - 1] + values[i]);
    return acc;
}

int ref_eig_hermitian(int32_t n, const double *a, double *evals, double *evecs) {
    return guarded([&] {
        Mat A(n, n);
        for (size_t i = 0; i < A.a.size(); ++i) A.a[i] = cd(a[2 * i], a[2 * i + 1]);
        std::vector<double> w;
        Mat V;
        eig_hermitian(A, w, V);
        std::copy(w.begin(), w.end(), evals);
        store(V, evecs);
    });
}

int ref_expm(int32_t n, const double *a, double *out) {
    return guarded([&] {
        Mat A(n, n);
        for (size_t i = 0; i < A.a.size(); ++i) A.a[i] = cd(a[2 * i], a[2 * i + 1]);
        store(expm(A), out);
    });
}

int ref_time_grid(double total_time, double dt, int64_t *n, double *times) {
    return guarded([&] {
        const std::vector<double> t = free_times(total_time, dt);
        *n = (int64_t)t.size();
        if (times) std::copy(t.begin(), t.end(), times);
    });
}

}  // extern "C"
