// Host-side resolution of spin systems into device operator formats.
#include "msg_resolve.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <cstring>

namespace msg {

namespace {

const double kCommon = 87.9410005;  // rad ns^-1 T^-1 (SpinSpace_hamiltonians.cpp:746-748)

// T_a = sum_c g[a][c] S_c (Spin::Tx/Ty/Tz with laboratory quantisation axes).
void t_matrix(const msg_spin_desc &sp, bool ignore, double g[9]) {
    for (int i = 0; i < 9; ++i) g[i] = ignore ? ((i % 4) == 0 ? 1.0 : 0.0) : sp.tensor[i];
}

// Spin-S matrices in the |+S>-first basis, row-major mult x mult, axis x/y/z.
std::vector<cd> spin_op(int mult, int axis) {
    const double S = 0.5 * (mult - 1);
    std::vector<cd> m((size_t)mult * mult, 0.0);
    for (int i = 0; i < mult; ++i) {
        if (axis == 2) { m[(size_t)i * mult + i] = S - i; continue; }
        // <i-1| S+ |i> = sqrt(S(S+1) - m(m+1)), m = S - i
        if (i > 0) {
            const double mm = S - i;
            const double v = std::sqrt(S * (S + 1) - mm * (mm + 1));
            // Sx = (S+ + S-)/2, Sy = (S+ - S-)/(2i)
            if (axis == 0) {
                m[(size_t)(i - 1) * mult + i] += 0.5 * v;
                m[(size_t)i * mult + (i - 1)] += 0.5 * v;
            } else {
                m[(size_t)(i - 1) * mult + i] += cd(0.0, -0.5 * v);
                m[(size_t)i * mult + (i - 1)] += cd(0.0, 0.5 * v);
            }
        }
    }
    return m;
}

// Electron-pair projector of a State (|v><v|/<v|v> on the listed electrons,
// identity on the other electron), or false if the State involves a nucleus.
bool pair_projector(const msg_system_desc *d, const Resolved &r, int state, Mat4 &P) {
    const msg_state_desc &st = d->states[state];
    const int32_t *sp = d->index_pool + st.spins_offset;
    const int ns = st.spins_count;
    if (ns < 1 || ns > 2) return false;
    int which[2] = {-1, -1};
    for (int k = 0; k < ns; ++k) {
        if (sp[k] == r.e1) which[k] = 0;
        else if (sp[k] == r.e2) which[k] = 1;
        else return false;
    }
    if (ns == 2 && which[0] == which[1]) return false;
    const int L = ns == 2 ? 4 : 2;
    std::vector<cd> v(L);
    double nn = 0.0;
    for (int i = 0; i < L; ++i) {
        v[i] = cd(d->amp_pool[2 * (st.amp_offset + i)], d->amp_pool[2 * (st.amp_offset + i) + 1]);
        nn += std::norm(v[i]);
    }
    if (!(nn > 0.0)) return false;
    P.fill(0.0);
    for (int row = 0; row < 4; ++row)
        for (int col = 0; col < 4; ++col) {
            const int ra[2] = {row >> 1, row & 1}, ca[2] = {col >> 1, col & 1};
            if (ns == 2) {
                // listed order (first most significant) -> sub index
                const int rs = ra[which[0]] * 2 + ra[which[1]];
                const int cs = ca[which[0]] * 2 + ca[which[1]];
                P[row * 4 + col] = v[rs] * std::conj(v[cs]) / nn;
            } else {
                const int other = 1 - which[0];
                if (ra[other] != ca[other]) continue;
                P[row * 4 + col] = v[ra[which[0]]] * std::conj(v[ca[which[0]]]) / nn;
            }
        }
    return true;
}

// Jacobi eigen-decomposition of a 4x4 Hermitian matrix (row-major).
void eig4(const Mat4 &Ain, double w[4], cd V[16]) {
    cd A[16];
    for (int i = 0; i < 16; ++i) A[i] = Ain[i];
    for (int i = 0; i < 16; ++i) V[i] = (i % 5 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                if (i != j) off += std::norm(A[i * 4 + j]);
        if (off < 1e-32) break;
        for (int p = 0; p < 3; ++p)
            for (int q = p + 1; q < 4; ++q) {
                const double g = std::abs(A[p * 4 + q]);
                if (g < 1e-300) continue;
                const cd ph = A[p * 4 + q] / g;
                const double th = 0.5 * std::atan2(2.0 * g, A[q * 4 + q].real() - A[p * 4 + p].real());
                const double c = std::cos(th), s = std::sin(th);
                const cd r00 = c, r01 = s, r10 = -s * std::conj(ph), r11 = c * std::conj(ph);
                for (int k = 0; k < 4; ++k) {
                    const cd x = A[k * 4 + p], y = A[k * 4 + q];
                    A[k * 4 + p] = x * r00 + y * r10;
                    A[k * 4 + q] = x * r01 + y * r11;
                }
                for (int k = 0; k < 4; ++k) {
                    const cd x = A[p * 4 + k], y = A[q * 4 + k];
                    A[p * 4 + k] = std::conj(r00) * x + std::conj(r10) * y;
                    A[q * 4 + k] = std::conj(r01) * x + std::conj(r11) * y;
                }
                for (int k = 0; k < 4; ++k) {
                    const cd x = V[k * 4 + p], y = V[k * 4 + q];
                    V[k * 4 + p] = x * r00 + y * r10;
                    V[k * 4 + q] = x * r01 + y * r11;
                }
            }
    }
    for (int i = 0; i < 4; ++i) w[i] = A[i * 5].real();
}

void matmul(int n, const std::vector<cd> &A, const std::vector<cd> &B, std::vector<cd> &C) {
    C.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            const cd a = A[(size_t)i * n + k];
            if (a == 0.0) continue;
            const cd *b = &B[(size_t)k * n];
            cd *c = &C[(size_t)i * n];
            for (int j = 0; j < n; ++j) c[j] += a * b[j];
        }
}

}  // namespace

std::string free_times(double total, double dt, std::vector<double> &t) {
    t.assign(1, 0.0);
    if (!(dt > 0.0) || !std::isfinite(dt)) return "timestep must be finite and positive";
    if (!(total >= 0.0) || !std::isfinite(total)) return "totaltime must be finite and non-negative";
    if (total == 0.0) return "";
    const double ratio = total / dt;
    if (!std::isfinite(ratio) || ratio > 1e9) return "too many time intervals";
    const double nearest = std::round(ratio);
    const double tol = 64.0 * 2.220446049250313e-16 * std::max(1.0, std::fabs(ratio));
    const int64_t n = (int64_t)((nearest >= 1.0 && std::fabs(ratio - nearest) <= tol) ? nearest : std::ceil(ratio));
    for (int64_t s = 1; s <= n; ++s) {
        const double target = s == n ? total : std::min(total, (double)s * dt);
        if (!(target > t.back())) return "non-increasing time grid";
        t.push_back(target);
    }
    return "";
}


namespace {
inline void philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t key, uint32_t o[4]) {
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}
// two uniforms in (0, 1) with 53-bit resolution from one Philox block
inline void two_uniforms(const uint32_t o[4], double &u1, double &u2) {
    u1 = ((double)(((uint64_t)o[0] << 21) ^ (o[1] >> 11)) + 0.5) * 0x1p-53;
    u2 = ((double)(((uint64_t)o[2] << 21) ^ (o[3] >> 11)) + 0.5) * 0x1p-53;
}
}  // namespace

std::string trace_samples(const Resolved &r, int state, int method, uint64_t seed, int m, double *out) {
    if (state < 0 || state >= (int)r.state_spins.size()) return "state index out of range";
    if (m < 1) return "the trace-sample count must be positive";
    const int64_t D = r.D;
    std::vector<int64_t> stride(r.nspins);
    {
        int64_t s = 1;
        for (int i = r.nspins - 1; i >= 0; --i) { stride[i] = s; s *= r.mult[i]; }
    }
    const std::vector<int> &sp = r.state_spins[state];
    const std::vector<cd> &v = r.state_amps[state];
    double vn = 0.0;
    for (const cd &x : v) vn += std::norm(x);
    if (!(vn > 0.0)) return "state with zero norm";
    std::vector<char> fixed(r.nspins, 0);
    for (int x : sp) fixed[x] = 1;
    std::vector<int> open;
    for (int i = 0; i < r.nspins; ++i)
        if (!fixed[i]) open.push_back(i);
    const int64_t L = (int64_t)v.size(), Z = D / L;
    // listed-spin sub-index <-> offset, open-spin index <-> offset
    std::vector<int64_t> off_listed(L), off_open(Z);
    for (int64_t c = 0; c < L; ++c) {
        int64_t rem = c, idx = 0;
        for (int q = (int)sp.size() - 1; q >= 0; --q) {
            idx += (rem % r.mult[sp[q]]) * stride[sp[q]];
            rem /= r.mult[sp[q]];
        }
        off_listed[c] = idx;
    }
    for (int64_t k = 0; k < Z; ++k) {
        int64_t rem = k, idx = 0;
        for (int q = (int)open.size() - 1; q >= 0; --q) {
            idx += (rem % r.mult[open[q]]) * stride[open[q]];
            rem /= r.mult[open[q]];
        }
        off_open[k] = idx;
    }
    std::vector<cd> col(D);
    for (int j = 0; j < m; ++j) {
        if (method == 0) {
            // product over every spin of exp(-i phi Sz) exp(-i theta Sy)|+S>: Wigner d^S_{mS}(theta)
            col.assign(1, cd(1.0));
            for (int s = 0; s < r.nspins; ++s) {
                uint32_t o[4];
                philox4x32((uint32_t)j, (uint32_t)s, 0u, 0x434f4852u, seed, o);
                double u1, u2;
                two_uniforms(o, u1, u2);
                const double th = std::acos(1.0 - 2.0 * u1), ph = 2.0 * M_PI * u2;
                const int mult = r.mult[s];
                const double S = 0.5 * (mult - 1);
                const double c = std::cos(0.5 * th), sn = std::sin(0.5 * th);
                std::vector<cd> loc(mult);
                for (int i = 0; i < mult; ++i) {
                    const double mz = S - i;
                    // sqrt(binomial(2S, S + m)) cos^{S+m} sin^{S-m}
                    const double lb = std::lgamma(2 * S + 1) - std::lgamma(S + mz + 1) - std::lgamma(S - mz + 1);
                    const double mag = std::exp(0.5 * lb) * std::pow(c, S + mz) * std::pow(sn, S - mz);
                    loc[i] = mag * std::polar(1.0, -ph * mz);
                }
                std::vector<cd> nxt(col.size() * mult);
                for (size_t a = 0; a < col.size(); ++a)
                    for (int b = 0; b < mult; ++b) nxt[a * mult + b] = col[a] * loc[b];
                col.swap(nxt);
            }
            // project on the State's support: (v v^dagger / |v|^2) (x) 1
            double nrm = 0.0;
            for (int64_t k = 0; k < Z; ++k) {
                cd ov = 0.0;
                for (int64_t c = 0; c < L; ++c) ov += std::conj(v[c]) * col[off_open[k] + off_listed[c]];
                ov /= vn;
                for (int64_t c = 0; c < L; ++c) {
                    col[off_open[k] + off_listed[c]] = ov * v[c];
                    nrm += std::norm(ov * v[c]);
                }
            }
            if (!(nrm > 1e-28)) return "a coherent sample has no overlap with the State";
            const double inv = 1.0 / std::sqrt(nrm);
            for (int64_t i = 0; i < D; ++i) col[i] *= inv;
        } else {
            // |v> (x) normalise(x + i y), x, y ~ N(0,1) by Box-Muller
            std::vector<cd> chi(Z);
            double cn = 0.0;
            for (int64_t k = 0; k < Z; ++k) {
                uint32_t o[4];
                philox4x32((uint32_t)j, (uint32_t)k, 1u, 0x53555A00u, seed, o);
                double u1, u2;
                two_uniforms(o, u1, u2);
                const double rr = std::sqrt(-2.0 * std::log(u1));
                chi[k] = cd(rr * std::cos(2.0 * M_PI * u2), rr * std::sin(2.0 * M_PI * u2));
                cn += std::norm(chi[k]);
            }
            const double sc = 1.0 / std::sqrt(cn * vn);
            std::fill(col.begin(), col.end(), cd(0.0));
            col.resize(D);
            for (int64_t k = 0; k < Z; ++k)
                for (int64_t c = 0; c < L; ++c) col[off_open[k] + off_listed[c]] = v[c] * chi[k] * sc;
        }
        for (int64_t i = 0; i < D; ++i) {
            out[2 * ((size_t)j * D + i)] = col[i].real();
            out[2 * ((size_t)j * D + i) + 1] = col[i].imag();
        }
    }
    return "";
}

std::string legacy_times(double total, double dt, std::vector<double> &t) {
    t.clear();
    if (!(dt > 0.0) || !std::isfinite(dt)) return "timestep must be finite and positive";
    if (!(total > 0.0) || !std::isfinite(total)) return "totaltime must be finite and positive";
    const double n = std::ceil(total / dt);
    if (!(n >= 1.0) || n > 1e9) return "too many time intervals";
    for (int64_t k = 0; k < (int64_t)n; ++k) t.push_back((double)k * dt);
    return "";
}

std::string resolve_system(const msg_system_desc *d, Resolved &r) {
    r = Resolved();
    if (!d || d->n_spins < 1 || !d->spins) return "system has no spins";
    r.nspins = d->n_spins;
    r.mult.resize(r.nspins);
    r.D = 1;
    for (int i = 0; i < r.nspins; ++i) {
        r.mult[i] = d->spins[i].multiplicity;
        if (r.mult[i] < 1 || r.mult[i] > 16) return "invalid spin multiplicity";
        r.D *= r.mult[i];
        if (r.D > (int64_t)1 << 30) return "Hilbert space too large";
    }
    auto pool_ok = [&](int off, int cnt) {
        return off >= 0 && cnt >= 0 && off + cnt <= d->index_pool_len;
    };
    // electrons
    std::vector<int> el;
    for (int i = 0; i < r.nspins; ++i)
        if (d->spins[i].is_electron) el.push_back(i);
    if (el.size() == 2 && r.mult[el[0]] == 2 && r.mult[el[1]] == 2) {
        r.e1 = el[0];
        r.e2 = el[1];
        r.pair_ok = true;
    }
    for (int st = 0; st < d->n_states; ++st) {
        const msg_state_desc &sd = d->states[st];
        if (!pool_ok(sd.spins_offset, sd.spins_count) || sd.amp_offset < 0) return "state out of range";
        std::vector<int> sp(d->index_pool + sd.spins_offset, d->index_pool + sd.spins_offset + sd.spins_count);
        int64_t L = 1;
        for (int x : sp) {
            if (x < 0 || x >= r.nspins) return "state spin index out of range";
            L *= r.mult[x];
        }
        if (sd.amp_offset + L > d->amp_pool_len) return "state amplitudes out of range";
        std::vector<cd> amps(L);
        for (int64_t i = 0; i < L; ++i)
            amps[i] = cd(d->amp_pool[2 * (sd.amp_offset + i)], d->amp_pool[2 * (sd.amp_offset + i) + 1]);
        r.state_spins.push_back(sp);
        r.state_amps.push_back(amps);
    }
    if (!r.pair_ok) return "";  // structure not representable on the device paths; caller reports
    // nuclei in device digit order: the spin-1/2 nuclei first (system order), then the
    // others, so that a single nucleus of higher spin is the most significant digit
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < r.nspins; ++i)
            if (i != r.e1 && i != r.e2 && ((r.mult[i] == 2) == (pass == 0))) r.nuc.push_back(i);
    r.zdim = r.D / 4;
    // MolSpin index -> device index (nu * 4 + 2a + b, nucleus 0 least significant)
    r.perm.resize(r.D);
    std::vector<int64_t> stride(r.nspins);
    {
        int64_t s = 1;
        for (int i = r.nspins - 1; i >= 0; --i) { stride[i] = s; s *= r.mult[i]; }
    }
    std::vector<int64_t> nstride(r.nuc.size());
    {
        int64_t s = 1;
        for (size_t k = 0; k < r.nuc.size(); ++k) { nstride[k] = s; s *= r.mult[r.nuc[k]]; }
    }
    for (int64_t idx = 0; idx < r.D; ++idx) {
        const int a = (int)((idx / stride[r.e1]) % 2), b = (int)((idx / stride[r.e2]) % 2);
        int64_t nu = 0;
        for (size_t k = 0; k < r.nuc.size(); ++k)
            nu += ((idx / stride[r.nuc[k]]) % r.mult[r.nuc[k]]) * nstride[k];
        r.perm[idx] = nu * 4 + 2 * a + b;
    }
    // interactions -> one- and two-spin terms (InteractionOperator)
    for (int it = 0; it < d->n_interactions; ++it) {
        const msg_interaction_desc &in = d->interactions[it];
        if (!pool_ok(in.group1_offset, in.group1_count) || !pool_ok(in.group2_offset, in.group2_count))
            return "interaction spin group out of range";
        double sc = in.prefactor;
        if (in.tensor_present && in.tensor_isotropic) sc *= in.tensor[0];
        if (in.common_prefactor) sc *= kCommon;
        const int32_t *g1 = d->index_pool + in.group1_offset;
        const int32_t *g2 = d->index_pool + in.group2_offset;
        for (int k = 0; k < in.group1_count; ++k)
            if (g1[k] < 0 || g1[k] >= r.nspins) return "interaction spin index out of range";
        for (int k = 0; k < in.group2_count; ++k)
            if (g2[k] < 0 || g2[k] >= r.nspins) return "interaction spin index out of range";
        if (in.type == MSG_INTERACTION_ZEEMAN) {
            for (int k = 0; k < in.group1_count; ++k) {
                Term1 t{};
                t.s = g1[k];
                t.swept = in.field_swept != 0;
                double G[9];
                t_matrix(d->spins[t.s], in.ignore_tensors != 0, G);
                for (int i = 0; i < 9; ++i) t.Mz[i] = sc * G[i];
                for (int a = 0; a < 3; ++a)
                    t.c[a] = t.Mz[3 * a] * in.field[0] + t.Mz[3 * a + 1] * in.field[1] + t.Mz[3 * a + 2] * in.field[2];
                if (in.time_dependence != MSG_TD_NONE) {
                    if (in.time_dependence != MSG_TD_LINEAR && in.time_dependence != MSG_TD_CIRCULAR)
                        return "unknown field time dependence (linearpolarized or circularpolarized)";
                    if (in.field_swept) return "a time-dependent Zeeman field cannot be the swept field";
                    t.td = in.time_dependence;
                    t.td_perp = in.td_perpendicular != 0;
                    t.td_freq = in.td_frequency;
                    t.td_phase = in.td_phase;
                    for (int a = 0; a < 3; ++a) {
                        t.td_axis[a] = in.td_axis[a];
                        t.field[a] = in.field[a];
                    }
                    r.has_td = true;
                }
                r.t1.push_back(t);
            }
        } else if (in.type == MSG_INTERACTION_DOUBLESPIN || in.type == MSG_INTERACTION_EXCHANGE) {
            const bool ex = in.type == MSG_INTERACTION_EXCHANGE;
            const bool aniso = in.tensor_present && !in.tensor_isotropic;
            for (int i = 0; i < in.group1_count; ++i)
                for (int j = 0; j < in.group2_count; ++j) {
                    Term2 t{};
                    t.s1 = g1[i];
                    t.s2 = g2[j];
                    if (t.s1 == t.s2) return "a two-spin interaction couples a spin to itself";
                    double ga[9], gb[9];
                    t_matrix(d->spins[t.s1], in.ignore_tensors != 0, ga);
                    t_matrix(d->spins[t.s2], in.ignore_tensors != 0, gb);
                    // S1_c S2_e coefficient of sum_ab A_ab T1a T2b
                    for (int c = 0; c < 3; ++c)
                        for (int e = 0; e < 3; ++e) {
                            double m = 0.0;
                            for (int aa = 0; aa < 3; ++aa)
                                for (int bb = 0; bb < 3; ++bb)
                                    m += (aniso ? in.tensor[3 * aa + bb] : (aa == bb ? 1.0 : 0.0)) * ga[3 * aa + c] *
                                         gb[3 * bb + e];
                            t.M[3 * c + e] = sc * (ex ? 2.0 : 1.0) * m;
                        }
                    r.t2.push_back(t);
                    if (ex) r.shift += 0.5 * sc;
                }
        } else {
            return "interaction type not supported";
        }
    }
    // structure of the matrix-free operator
    r.fast_ok = true;
    r.mix_mult = 0;
    {
        int nhalf = 0, nbig = 0;
        for (int k : r.nuc) {
            if (r.mult[k] == 2) ++nhalf;
            else { ++nbig; r.mix_mult = r.mult[k]; }
        }
        // one nucleus of spin 1 or 3/2 rides on the group index of the matrix-free kernel,
        // beside 6 to 10 spin-1/2 nuclei on the thread index
        if (nbig > 1) { r.fast_ok = false; r.why_not_fast = "more than one nucleus with spin > 1/2"; }
        else if (nbig == 1 && r.mix_mult > 4) { r.fast_ok = false; r.why_not_fast = "nucleus with spin > 3/2"; }
        else if (nbig == 1 && (nhalf < 6 || nhalf > 10)) {
            r.fast_ok = false;
            r.why_not_fast = "a nucleus with spin > 1/2 needs 6 to 10 spin-1/2 nuclei beside it";
        }
    }
    if ((int)r.nuc.size() > kMaxNuclei) { r.fast_ok = false; r.why_not_fast = "too many nuclei"; }
    r.nn_pairs.clear();
    {
        std::vector<int> nucpos(r.nspins, -1);
        for (size_t k = 0; k < r.nuc.size(); ++k) nucpos[r.nuc[k]] = (int)k;
        for (const Term2 &t : r.t2) {
            const bool e_a = t.s1 == r.e1 || t.s1 == r.e2, e_b = t.s2 == r.e1 || t.s2 == r.e2;
            if (e_a && e_b) r.has_ee = true;
            if (e_a || e_b) continue;
            // nucleus-nucleus coupling: a 4x4 block on the two nuclei's index bits
            if (r.mult[t.s1] != 2 || r.mult[t.s2] != 2) {
                r.fast_ok = false;
                r.why_not_fast = "nucleus-nucleus coupling with a nucleus of spin > 1/2";
                continue;
            }
            const int k1 = std::min(nucpos[t.s1], nucpos[t.s2]), k2 = std::max(nucpos[t.s1], nucpos[t.s2]);
            const std::array<int, 2> pr{k1, k2};
            if (std::find(r.nn_pairs.begin(), r.nn_pairs.end(), pr) == r.nn_pairs.end()) r.nn_pairs.push_back(pr);
        }
        if ((int)r.nn_pairs.size() > kMaxPairs) {
            r.fast_ok = false;
            r.why_not_fast = "more than " + std::to_string(kMaxPairs) + " nucleus-nucleus couplings";
        }
    }
    // states, transitions, reaction operator
    for (int s = 0; s < d->n_states; ++s) {
        if (!pool_ok(d->states[s].spins_offset, d->states[s].spins_count)) return "state spin list out of range";
        Mat4 P;
        if (!pair_projector(d, r, s, P)) {
            r.pair_ok = false;
            return "";
        }
        r.P_states.push_back(P);
    }
    r.K.fill(0.0);
    for (int t = 0; t < d->n_transitions; ++t) {
        const int src = d->transitions[t].source_state;
        if (src < 0) continue;
        if (src >= d->n_states) return "transition source out of range";
        r.P_trans.push_back(r.P_states[src]);
        r.rates.push_back(d->transitions[t].rate);
        r.kmax = std::max(r.kmax, d->transitions[t].rate);
        for (int i = 0; i < 16; ++i) r.K[i] += 0.5 * d->transitions[t].rate * r.P_states[src][i];
    }
    {
        const cd kap = 0.25 * (r.K[0] + r.K[5] + r.K[10] + r.K[15]);
        double dev = 0.0;
        for (int i = 0; i < 16; ++i) dev += std::abs(r.K[i] - ((i % 5) == 0 ? kap : cd(0.0)));
        r.symmetric = dev <= 1e-13 * std::max(1.0, std::abs(kap)) && std::fabs(kap.imag()) < 1e-300 + 1e-13;
        r.kappa = kap.real();
    }
    // relaxation: relaxationdephasing channels on the radical-pair electrons; a zero
    // rate is no channel (HasNonzeroRelaxationRate, SpinSpace_relaxation.cpp:1110-1125)
    r.relax_rates.clear();
    r.relax_rate = 0.0;
    for (int i = 0; i < d->n_relaxations; ++i) {
        const auto &rl = d->relaxations[i];
        if (rl.rate == 0.0) continue;
        if (!((rl.spin1 == r.e1 && rl.spin2 == r.e2) || (rl.spin1 == r.e2 && rl.spin2 == r.e1)))
            return "relaxationdephasing must act on the radical-pair electrons";
        if (!(rl.rate > 0.0) || !std::isfinite(rl.rate)) return "relaxationdephasing rate must be positive and finite";
        r.relax_rates.push_back(rl.rate);
        r.relax_rate += rl.rate;
    }
    r.relax = !r.relax_rates.empty();
    if (d->n_transitions > 0) r.first_rate = d->transitions[0].rate;
    for (const Term1 &t : r.t1)
        if (t.td && t.s != r.e1 && t.s != r.e2) r.td_nuclear = true;
    // direct factors of rho0 = rho_e (x) 1/Z (HSStatePreparation.cpp:77-165)
    if (d->n_initial >= 1) {
        if (d->n_initial > 16) return "too many initial states";
        std::vector<double> w(d->n_initial, 1.0 / d->n_initial);
        if (d->initial_weights) {
            double tot = 0.0;
            for (int i = 0; i < d->n_initial; ++i) {
                w[i] = d->initial_weights[i];
                if (!(w[i] >= 0.0)) return "initial-state weights must be non-negative";
                tot += w[i];
            }
            if (!(tot > 0.0)) return "initial-state weights must have a positive sum";
            for (double &x : w) x /= tot;
        }
        Mat4 rho{};
        for (int i = 0; i < d->n_initial; ++i) {
            const int s = d->initial_states[i];
            if (s < 0 || s >= d->n_states) return "initial state out of range";
            r.initial.push_back(s);
            const Mat4 &P = r.P_states[s];
            const double tr = (P[0] + P[5] + P[10] + P[15]).real();
            for (int k = 0; k < 16; ++k) rho[k] += w[i] * P[k] / tr;
        }
        double ev[4];
        cd V[16];
        eig4(rho, ev, V);
        for (int k = 0; k < 4; ++k)
            if (ev[k] > 1e-12) {
                std::array<cd, 4> v;
                for (int i = 0; i < 4; ++i) v[i] = std::sqrt(ev[k]) * V[i * 4 + k];
                r.evec.push_back(v);
            }
    }
    return "";
}

void td_field(const Term1 &t, double time, double B[3]) {
    if (t.td == MSG_TD_LINEAR) {  // linearpolarized: the field scaled by cos(w t + phase)
        const double c = std::cos(t.td_freq * time + t.td_phase);
        for (int a = 0; a < 3; ++a) B[a] = t.field[a] * c;
        return;
    }
    // circularpolarized: Rodrigues rotation by x = w t + phase about the axis as given
    const double *u = t.td_axis;
    const double x = time * t.td_freq + t.td_phase, sx = std::sin(x), hx = 2.0 * std::sin(0.5 * x) * std::sin(0.5 * x);
    const double W[3][3] = {{0.0, -u[2], u[1]}, {u[2], 0.0, -u[0]}, {-u[1], u[0], 0.0}};
    double b[3] = {t.field[0], t.field[1], t.field[2]};
    if (t.td_perp) {
        const double d = u[0] * b[0] + u[1] * b[1] + u[2] * b[2];
        for (int a = 0; a < 3; ++a) b[a] -= d * u[a];
    }
    for (int i = 0; i < 3; ++i) {
        double acc = b[i];
        for (int j = 0; j < 3; ++j) {
            double w2 = 0.0;
            for (int k = 0; k < 3; ++k) w2 += W[i][k] * W[k][j];
            acc += (sx * W[i][j] + hx * w2) * b[j];
        }
        B[i] = acc;
    }
}

bool term1_value(const Term1 &t, const double *field, const TimePoint *tp, double v[3]) {
    if (t.td && tp) {
        if (tp->skip) return false;
        double B[3];
        td_field(t, tp->t, B);
        for (int a = 0; a < 3; ++a) v[a] = t.Mz[3 * a] * B[0] + t.Mz[3 * a + 1] * B[1] + t.Mz[3 * a + 2] * B[2];
        return true;
    }
    for (int a = 0; a < 3; ++a)
        v[a] = (t.swept && field) ? t.Mz[3 * a] * field[0] + t.Mz[3 * a + 1] * field[1] + t.Mz[3 * a + 2] * field[2]
                                  : t.c[a];
    return true;
}

void fast_coefficients(const Resolved &r, const double *field, std::vector<float> &out, int &op_flags,
                       int nuc_flags[kMaxNuclei], const TimePoint *tp) {
    const int N = (int)r.nuc.size(), P = (int)r.nn_pairs.size();
    std::vector<double> c((size_t)kCoefHead + (size_t)N * kNucStride + (size_t)P * kPairStride, 0.0);
    std::vector<int> nucpos(r.nspins, -1);
    for (int k = 0; k < N; ++k) nucpos[r.nuc[k]] = k;
    for (int k = 0; k < kMaxNuclei; ++k) nuc_flags[k] = 0;
    op_flags = 0;
    c[6] = r.shift;
    // electron-pair matrix E (e-e couplings); S1 acts on the high bit of c
    std::vector<cd> sx = spin_op(2, 0), sy = spin_op(2, 1), sz = spin_op(2, 2);
    const std::vector<cd> *S[3] = {&sx, &sy, &sz};
    cd E[16] = {};
    for (const Term1 &t : r.t1) {
        double v[3];
        if (!term1_value(t, field, tp, v)) continue;
        if (t.s == r.e1) for (int a = 0; a < 3; ++a) c[a] += v[a];
        else if (t.s == r.e2) for (int a = 0; a < 3; ++a) c[3 + a] += v[a];
        else {
            const int k = nucpos[t.s];
            double *cn = &c[kCoefHead + (size_t)k * kNucStride];
            cn[6] += 0.5 * v[2];            // <up|c.I|up> = c_z/2
            cn[23] += 0.5 * v[0];           // <up|c.I|down> = (c_x - i c_y)/2
            cn[24] += -0.5 * v[1];
            nuc_flags[k] |= NUC_T;
        }
    }
    for (const Term2 &t : r.t2) {
        const bool e1a = t.s1 == r.e1 || t.s1 == r.e2, e2b = t.s2 == r.e1 || t.s2 == r.e2;
        if (e1a && e2b) {
            // sum_ab M_ab S_{s1,a} S_{s2,b} on the pair, kron(e1 op, e2 op)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double m = t.M[3 * a + b];
                    if (m == 0.0) continue;
                    const std::vector<cd> &A1 = t.s1 == r.e1 ? *S[a] : *S[b];
                    const std::vector<cd> &A2 = t.s1 == r.e1 ? *S[b] : *S[a];
                    for (int i = 0; i < 4; ++i)
                        for (int j = 0; j < 4; ++j)
                            E[i * 4 + j] += m * A1[(i >> 1) * 2 + (j >> 1)] * A2[(i & 1) * 2 + (j & 1)];
                }
            op_flags |= OP_EREST;
            continue;
        }
        if (!e1a && !e2b) {
            // nucleus-nucleus: sum_ab M_ab I_{k1,a} I_{k2,b} as kron(k1 op, k2 op)
            const int k1 = nucpos[t.s1], k2 = nucpos[t.s2];
            const std::array<int, 2> pr{std::min(k1, k2), std::max(k1, k2)};
            const int pc = (int)(std::find(r.nn_pairs.begin(), r.nn_pairs.end(), pr) - r.nn_pairs.begin());
            double *ch = &c[kCoefHead + (size_t)N * kNucStride + (size_t)pc * kPairStride];
            const bool first = k1 == pr[0];  // t.s1 is the pair's first nucleus
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double m = t.M[3 * a + b];
                    if (m == 0.0) continue;
                    const std::vector<cd> &A1 = first ? *S[a] : *S[b];
                    const std::vector<cd> &A2 = first ? *S[b] : *S[a];
                    for (int i = 0; i < 4; ++i)
                        for (int j = 0; j < 4; ++j) {
                            const cd v = m * A1[(i >> 1) * 2 + (j >> 1)] * A2[(i & 1) * 2 + (j & 1)];
                            ch[8 * i + 2 * j] += v.real();
                            ch[8 * i + 2 * j + 1] += v.imag();
                        }
                }
            continue;
        }
        // electron-nucleus: M_ab with a on the electron, b on the nucleus
        int e, n;
        double M[9];
        if (e1a) { e = t.s1; n = t.s2; for (int i = 0; i < 9; ++i) M[i] = t.M[i]; }
        else { e = t.s2; n = t.s1; for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) M[3 * a + b] = t.M[3 * b + a]; }
        const int k = nucpos[n];
        double *cn = &c[kCoefHead + (size_t)k * kNucStride];
        const int hoff = e == r.e1 ? 0 : 3, boff = e == r.e1 ? 7 : 15;
        nuc_flags[k] |= e == r.e1 ? NUC_E1 : NUC_E2;
        for (int a = 0; a < 3; ++a) cn[hoff + a] += 0.5 * M[3 * a + 2];
        // u_a = (M_ax - i M_ay)/2;  block = sum_a u_a sigma_a / 2
        cd u[3];
        for (int a = 0; a < 3; ++a) u[a] = cd(0.5 * M[3 * a + 0], -0.5 * M[3 * a + 1]);
        const cd I(0.0, 1.0);
        const cd b00 = 0.5 * u[2], b01 = 0.5 * (u[0] - I * u[1]), b10 = 0.5 * (u[0] + I * u[1]), b11 = -0.5 * u[2];
        const cd blk[4] = {b00, b01, b10, b11};
        for (int q = 0; q < 4; ++q) {
            cn[boff + 2 * q] += blk[q].real();
            cn[boff + 2 * q + 1] += blk[q].imag();
        }
    }
    for (int i = 0; i < 16; ++i) {
        c[8 + 2 * i] = E[i].real();
        c[8 + 2 * i + 1] = E[i].imag();
    }
    // axial couplings: the block C = [[0, 0], [b, 0]] with b real and h along z; entries
    // below 1e-12 of the block's largest (the rounding of a rotated isotropic tensor) are
    // cleared, far below the FP32 rounding of the kernels' sums
    for (int k = 0; k < N; ++k) {
        double *cn = &c[kCoefHead + (size_t)k * kNucStride];
        for (int e = 0; e < 2; ++e) {
            if (!(nuc_flags[k] & (e ? NUC_E2 : NUC_E1))) continue;
            double *h = cn + 3 * e, *b = cn + (e ? 15 : 7);
            double big = 0.0;
            for (int q = 0; q < 3; ++q) big = std::max(big, std::abs(h[q]));
            for (int q = 0; q < 8; ++q) big = std::max(big, std::abs(b[q]));
            const double eps = 1e-12 * big;
            const int zero[] = {0, 1, 2, 3, 5, 6, 7};  // C[0], C[1], Im C[2], C[3]
            bool axial = std::abs(h[0]) <= eps && std::abs(h[1]) <= eps;
            for (int q : zero) axial = axial && std::abs(b[q]) <= eps;
            if (!axial) continue;
            h[0] = h[1] = 0.0;
            for (int q : zero) b[q] = 0.0;
            nuc_flags[k] |= e ? NUC_A2 : NUC_A1;
        }
    }
    out.resize(c.size());
    for (size_t i = 0; i < c.size(); ++i) out[i] = (float)c[i];
}

void spin_operator_dense(const Resolved &r, int spin, int axis, std::vector<cd> &H) {
    const int64_t D = r.D;
    H.assign((size_t)(D * D), 0.0);
    int64_t stride = 1;
    for (int i = r.nspins - 1; i > spin; --i) stride *= r.mult[i];
    const int m = r.mult[spin];
    const std::vector<cd> S = spin_op(m, axis);
    for (int64_t row = 0; row < D; ++row) {
        const int dr = (int)((row / stride) % m);
        for (int dc = 0; dc < m; ++dc) {
            const cd x = S[(size_t)dr * m + dc];
            if (x != 0.0) H[(size_t)(r.perm[row] * D + r.perm[row + (int64_t)(dc - dr) * stride])] += x;
        }
    }
}

void dense_hamiltonian(const Resolved &r, const double *field, std::vector<cd> &H, const TimePoint *tp) {
    const int64_t D = r.D;
    H.assign((size_t)(D * D), 0.0);
    std::vector<int64_t> stride(r.nspins);
    {
        int64_t s = 1;
        for (int i = r.nspins - 1; i >= 0; --i) { stride[i] = s; s *= r.mult[i]; }
    }
    // cached per (mult, axis); sized once so references stay valid
    std::vector<std::vector<cd>> ops((size_t)17 * 3);
    auto op = [&](int mult, int axis) -> const std::vector<cd> & {
        const size_t key = (size_t)mult * 3 + axis;
        if (ops[key].empty()) ops[key] = spin_op(mult, axis);
        return ops[key];
    };
    auto digit = [&](int64_t idx, int s) { return (int)((idx / stride[s]) % r.mult[s]); };
    // in MolSpin order, then permuted
    auto add = [&](int64_t row, int64_t col, cd v) { H[(size_t)(r.perm[row] * D + r.perm[col])] += v; };
    for (int64_t row = 0; row < D; ++row) {
        add(row, row, r.shift);
        for (const Term1 &t : r.t1) {
            double v[3];
            if (!term1_value(t, field, tp, v)) continue;
            const int m = r.mult[t.s], dr = digit(row, t.s);
            for (int a = 0; a < 3; ++a) {
                if (v[a] == 0.0) continue;
                const std::vector<cd> &S = op(m, a);
                for (int dc = 0; dc < m; ++dc) {
                    const cd x = S[(size_t)dr * m + dc];
                    if (x != 0.0) add(row, row + (int64_t)(dc - dr) * stride[t.s], v[a] * x);
                }
            }
        }
        for (const Term2 &t : r.t2) {
            const int m1 = r.mult[t.s1], m2 = r.mult[t.s2];
            const int d1 = digit(row, t.s1), d2 = digit(row, t.s2);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double mab = t.M[3 * a + b];
                    if (mab == 0.0) continue;
                    const std::vector<cd> &A = op(m1, a), &B = op(m2, b);
                    for (int c1 = 0; c1 < m1; ++c1) {
                        const cd x = A[(size_t)d1 * m1 + c1];
                        if (x == 0.0) continue;
                        for (int c2 = 0; c2 < m2; ++c2) {
                            const cd y = B[(size_t)d2 * m2 + c2];
                            if (y == 0.0) continue;
                            add(row, row + (int64_t)(c1 - d1) * stride[t.s1] + (int64_t)(c2 - d2) * stride[t.s2],
                                mab * x * y);
                        }
                    }
                }
        }
    }
}

std::string uncoupled_split(const Resolved &r, std::vector<int> &A, std::vector<int> &B) {
    if (!r.pair_ok) return "needs two spin-1/2 electrons with States on them";
    if (r.has_ee) return "the electrons are coupled (exchange or electron-electron terms)";
    // connected components of the two-spin coupling graph
    std::vector<int> comp(r.nspins);
    for (int i = 0; i < r.nspins; ++i) comp[i] = i;
    std::function<int(int)> find = [&](int x) { return comp[x] == x ? x : comp[x] = find(comp[x]); };
    for (const Term2 &t : r.t2) comp[find(t.s1)] = find(t.s2);
    const int ca = find(r.e1), cb = find(r.e2);
    if (ca == cb) return "a chain of couplings links the two electrons";
    A.assign(1, r.e1);
    B.assign(1, r.e2);
    for (int i = 0; i < r.nspins; ++i) {
        if (i == r.e1 || i == r.e2) continue;
        (find(i) == cb ? B : A).push_back(i);
    }
    return "";
}

void subsystem_hamiltonian(const Resolved &r, const std::vector<int> &spins, const double *field,
                           std::vector<cd> &H) {
    const int ns = (int)spins.size();
    std::vector<int> pos(r.nspins, -1);
    for (int i = 0; i < ns; ++i) pos[spins[i]] = i;
    std::vector<int64_t> stride(ns);
    int64_t n = 1;
    for (int i = ns - 1; i >= 0; --i) { stride[i] = n; n *= r.mult[spins[i]]; }
    H.assign((size_t)(n * n), 0.0);
    std::vector<std::vector<cd>> ops((size_t)17 * 3);
    auto op = [&](int mult, int axis) -> const std::vector<cd> & {
        const size_t key = (size_t)mult * 3 + axis;
        if (ops[key].empty()) ops[key] = spin_op(mult, axis);
        return ops[key];
    };
    auto digit = [&](int64_t idx, int i) { return (int)((idx / stride[i]) % r.mult[spins[i]]); };
    for (int64_t row = 0; row < n; ++row) {
        for (const Term1 &t : r.t1) {
            const int i = pos[t.s];
            if (i < 0) continue;
            double v[3];
            term1_value(t, field, nullptr, v);
            const int m = r.mult[t.s], dr = digit(row, i);
            for (int a = 0; a < 3; ++a) {
                if (v[a] == 0.0) continue;
                const std::vector<cd> &S = op(m, a);
                for (int dc = 0; dc < m; ++dc) {
                    const cd x = S[(size_t)dr * m + dc];
                    if (x != 0.0) H[(size_t)(row * n + row + (int64_t)(dc - dr) * stride[i])] += v[a] * x;
                }
            }
        }
        for (const Term2 &t : r.t2) {
            const int i1 = pos[t.s1], i2 = pos[t.s2];
            if (i1 < 0 || i2 < 0) continue;
            const int m1 = r.mult[t.s1], m2 = r.mult[t.s2];
            const int d1 = digit(row, i1), d2 = digit(row, i2);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double mab = t.M[3 * a + b];
                    if (mab == 0.0) continue;
                    const std::vector<cd> &Sa = op(m1, a), &Sb = op(m2, b);
                    for (int c1 = 0; c1 < m1; ++c1) {
                        const cd x = Sa[(size_t)d1 * m1 + c1];
                        if (x == 0.0) continue;
                        for (int c2 = 0; c2 < m2; ++c2) {
                            const cd y = Sb[(size_t)d2 * m2 + c2];
                            if (y == 0.0) continue;
                            H[(size_t)(row * n + row + (int64_t)(c1 - d1) * stride[i1] + (int64_t)(c2 - d2) * stride[i2])] +=
                                mab * x * y;
                        }
                    }
                }
        }
    }
}

void expm_dense(int n, const std::vector<cd> &A, std::vector<cd> &E) {
    // Pade(13) coefficients and theta_13 (Higham, SIAM J. Matrix Anal. Appl. 26 (2005) 1179).
    static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
                                 129060195264000.0,   10559470521600.0,    670442572800.0,     33522128640.0,
                                 1323241920.0,        40840800.0,          960960.0,           16380.0,
                                 182.0,               1.0};
    const double theta = 5.371920351148152;
    double nrm = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += std::abs(A[(size_t)i * n + j]);
        nrm = std::max(nrm, s);
    }
    int sq = 0;
    if (nrm > theta) sq = (int)std::ceil(std::log2(nrm / theta));
    const double f = std::ldexp(1.0, -sq);
    std::vector<cd> X(A.size());
    for (size_t i = 0; i < A.size(); ++i) X[i] = A[i] * f;
    std::vector<cd> X2, X4, X6;
    matmul(n, X, X, X2);
    matmul(n, X2, X2, X4);
    matmul(n, X4, X2, X6);
    const size_t nn = (size_t)n * n;
    std::vector<cd> t1(nn), t2(nn), U(nn), V(nn), tmp;
    for (size_t i = 0; i < nn; ++i) t1[i] = b[13] * X6[i] + b[11] * X4[i] + b[9] * X2[i];
    matmul(n, X6, t1, tmp);
    for (size_t i = 0; i < nn; ++i) t2[i] = tmp[i] + b[7] * X6[i] + b[5] * X4[i] + b[3] * X2[i];
    for (int i = 0; i < n; ++i) t2[(size_t)i * n + i] += b[1];
    matmul(n, X, t2, U);
    for (size_t i = 0; i < nn; ++i) t1[i] = b[12] * X6[i] + b[10] * X4[i] + b[8] * X2[i];
    matmul(n, X6, t1, tmp);
    for (size_t i = 0; i < nn; ++i) V[i] = tmp[i] + b[6] * X6[i] + b[4] * X4[i] + b[2] * X2[i];
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] += b[0];
    // solve (V - U) E = (V + U) by Gaussian elimination with partial pivoting
    std::vector<cd> L(nn), R(nn);
    for (size_t i = 0; i < nn; ++i) { L[i] = V[i] - U[i]; R[i] = V[i] + U[i]; }
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::abs(L[(size_t)i * n + k]) > std::abs(L[(size_t)p * n + k])) p = i;
        if (p != k)
            for (int j = 0; j < n; ++j) {
                std::swap(L[(size_t)k * n + j], L[(size_t)p * n + j]);
                std::swap(R[(size_t)k * n + j], R[(size_t)p * n + j]);
            }
        const cd piv = L[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) {
            const cd fct = L[(size_t)i * n + k] / piv;
            if (fct == 0.0) continue;
            for (int j = k; j < n; ++j) L[(size_t)i * n + j] -= fct * L[(size_t)k * n + j];
            for (int j = 0; j < n; ++j) R[(size_t)i * n + j] -= fct * R[(size_t)k * n + j];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        for (int k = i + 1; k < n; ++k) {
            const cd fct = L[(size_t)i * n + k];
            if (fct == 0.0) continue;
            for (int j = 0; j < n; ++j) R[(size_t)i * n + j] -= fct * R[(size_t)k * n + j];
        }
        const cd inv = 1.0 / L[(size_t)i * n + i];
        for (int j = 0; j < n; ++j) R[(size_t)i * n + j] *= inv;
    }
    E.swap(R);
    for (int i = 0; i < sq; ++i) {
        matmul(n, E, E, tmp);
        E.swap(tmp);
    }
}

void relax_thresholds(const std::vector<double> &rates, double h, uint32_t t[4]) {
    t[0] = t[1] = t[2] = t[3] = 0u;
    if (rates.empty() || !(h > 0.0)) return;
    if (rates.size() == 1) {  // the single-channel branch (:1224-1232)
        const double p = -0.5 * std::expm1(-2.0 * (0.5 * rates[0]) * h);
        t[1] = (uint32_t)std::floor(p * 4294967296.0);
        return;
    }
    double mu = 0.0;  // superposed Poisson process (:1233-1252)
    for (double r : rates) mu += 0.5 * r * h;
    // T_n = ceil(F(n) 2^32) (count > n exactly when x >= T_n), held in 32 bits
    double term = std::exp(-mu), F = term;
    for (int n = 0; n < 3; ++n) {
        t[n] = (uint32_t)std::min(4294967295.0, std::ceil(F * 4294967296.0));
        term *= mu / (n + 1);
        F += term;
    }
    t[3] = 0xFFFFFFFFu;  // count 1 on [T_0, T_1), count 3 on [T_2, 2^32 - 1)
}

}  // namespace msg
