// Test-only host shim over the backend's host resolution (molspin_amd/csrc/msg_resolve.cpp):
// exposes the dense Hamiltonian in device order, the basis permutation and a host
// restatement of the matrix-free operator applied from the per-item coefficient
// blocks, so the CPU suite can check them against the oracle without a GPU.
#include <cstring>
#include <vector>

#include "../molspin_amd/csrc/msg_resolve.hpp"

using namespace msg;

namespace {
void apply_block(const float *p, bool herm, std::complex<float> m[4]) {
    std::complex<float> a(p[0], p[1]), b(p[2], p[3]), c(p[4], p[5]), d(p[6], p[7]);
    if (herm) { m[0] = std::conj(a); m[1] = std::conj(c); m[2] = std::conj(b); m[3] = std::conj(d); }
    else { m[0] = a; m[1] = b; m[2] = c; m[3] = d; }
}
}  // namespace

extern "C" {

int shim_resolve(const msg_system_desc *d, int64_t *dim, int *fast_ok, int *pair_ok, int *symmetric) {
    Resolved r;
    if (!resolve_system(d, r).empty()) return 1;
    *dim = r.D;
    *fast_ok = r.fast_ok;
    *pair_ok = r.pair_ok;
    *symmetric = r.symmetric;
    return 0;
}

int shim_trace_samples(const msg_system_desc *d, int state, int method, uint64_t seed, int m, double *out) {
    Resolved r;
    if (!resolve_system(d, r).empty()) return 1;
    return trace_samples(r, state, method, seed, m, out).empty() ? 0 : 2;
}

int shim_perm(const msg_system_desc *d, int64_t *perm) {
    Resolved r;
    if (!resolve_system(d, r).empty() || !r.pair_ok) return 1;
    std::memcpy(perm, r.perm.data(), r.perm.size() * sizeof(int64_t));
    return 0;
}

int shim_dense_hamiltonian(const msg_system_desc *d, const double *field, double *out) {
    Resolved r;
    if (!resolve_system(d, r).empty() || !r.pair_ok) return 1;
    std::vector<cd> H;
    dense_hamiltonian(r, field, H);
    std::memcpy(out, H.data(), H.size() * sizeof(cd));
    return 0;
}

// The per-nucleus flags of the matrix-free coefficient blocks at one field.
int shim_nuc_flags(const msg_system_desc *d, const double *field, int *flags) {
    Resolved r;
    if (!resolve_system(d, r).empty() || !r.fast_ok) return 1;
    std::vector<float> co;
    int opf = 0;
    fast_coefficients(r, field, co, opf, flags);
    return 0;
}

// y = H x with the matrix-free coefficient blocks (same algebra as apply_h in
// msg_kernels.hip); x, y in device order, interleaved complex float.
int shim_fast_apply(const msg_system_desc *d, const double *field, const float *x, float *y) {
    Resolved r;
    if (!resolve_system(d, r).empty() || !r.fast_ok) return 1;
    std::vector<float> co;
    int opf = 0, nf[kMaxNuclei];
    fast_coefficients(r, field, co, opf, nf);
    const int N = (int)r.nuc.size();
    const int ng = 1 << N;
    auto X = [&](int nu, int c) { return std::complex<float>(x[2 * (4 * nu + c)], x[2 * (4 * nu + c) + 1]); };
    for (int nu = 0; nu < ng; ++nu) {
        std::complex<float> out[4] = {};
        float f1[3] = {co[0], co[1], co[2]}, f2[3] = {co[3], co[4], co[5]}, rr = co[6];
        for (int n = 0; n < N; ++n) {
            const float *cn = &co[kCoefHead + n * kNucStride];
            const bool down = (nu >> n) & 1;
            const float s = down ? -1.f : 1.f;
            const int nb = nu ^ (1 << n);
            std::complex<float> m[4];
            // an axial coupling's block: [[0, 0], [b, 0]] with b = Re C[2], transposed for down
            auto axial = [&](float b) { m[0] = m[3] = 0.f; m[1] = down ? b : 0.f; m[2] = down ? 0.f : b; };
            if (nf[n] & NUC_E1) {
                if (nf[n] & NUC_A1) { f1[2] += s * cn[2]; axial(cn[11]); }
                else {
                    for (int a = 0; a < 3; ++a) f1[a] += s * cn[a];
                    apply_block(cn + 7, down, m);
                }
                for (int b = 0; b < 2; ++b) {
                    out[b] += m[0] * X(nb, b) + m[1] * X(nb, 2 + b);
                    out[2 + b] += m[2] * X(nb, b) + m[3] * X(nb, 2 + b);
                }
            }
            if (nf[n] & NUC_E2) {
                if (nf[n] & NUC_A2) { f2[2] += s * cn[5]; axial(cn[19]); }
                else {
                    for (int a = 0; a < 3; ++a) f2[a] += s * cn[3 + a];
                    apply_block(cn + 15, down, m);
                }
                for (int a = 0; a < 2; ++a) {
                    out[2 * a] += m[0] * X(nb, 2 * a) + m[1] * X(nb, 2 * a + 1);
                    out[2 * a + 1] += m[2] * X(nb, 2 * a) + m[3] * X(nb, 2 * a + 1);
                }
            }
            if (nf[n] & NUC_T) {
                rr += s * cn[6];
                const std::complex<float> tt(cn[23], down ? -cn[24] : cn[24]);
                for (int c = 0; c < 4; ++c) out[c] += tt * X(nb, c);
            }
        }
        for (int c = 0; c < 4; ++c) out[c] += rr * X(nu, c);
        // nucleus-nucleus couplings: row 2 b1 + b2 of each pair's 4x4 block
        for (size_t pc = 0; pc < r.nn_pairs.size(); ++pc) {
            const int k1 = r.nn_pairs[pc][0], k2 = r.nn_pairs[pc][1];
            const float *h = &co[kCoefHead + N * kNucStride + pc * kPairStride];
            const int row = (((nu >> k1) & 1) << 1) | ((nu >> k2) & 1);
            for (int col = 0; col < 4; ++col) {
                const int np = (nu & ~((1 << k1) | (1 << k2))) | ((col >> 1) << k1) | ((col & 1) << k2);
                const std::complex<float> hv(h[8 * row + 2 * col], h[8 * row + 2 * col + 1]);
                for (int c = 0; c < 4; ++c) out[c] += hv * X(np, c);
            }
        }
        auto fm = [](const float f[3], std::complex<float> m[4]) {
            m[0] = {0.5f * f[2], 0.f}; m[1] = {0.5f * f[0], -0.5f * f[1]};
            m[2] = {0.5f * f[0], 0.5f * f[1]}; m[3] = {-0.5f * f[2], 0.f};
        };
        std::complex<float> m[4];
        fm(f1, m);
        for (int b = 0; b < 2; ++b) {
            out[b] += m[0] * X(nu, b) + m[1] * X(nu, 2 + b);
            out[2 + b] += m[2] * X(nu, b) + m[3] * X(nu, 2 + b);
        }
        fm(f2, m);
        for (int a = 0; a < 2; ++a) {
            out[2 * a] += m[0] * X(nu, 2 * a) + m[1] * X(nu, 2 * a + 1);
            out[2 * a + 1] += m[2] * X(nu, 2 * a) + m[3] * X(nu, 2 * a + 1);
        }
        if (opf & OP_EREST)
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j)
                    out[i] += std::complex<float>(co[8 + 8 * i + 2 * j], co[8 + 8 * i + 2 * j + 1]) * X(nu, j);
        for (int c = 0; c < 4; ++c) {
            y[2 * (4 * nu + c)] = out[c].real();
            y[2 * (4 * nu + c) + 1] = out[c].imag();
        }
    }
    return 0;
}

}  // extern "C"
