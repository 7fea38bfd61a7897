// C ABI of the MI355X propagation backend (include/molspin_gpu.h).
//
// A msg_propagate call replaces, for a whole batch of items (field
// orientations / strengths), the orientation and time loops of
// TaskHSGeneral::RunLocal (TaskHSGeneral.cpp:347-577): observable sampling
// (HSObservableCollector::Evaluate), HSPropagator::StepStochastic and the
// trapezoid / finite-time correction (HSObservableCollector.cpp:304-392).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/molspin_gpu.h"
#include "msg_internal.hpp"
#include "msg_resolve.hpp"

namespace msg {
hipError_t launch_lanczos(const StepArgs &a, const StepArgs *da, int group_per_thread, int threads, int slots,
                          hipStream_t st);
hipError_t set_lanczos_lds_limits();
hipError_t launch_dense(const StepArgs &a, const StepArgs *da, int slots, hipStream_t st);
hipError_t launch_init_norms(const StepArgs &a, const StepArgs *da, hipStream_t st);
hipError_t launch_zero_int(int *p, int n, hipStream_t st);
hipError_t launch_reduce_columns(const double *in, double *out, int n_items, int ncols, int nk, int in_stride, int pick,
                                 hipStream_t st);
int lanczos_occupancy(int group_per_thread, int threads, size_t lds, bool arnoldi);
int rk4_occupancy(int group_per_thread, int threads, size_t lds);
int mix_occupancy(int mult, int threads, size_t lds);
bool mix_dbuf(int mult, int threads);
hipError_t launch_zlincomb(const void *X, const void *X2, const void *X3, void *out, int n, int batch,
                           const double *c, hipStream_t st);
hipError_t launch_blockdiag_right(const void *M, const double *B4, double scale, void *Y, int n, hipStream_t st);
hipError_t launch_lorentz_sum(const void *Pt, const void *Rt, const double *lam, double k, int n, double *rows,
                              double *out, hipStream_t st);
hipError_t launch_timeinf_sum(const void *Pt, const void *Rt, const double *lam, double a, int n, double *rows,
                              double *out, hipStream_t st);
hipError_t launch_herm_update(const void *Y, const void *R, void *S, int n, int batch, double c, double s,
                              hipStream_t st);
hipError_t launch_dyn_generator(const void *G, const void *Z, const double *coef, const int *item, const double *dt,
                               int nz, int n, int batch, void *out, hipStream_t st);
hipError_t launch_accumulate_fro(const void *X, const void *Delta, void *W, int n, int batch, int parts, double *part,
                                 hipStream_t st);
hipError_t launch_pair_block_trace(const void *X, int n, int batch, void *out, hipStream_t st);
hipError_t launch_pauli_right(const void *M, void *Y, int n, int mu, int batch, hipStream_t st);
hipError_t launch_corr(const void *S, size_t plane, const double *lam, const double *times, int n, int ns, int batch,
                       int nparts, size_t pstride, double *C, hipStream_t st);
hipError_t launch_unc_combine(const double *CA, const double *CB, int nparts, size_t pstride, const double *Q,
                              const double *wy, const double *wt, int n_items, int ns, int nobs, double dimA,
                              double dimB, double *yields, double *trace, hipStream_t st);
hipError_t launch_to_planes(const void *U, float *dst, int n, int Dp, size_t stride, int batch, hipStream_t st);
int dense_occupancy(int dpad, size_t lds);
int dense_threads(int dpad);
hipError_t launch_uneven(const UnevenArgs *du, int workgroups, hipStream_t st);
int uneven_occupancy();
hipError_t scratch_address(int device, void **ptr);
}  // namespace msg

using namespace msg;

struct msg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 0;
    std::string err;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double kernel_ms = 0.0;
    int launches = 0;
    rocblas_handle blas = nullptr;  // created on first use (eigensolver), bound to `stream`
    // Lanczos launch shapes chosen by msg_autotune, per Hilbert dimension
    struct Shape {
        int G = 0, T = 0, slots = 0;
    };
    std::map<int64_t, Shape> tuned;
    const Shape *force = nullptr;  // candidate under trial inside msg_autotune
    // pinned host staging of given columns (grown on demand, freed with the context)
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    // this device's copy of the allocation-free entry point's scratch (kScratchBytes),
    // looked up once at set-up
    void *scratch = nullptr;
};

struct msg_system {
    msg_ctx *ctx = nullptr;
    Resolved r;
};

namespace {

std::string g_create_error;

msg_status fail(msg_ctx *ctx, msg_status st, const std::string &m) {
    if (ctx) ctx->err = m;
    else g_create_error = m;
    return st;
}

#define HIPCHECK(call)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            throw DeviceError(std::string(#call) + ": " + hipGetErrorString(e_));                  \
    } while (0)

struct DeviceError {
    std::string what;
    explicit DeviceError(std::string w) : what(std::move(w)) {}
};
struct ArgError {
    msg_status st;
    std::string what;
};
// The device refused a cooperative launch (split columns wider than the resident
// capacity): nothing was launched.  Only the autotuner recovers from it.
struct CoopRefused {
    std::string what;
};

// Bump allocator over the allocation-free entry point's scratch: while an Arena is
// installed on the calling thread (ArenaScope), DBuf carves its buffers out of it
// instead of allocating, and never frees them.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, off = 0;
    size_t left() const { return cap - off; }
};
thread_local Arena *t_arena = nullptr;
struct ArenaScope {
    explicit ArenaScope(Arena *a) { t_arena = a; }
    ~ArenaScope() { t_arena = nullptr; }
};

// RAII device allocation, stream-ordered so that a batch never synchronises the
// device (other requests keep running while this one allocates or frees).
struct DBuf {
    void *p = nullptr;
    size_t n = 0;
    hipStream_t st = nullptr;
    bool carved = false;  // from the arena: nothing to free
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { reset(); }
    void reset() {
        if (p && !carved) (void)hipFreeAsync(p, st);
        p = nullptr;
        n = 0;
        carved = false;
    }
    template <class T>
    T *alloc(size_t count, hipStream_t stream) {
        reset();
        if (count == 0) return nullptr;
        st = stream;
        if (t_arena) {
            const size_t bytes = (count * sizeof(T) + 255) / 256 * 256;
            if (bytes > t_arena->left())
                throw ArgError{MSG_OOM_ADMISSION, "the run needs more than the " +
                                                      std::to_string(t_arena->cap >> 20) +
                                                      " MiB of the allocation-free scratch"};
            p = t_arena->base + t_arena->off;
            t_arena->off += bytes;
            carved = true;
        } else {
            HIPCHECK(hipMallocAsync(&p, count * sizeof(T), stream));
        }
        n = count * sizeof(T);
        return static_cast<T *>(p);
    }
    template <class T>
    T *get() const { return static_cast<T *>(p); }
};

enum class Path { Lanczos, Dense };

// One propagation batch: all device state of the items of one call or request.
struct Batch {
    msg_ctx *ctx = nullptr;
    const msg_system *sys = nullptr;
    msg_run_params prm{};
    Path path = Path::Lanczos;
    int n_items = 0, ncols = 0, n_units = 0, nobs = 0, nsamples = 0;
    int G = 1, T = 64, slots = 1;
    int share = 1;  // the launch keeps 1/share of the resident workgroups it could hold (server space sharing)
    bool no_split = false;  // no split-column shape (G = -3): its cooperative launch needs the whole device
    DBuf xred, bar;         // split columns: partial sums and barrier counters per column slot
    std::vector<double> times, scale;
    StepArgs a{};
    DBuf coef, U, dts, wy, wt, thr, yacc, trace, state, basis, queue, status, given, norm0, yout, tout;
    DBuf obs, evec, nflags, args;
    int s_next = 0;
    bool keep_state = false;
    hipStream_t stream = nullptr;
    size_t footprint = 0;
    StepArgs args_host{};  // host copy of the last launch arguments
    bool timed = false;    // bracket the propagation kernels with the context events
    int launches = 0;      // propagation kernels launched so far
    double kernel_ms = 0.0;  // device time of the launches before the last one
    // dynamics=dynamic on the dense path: one propagator per interval, built per chunk
    // of intervals from G_item (static generator) + sum_k c_k(t) Z_k
    bool arnoldi = false;  // matrix-free path on G = -iH - K (unequal recombination rates)
    bool dyn_dense = false;
    int dyn_chunk = 0;                      // intervals per propagator chunk
    DBuf dynG, dynZ, dyn_head;
    std::vector<double> dyn_gnorm, dyn_znorm, item_fields;
    std::vector<std::pair<int, int>> dyn_terms;  // (spin, axis) of each Z_k
    bool field_given = false;
    bool uneven = false;  // part of an uneven launch: the launch owns the Krylov bases
    Batch() = default;
    Batch(const Batch &) = delete;
};

const int kObsMax = kMaxObs;

void check_params(const msg_system *sys, const msg_run_params *p) {
    if (!p) throw ArgError{MSG_INVALID_ARG, "null run parameters"};
    if (p->calculation != MSG_CALC_YIELDS && p->calculation != MSG_CALC_TIMEEVOLUTION)
        throw ArgError{MSG_INVALID_ARG, "calculation must be yields or timeevolution"};
    if (p->observables != MSG_OBS_TRANSITION_YIELDS && p->observables != MSG_OBS_STATE_POPULATIONS)
        throw ArgError{MSG_INVALID_ARG, "unknown observable mode"};
    if (p->method != MSG_METHOD_AUTO && p->method != MSG_METHOD_KRYLOV && p->method != MSG_METHOD_EXPONENTIAL &&
        p->method != MSG_METHOD_RK4 && p->method != MSG_METHOD_UNCOUPLED)
        throw ArgError{MSG_INVALID_ARG, "propagationmethod must be auto, krylov, exponential, rk4 or uncoupled"};
    if (p->krylov_size < 0 || p->krylov_size > kMaxKrylov) throw ArgError{MSG_UNSUPPORTED, "krylovsize above 32"};
    if (p->sampling != MSG_SAMPLING_DIRECT && p->sampling != MSG_SAMPLING_GIVEN)
        throw ArgError{MSG_INVALID_ARG, "unknown sampling mode"};
    if (p->dynamics != 0 && p->dynamics != 1) throw ArgError{MSG_INVALID_ARG, "dynamics must be static or dynamic"};
    if (p->dynamics && p->task != MSG_TASK_HSGENERAL)
        throw ArgError{MSG_INVALID_ARG, "the legacy StaticHS tasks take static Hamiltonians"};
    if (!sys->r.pair_ok)
        throw ArgError{MSG_UNSUPPORTED,
                       "the device paths need exactly two spin-1/2 electrons and States over those electrons"};
}

int observable_count(const Resolved &r, const msg_run_params *p) {
    return p->observables == MSG_OBS_TRANSITION_YIELDS ? (int)r.P_trans.size() : (int)r.P_states.size();
}

// Symmetric recombination: K = kappa 1 (HSGeneral); the legacy tasks call any set of
// equal rates symmetric unless a source State is the identity
// (TaskStaticHSDirectYields.cpp:257-292) and then use kappa = k/2.
bool sym_for(const Resolved &r, const msg_run_params *p) {
    if (p->task == MSG_TASK_HSGENERAL) return r.symmetric;
    if (r.rates.empty()) return true;
    for (double k : r.rates)
        if (k != r.rates[0]) return false;
    for (const Mat4 &P : r.P_trans) {
        bool ident = true;
        for (int i = 0; i < 16; ++i) ident &= std::abs(P[i] - ((i % 5) == 0 ? cd(1.0) : cd(0.0))) < 1e-14;
        if (ident) return false;
    }
    return true;
}
double kappa_for(const Resolved &r, const msg_run_params *p) {
    if (p->task == MSG_TASK_HSGENERAL) return r.kappa;
    return r.rates.empty() ? 0.0 : 0.5 * r.rates[0];
}

Path choose_path(const Resolved &r, const msg_run_params *p) {
    if (p->method == MSG_METHOD_UNCOUPLED)
        throw ArgError{MSG_UNSUPPORTED, "propagationmethod=uncoupled is available through msg_propagate only"};
    const bool dense_ok = r.D <= 512;
    // a time-dependent field on a nucleus changes per-nucleus coefficients, which the
    // matrix-free head cannot carry; the dense path builds each interval's propagator
    const bool dyn_nuclear = p->dynamics && r.has_td && r.td_nuclear;
    const bool mixed = r.mix_mult > 0;  // one nucleus of spin > 1/2 on the group index (k_lanczos_mix)
    const bool lanczos_ok = r.fast_ok && sym_for(r, p) && r.D >= 256 && r.D <= (int64_t)4 << kMaxNuclei &&
                            !dyn_nuclear && !(mixed && p->dynamics && r.has_td);
    if (dyn_nuclear && !dense_ok)
        throw ArgError{MSG_UNSUPPORTED, "a time-dependent field on a nucleus needs D <= 512 (dense path)"};
    if (p->method == MSG_METHOD_RK4 && !dense_ok) {
        // matrix-free RK4 stages of -iH - K for a static Hamiltonian
        if (!r.fast_ok) throw ArgError{MSG_UNSUPPORTED, "matrix-free RK4: " + r.why_not_fast};
        if (mixed) throw ArgError{MSG_UNSUPPORTED, "matrix-free RK4 takes spin-1/2 nuclei only"};
        if (p->dynamics && r.has_td)
            throw ArgError{MSG_UNSUPPORTED, "propagationmethod=rk4 with a time-dependent field needs D <= 512"};
        if (r.D > 16384) throw ArgError{MSG_UNSUPPORTED, "matrix-free RK4 is limited to D <= 16384"};
        return Path::Lanczos;
    }
    if (p->method == MSG_METHOD_EXPONENTIAL || p->method == MSG_METHOD_RK4) {
        if (!dense_ok)
            throw ArgError{MSG_UNSUPPORTED, "propagationmethod=exponential needs D <= 512 on the device"};
        return Path::Dense;
    }
    if (p->method == MSG_METHOD_AUTO && dense_ok) return Path::Dense;
    if (lanczos_ok) return Path::Lanczos;
    if (dense_ok) return Path::Dense;  // exact exponential: same result as Krylov within truncation
    // unequal rates: matrix-free Arnoldi on G = -iH - K while the Krylov vector fits in LDS
    const bool arnoldi_ok = r.fast_ok && !mixed && !sym_for(r, p) && r.D <= 16384 && !dyn_nuclear;
    if (arnoldi_ok) return Path::Lanczos;
    if (!sym_for(r, p))
        throw ArgError{MSG_UNSUPPORTED, "unequal recombination rates need D <= 16384 (matrix-free Arnoldi) with "
                                        "spin-1/2 nuclei, or D <= 512 (dense exponential path)"};
    if (!r.fast_ok) throw ArgError{MSG_UNSUPPORTED, "matrix-free path: " + r.why_not_fast};
    throw ArgError{MSG_UNSUPPORTED, "Hilbert dimension outside the matrix-free range"};
}

const StepArgs *push_args(Batch &b, const StepArgs &a);

// The context's rocBLAS handle (created on first use, destroyed with the context),
// bound to the stream of the calling batch.  A context is used by one host thread at
// a time, so rebinding the stream per call is safe.
rocblas_handle blas_on(msg_ctx *ctx, hipStream_t st) {
    if (!ctx->blas && rocblas_create_handle(&ctx->blas) != rocblas_status_success) {
        ctx->blas = nullptr;
        throw DeviceError("rocblas_create_handle failed");
    }
    if (rocblas_set_stream(ctx->blas, st) != rocblas_status_success) throw DeviceError("rocblas_set_stream failed");
    return ctx->blas;
}

// Batched complex FP64 GEMM of row-major n x n operands (contiguous batches) on
// rocBLAS: C = coef A op(B) + E with op(B) = B or B^dagger.  A row-major array read
// column-major is the transpose, so C^T = op(B)^T A^T is one column-major call with
// the operands swapped (conj(B) A^T for op(B) = B^dagger).
hipError_t zgemm(msg_ctx *ctx, const void *A, const void *B, void *C, int n, int batch, double coef, hipStream_t st,
                 bool conj_b = false, const void *E = nullptr) {
    rocblas_handle h = blas_on(ctx, st);
    const size_t nn = (size_t)n * n;
    if (E && E != C) {
        const hipError_t e = hipMemcpyAsync(C, E, (size_t)batch * nn * sizeof(cd), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    }
    const rocblas_double_complex alpha(coef, 0.0), beta(E ? 1.0 : 0.0, 0.0);
    const rocblas_status rs = rocblas_zgemm_strided_batched(
        h, conj_b ? rocblas_operation_conjugate_transpose : rocblas_operation_none, rocblas_operation_none, n, n, n,
        &alpha, static_cast<const rocblas_double_complex *>(B), n, (rocblas_stride)nn,
        static_cast<const rocblas_double_complex *>(A), n, (rocblas_stride)nn, &beta,
        static_cast<rocblas_double_complex *>(C), n, (rocblas_stride)nn, batch);
    return rs == rocblas_status_success ? hipSuccess : hipErrorUnknown;
}

// exp(X) for a batch of nb row-major D x D matrices scaled to ||X||_1 <= 1/2: the
// degree-18 Taylor polynomial by Paterson-Stockmeyer in X^4 (powers X^2, X^3, X^4;
// then P = B_4 and P = P X^4 + B_k for k = 3..0 with B_k = sum_i X^i / (4k+i)!).
// With rk4 the polynomial is I + X + X^2/2 + X^3/6 + X^4/24 (one RK4 step of a static
// generator).  The result is left in *dP; dX2, dX3, dX4, dB and *dQ are workspace.
void taylor_batch(msg_ctx *ctx, cd *dX, cd *dX2, cd *dX3, cd *dX4, cd *dB, cd *&dP, cd *&dQ, int D, int nb,
                  bool rk4, hipStream_t stream) {
    double fact[19];  // 1/j!
    fact[0] = 1.0;
    for (int j = 1; j <= 18; ++j) fact[j] = fact[j - 1] / j;
    auto taylor_block = [&](int k, double c[4]) {  // coefficients of X^0..X^3 in block k
        for (int i = 0; i < 4; ++i) c[i] = 4 * k + i <= 18 ? fact[4 * k + i] : 0.0;
    };
    HIPCHECK(zgemm(ctx, dX, dX, dX2, D, nb, 1.0, stream));
    HIPCHECK(zgemm(ctx, dX2, dX, dX3, D, nb, 1.0, stream));
    double c[4];
    if (rk4) {
        taylor_block(0, c);
        HIPCHECK(launch_zlincomb(dX, dX2, dX3, dB, D, nb, c, stream));
        HIPCHECK(zgemm(ctx, dX2, dX2, dP, D, nb, 1.0 / 24.0, stream, false, dB));
        return;
    }
    HIPCHECK(zgemm(ctx, dX2, dX2, dX4, D, nb, 1.0, stream));
    taylor_block(4, c);
    HIPCHECK(launch_zlincomb(dX, dX2, dX3, dP, D, nb, c, stream));
    for (int k = 3; k >= 0; --k) {
        taylor_block(k, c);
        HIPCHECK(launch_zlincomb(dX, dX2, dX3, dB, D, nb, c, stream));
        HIPCHECK(zgemm(ctx, dP, dX4, dQ, D, nb, 1.0, stream, false, dB));
        std::swap(dP, dQ);
    }
}

double norm1(const cd *g, int D) {  // max column sum of a row-major D x D matrix
    std::vector<double> cs(D, 0.0);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) cs[j] += std::abs(g[(size_t)i * D + j]);
    return *std::max_element(cs.begin(), cs.end());
}

// Generators of a field sweep.  H(f) = H(0) + sum_a f_a Z_a is linear in the swept
// field, so the host builds H(0) and the three unit-field operators once and the device
// forms every item's scaled generator s_q (G0 + sum_a f_a Z_a) (k_dyn_generator), with
// G0 = -i H(0) - K (x) 1 (with_k) or -i H(0), and Z_a = -i (H(e_a) - H(0)).
struct SweepGenerators {
    DBuf G0, Z, coef, item, scale;
    double g0norm = 0.0, znorm[3] = {0.0, 0.0, 0.0};
    std::vector<double> hcoef, hscale;

    void build(const Resolved &r, int D, bool with_k, int chunk, hipStream_t st) {
        const size_t DD = (size_t)D * D;
        const double zero[3] = {0.0, 0.0, 0.0};
        std::vector<cd> H, H1, g(DD), z(3 * DD);
        dense_hamiltonian(r, zero, H);
        for (size_t k = 0; k < DD; ++k) g[k] = cd(H[k].imag(), -H[k].real());  // -i H(0)
        if (with_k)
            for (int nu = 0; nu < D / 4; ++nu)
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j) g[(size_t)(4 * nu + i) * D + 4 * nu + j] -= r.K[i * 4 + j];
        g0norm = norm1(g.data(), D);
        for (int ax = 0; ax < 3; ++ax) {
            double e[3] = {0.0, 0.0, 0.0};
            e[ax] = 1.0;
            dense_hamiltonian(r, e, H1);
            cd *za = &z[(size_t)ax * DD];
            for (size_t k = 0; k < DD; ++k) {
                const cd d = H1[k] - H[k];
                za[k] = cd(d.imag(), -d.real());  // -i Z_a
            }
            znorm[ax] = norm1(za, D);
        }
        HIPCHECK(hipMemcpyAsync(G0.alloc<cd>(DD, st), g.data(), DD * sizeof(cd), hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(Z.alloc<cd>(3 * DD, st), z.data(), 3 * DD * sizeof(cd), hipMemcpyHostToDevice, st));
        coef.alloc<double>((size_t)3 * chunk, st);
        scale.alloc<double>(chunk, st);
        item.alloc<int>(chunk, st);
        HIPCHECK(hipMemsetAsync(item.p, 0, (size_t)chunk * sizeof(int), st));  // every entry uses G0
        HIPCHECK(hipStreamSynchronize(st));  // g and z are released at scope end
        hcoef.resize((size_t)3 * chunk);
        hscale.resize(chunk);
    }
    // upper bound of ||G(f)||_1
    double bound(const double *f) const {
        return g0norm + std::fabs(f[0]) * znorm[0] + std::fabs(f[1]) * znorm[1] + std::fabs(f[2]) * znorm[2];
    }
    // out[q] = s (G0 + sum_a f_a Z_a), f = fields + 3 q, for q < nb
    void form(const double *fields, int nb, double s, int D, cd *out, hipStream_t st) {
        for (int q = 0; q < nb; ++q) {
            for (int ax = 0; ax < 3; ++ax) hcoef[3 * (size_t)q + ax] = fields[3 * (size_t)q + ax];
            hscale[q] = s;
        }
        HIPCHECK(hipMemcpyAsync(coef.p, hcoef.data(), (size_t)3 * nb * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(scale.p, hscale.data(), (size_t)nb * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHECK(launch_dyn_generator(G0.p, Z.p, coef.get<double>(), item.get<int>(), scale.get<double>(), 3, D, nb, out,
                                      st));
        HIPCHECK(hipStreamSynchronize(st));  // hcoef / hscale are refilled by the next call
    }
};

// Padded dimension of the dense path: 16 x (one of the row-tile counts the
// kernel is compiled for); 0 above 512.
int dense_padding(int D) {
    for (int nt : {1, 2, 3, 4, 6, 8, 12, 16, 24, 32})
        if (16 * nt >= D) return 16 * nt;
    return 0;
}

// Builds a batch and uploads everything a launch needs.
void prepare(Batch &b, msg_ctx *ctx, const msg_system *sys, const msg_run_params *p, int n_items,
             const double *fields, const double *columns, int n_columns, bool want_trace, bool keep_state,
             hipStream_t stream) {
    const Resolved &r = sys->r;
    check_params(sys, p);
    if (n_items < 1) throw ArgError{MSG_INVALID_ARG, "n_items must be positive"};
    b.ctx = ctx;
    b.sys = sys;
    b.prm = *p;
    b.stream = stream;
    b.keep_state = keep_state;
    b.path = choose_path(r, p);
    b.n_items = n_items;
    const int D = (int)r.D;
    // launch shape from the autotuner (a candidate under trial, or the tuned one)
    const msg_ctx::Shape *ts = ctx->force;
    if (!ts) {
        auto it = ctx->tuned.find(r.D);
        if (it != ctx->tuned.end()) ts = &it->second;
    }
    // columns
    if (p->sampling == MSG_SAMPLING_GIVEN) {
        if (!columns || n_columns < 1) throw ArgError{MSG_INVALID_ARG, "given sampling needs columns"};
        b.ncols = n_columns;
    } else {
        if (r.evec.empty()) throw ArgError{MSG_INVALID_ARG, "the system defines no initial state"};
        if ((int)r.evec.size() > kMaxInitial) throw ArgError{MSG_UNSUPPORTED, "initial density of rank above 4"};
        if (r.zdim * (int64_t)r.evec.size() > (1 << 30)) throw ArgError{MSG_UNSUPPORTED, "too many columns"};
        b.ncols = (int)(r.zdim * (int64_t)r.evec.size());
    }
    // observables
    b.nobs = observable_count(r, p);
    if (b.nobs < 1) throw ArgError{MSG_INVALID_ARG, "no observables (no transitions / States)"};
    if (b.nobs > kObsMax) throw ArgError{MSG_UNSUPPORTED, "more than 8 observables"};
    const std::vector<Mat4> &P = p->observables == MSG_OBS_TRANSITION_YIELDS ? r.P_trans : r.P_states;
    b.scale.assign(b.nobs, 1.0);
    if (p->observables == MSG_OBS_TRANSITION_YIELDS) b.scale = r.rates;
    // time grid
    const bool legacy = p->task == MSG_TASK_LEGACY_DIRECT_YIELDS || p->task == MSG_TASK_LEGACY_STOCH_YIELDS;
    if (!(p->task == MSG_TASK_HSGENERAL || legacy)) throw ArgError{MSG_INVALID_ARG, "unknown task"};
    const std::string te = legacy ? legacy_times(p->total_time, p->time_step, b.times)
                                  : free_times(p->total_time, p->time_step, b.times);
    if (!te.empty()) throw ArgError{MSG_INVALID_ARG, te};
    b.nsamples = (int)b.times.size();
    const int ns = b.nsamples;
    std::vector<double> dts(std::max(ns - 1, 1), 0.0), wyh((size_t)ns * b.nobs), wth((size_t)ns * b.nobs);
    std::vector<uint32_t> thr(4 * (size_t)std::max(ns - 1, 1), 0u);
    // symmetric K = kappa 1 is factored out as e^{-2 kappa t} -- exact for the exponential
    // and Krylov propagators, not for RK4's polynomial, which keeps K inside
    const bool factor_decay = sym_for(r, p) && p->method != MSG_METHOD_RK4;
    const double kappa = factor_decay ? kappa_for(r, p) : 0.0;
    for (int s = 0; s + 1 < ns; ++s) {
        dts[s] = b.times[s + 1] - b.times[s];
        if (r.relax) relax_thresholds(r.relax_rates, 0.5 * dts[s], &thr[4 * (size_t)s]);
    }
    for (int s = 0; s < ns; ++s) {
        double w = 0.0;
        if (s > 0) w += 0.5 * (b.times[s] - b.times[s - 1]);
        if (s + 1 < ns) w += 0.5 * (b.times[s + 1] - b.times[s]);
        const double decay = factor_decay ? std::exp(-2.0 * kappa * b.times[s]) : 1.0;
        for (int c = 0; c < b.nobs; ++c) {
            wyh[(size_t)s * b.nobs + c] = w * decay * b.scale[c];
            wth[(size_t)s * b.nobs + c] = decay * b.scale[c];
        }
    }
    StepArgs &a = b.a;
    a = StepArgs();
    a.dim = D;
    a.nobs = b.nobs;
    a.ncols = b.ncols;
    a.nsamples = ns;
    {
        std::vector<float> ob((size_t)32 * b.nobs);
        for (int c = 0; c < b.nobs; ++c)
            for (int i = 0; i < 16; ++i) {
                ob[32 * c + 2 * i] = (float)P[c][i].real();
                ob[32 * c + 2 * i + 1] = (float)P[c][i].imag();
            }
        float *d = b.obs.alloc<float>(ob.size(), stream);
        HIPCHECK(hipMemcpyAsync(d, ob.data(), ob.size() * sizeof(float), hipMemcpyHostToDevice, stream));
        a.obs = d;
    }
    a.relax = r.relax ? 1 : 0;
    a.seed = p->seed;
    if (!(p->krylov_tol >= 0.0) || !std::isfinite(p->krylov_tol))
        throw ArgError{MSG_INVALID_ARG, "krylov_tol must be finite and non-negative"};
    // basis reuse: legacy direct yields with symmetric recombination (the decay factored out)
    a.ktol = p->task == MSG_TASK_LEGACY_DIRECT_YIELDS && factor_decay && !r.relax ? p->krylov_tol : 0.0;
    a.m = p->krylov_size > 0 ? p->krylov_size : 16;
    a.renorm = 0;
    // initial states
    if (p->sampling == MSG_SAMPLING_GIVEN) {
        a.init_mode = 1;
        const size_t gbytes = (size_t)n_columns * 2 * D * sizeof(float);
        std::vector<float> pageable;  // allocation-free runs: no pinned staging
        if (t_arena) {
            pageable.resize(gbytes / sizeof(float));
        } else if (ctx->pinned_bytes < gbytes) {  // page-locked, so the copy runs at DMA speed
            if (ctx->pinned) (void)hipHostFree(ctx->pinned);
            ctx->pinned = nullptr;
            ctx->pinned_bytes = 0;
            HIPCHECK(hipHostMalloc(&ctx->pinned, gbytes, hipHostMallocDefault));
            ctx->pinned_bytes = gbytes;
        }
        float *g = t_arena ? pageable.data() : static_cast<float *>(ctx->pinned);
        // MolSpin order (complex128) -> device planes (complex64); columns split over host threads
        auto convert = [&](int j0, int j1) {
            for (int j = j0; j < j1; ++j)
                for (int i = 0; i < D; ++i) {
                    const int64_t dev = r.perm[i];
                    const int64_t nu = dev >> 2, c = dev & 3;
                    const size_t off = (size_t)j * 2 * D + (c < 2 ? 4 * nu + 2 * c : D + 4 * nu + 2 * (c - 2));
                    g[off] = (float)columns[2 * ((size_t)j * D + i)];
                    g[off + 1] = (float)columns[2 * ((size_t)j * D + i) + 1];
                }
        };
        const int64_t work = (int64_t)n_columns * D;
        const int nth = (int)std::min<int64_t>({16, std::max(1u, std::thread::hardware_concurrency()),
                                                 std::max<int64_t>(1, work >> 20), (int64_t)n_columns});
        if (nth <= 1) {
            convert(0, n_columns);
        } else {
            std::vector<std::thread> pool;
            for (int k = 0; k < nth; ++k)
                pool.emplace_back(convert, (int)((int64_t)n_columns * k / nth), (int)((int64_t)n_columns * (k + 1) / nth));
            for (auto &th : pool) th.join();
        }
        float *dg = b.given.alloc<float>((size_t)n_columns * 2 * D, stream);
        HIPCHECK(hipMemcpyAsync(dg, g, gbytes, hipMemcpyHostToDevice, stream));
        HIPCHECK(hipStreamSynchronize(stream));  // the staging buffer is reused by the next call
        a.given = dg;
    } else {
        a.init_mode = 0;
        a.n_init = (int)r.evec.size();
        a.zdim = (int)r.zdim;
        const double inv = 1.0 / std::sqrt((double)r.zdim);
        std::vector<float> ev((size_t)8 * a.n_init);
        for (int i = 0; i < a.n_init; ++i)
            for (int c = 0; c < 4; ++c) {
                ev[i * 8 + 2 * c] = (float)(r.evec[i][c].real() * inv);
                ev[i * 8 + 2 * c + 1] = (float)(r.evec[i][c].imag() * inv);
            }
        float *d = b.evec.alloc<float>(ev.size(), stream);
        HIPCHECK(hipMemcpyAsync(d, ev.data(), ev.size() * sizeof(float), hipMemcpyHostToDevice, stream));
        a.evec = d;
    }
    // operator
    if (b.path == Path::Lanczos) {
        const int ng = D / 4;
        // launch shape: groups per thread G, threads T (one workgroup per column);
        // G = 0 selects the streaming variant once two Krylov vectors exceed LDS,
        // G = -1 the streaming variant with v_j mirrored in LDS (one vector fits)
        if (ng <= 512) { b.T = ng; b.G = 1; }
        else if (ng <= 2048) { b.G = 2; b.T = ng / 2; }
        else if (ng <= 4096) { b.G = -1; b.T = 1024; }
        else { b.G = -2; b.T = 1024; }  // tiled streaming
        if (ts) {
            b.G = ts->G;
            b.T = ts->T;
        }
        if (const char *env = std::getenv("MSG_LANCZOS_G")) {  // launch-shape override (tuning runs)
            const int g = std::atoi(env);
            if ((g == 1 && ng >= 64 && ng <= 1024) || (g == 2 && ng >= 256 && ng <= 2048) || (g == 4 && ng == 1024)) {
                b.G = g;
                b.T = ng / g;
            }
            if ((g == 0 && ng > 2048) || (g == -1 && ng == 4096)) { b.G = g; b.T = 1024; }
            if (g == -3 && ng >= 4096) {  // split columns: 512 threads, 1024 with MSG_SPLIT_T=1024
                const char *st = std::getenv("MSG_SPLIT_T");
                b.G = -3;
                b.T = st && std::atoi(st) == 1024 ? 1024 : 512;
            }
            if (g == -2 && ng > kTileGroups) { b.G = -2; b.T = 1024; }
        }
        if (b.G == -3 && (b.no_split || ng < 4096)) {  // back to the default streaming shape
            b.G = ng <= 4096 ? -1 : -2;
            b.T = 1024;
        }
        const bool rk4 = p->method == MSG_METHOD_RK4;
        b.arnoldi = !sym_for(r, p) && !rk4;
        if (r.mix_mult > 0) {  // spin-1/2 nuclei on the thread index, the other nucleus on the group index
            b.G = r.mix_mult;
            b.T = ng / r.mix_mult;
            a.mixm = r.mix_mult;
        }
        if (rk4) {  // one column per workgroup, the stage argument in LDS: 1, 2 or 4 groups per thread
            b.G = ng <= 512 ? 1 : ng <= 2048 ? 2 : 4;
            b.T = ng / b.G;
            a.rk4 = 1;
            for (int i = 0; i < 16; ++i) {
                a.kmat[2 * i] = (float)r.K[i].real();
                a.kmat[2 * i + 1] = (float)r.K[i].imag();
            }
        }
        if (b.arnoldi) {  // G = -iH - K: Arnoldi with the Krylov vector in LDS, one group per thread up to D = 4096
            if (ng > 4096) throw ArgError{MSG_UNSUPPORTED, "matrix-free Arnoldi is limited to D <= 16384"};
            b.G = ng <= 1024 ? 1 : ng <= 2048 ? 2 : 4;
            b.T = ng / b.G;
            // a tuned, forced or requested shape with 1 or 2 groups per thread: D = 1024,
            // 2048 and 4096 run either
            int want = ts && ts->G > 0 ? ts->G : 0;
            if (const char *env = std::getenv("MSG_LANCZOS_G")) want = std::atoi(env);
            if (want == 2 && ng >= 256 && ng <= 1024) { b.G = 2; b.T = ng / 2; }
            if (want == 1 && ng <= 1024) { b.G = 1; b.T = ng; }
            a.arnoldi = 1;
            for (int i = 0; i < 16; ++i) {
                a.kmat[2 * i] = (float)r.K[i].real();
                a.kmat[2 * i + 1] = (float)r.K[i].imag();
            }
        }
        if (b.G > 0 && (b.T < 64 || b.T > 1024 || b.G * b.T != ng))
            throw ArgError{MSG_UNSUPPORTED, "no matrix-free launch shape for this dimension"};
        // nucleus-nucleus couplings read three partner groups across any two index bits:
        // the shapes that hold the whole Krylov vector in LDS (the grouped operator)
        if (!r.nn_pairs.empty() && b.G <= 0)
            throw ArgError{MSG_UNSUPPORTED, "nucleus-nucleus couplings on the matrix-free path need D <= 8192 "
                                            "(D <= 16384 with unequal rates or RK4), or D <= 512 (dense path)"};
        a.nnuc = (int)r.nuc.size();
        std::vector<float> all, one;
        int flags = 0;
        std::vector<int> nfl(kMaxNuclei, 0), axial(kMaxNuclei, NUC_A1 | NUC_A2);
        const bool dyn = p->dynamics && r.has_td;
        TimePoint static_part;
        static_part.skip = true;
        for (int it = 0; it < n_items; ++it) {
            int nf[kMaxNuclei];
            fast_coefficients(r, fields ? fields + 3 * it : nullptr, one, flags, nf, dyn ? &static_part : nullptr);
            if (it == 0) {
                a.coef_stride = (int)((one.size() + 15) / 16 * 16);
                all.assign((size_t)n_items * a.coef_stride, 0.f);
            }
            for (int k = 0; k < kMaxNuclei; ++k) {
                nfl[k] |= nf[k] & ~(NUC_A1 | NUC_A2);
                axial[k] &= nf[k];  // the short form only where every item's coupling is axial
            }
            a.op_flags |= flags;
            std::copy(one.begin(), one.end(), all.begin() + (size_t)it * a.coef_stride);
        }
        bool all_axial = true;
        for (int k = 0; k < kMaxNuclei; ++k) {
            nfl[k] |= axial[k];
            if (((nfl[k] & NUC_E1) && !(nfl[k] & NUC_A1)) || ((nfl[k] & NUC_E2) && !(nfl[k] & NUC_A2)))
                all_axial = false;
        }
        if (all_axial) a.op_flags |= OP_AXIAL;
        for (const auto &pr : r.nn_pairs) {  // the couplings' nucleus bits follow the flags
            nfl.push_back(pr[0]);
            nfl.push_back(pr[1]);
        }
        int *dfl = b.nflags.alloc<int>(nfl.size(), stream);
        HIPCHECK(hipMemcpyAsync(dfl, nfl.data(), nfl.size() * sizeof(int), hipMemcpyHostToDevice, stream));
        a.nuc_flags = dfl;
        a.npairs = (int)r.nn_pairs.size();
        a.pair_bits = a.npairs ? dfl + kMaxNuclei : nullptr;
        float *dc = b.coef.alloc<float>(all.size(), stream);
        HIPCHECK(hipMemcpyAsync(dc, all.data(), all.size() * sizeof(float), hipMemcpyHostToDevice, stream));
        if (dyn) {
            // per-interval head: the static w1, w2, shift plus the time-dependent electron
            // fields at the interval's midpoint (TaskHSGeneral.cpp:527-537)
            std::vector<float> heads((size_t)n_items * (ns - 1) * 8, 0.f);
            for (int it = 0; it < n_items; ++it) {
                const double *f = fields ? fields + 3 * it : nullptr;
                for (int s = 0; s + 1 < ns; ++s) {
                    TimePoint tp;
                    tp.t = b.times[s] + 0.5 * (b.times[s + 1] - b.times[s]);
                    double h[8];
                    for (int k = 0; k < 8; ++k) h[k] = all[(size_t)it * a.coef_stride + k];
                    for (const Term1 &t : r.t1) {
                        if (!t.td) continue;
                        double v[3];
                        term1_value(t, f, &tp, v);
                        const int off = t.s == r.e1 ? 0 : 3;
                        for (int k = 0; k < 3; ++k) h[off + k] += v[k];
                    }
                    float *dst = &heads[((size_t)it * (ns - 1) + s) * 8];
                    for (int k = 0; k < 8; ++k) dst[k] = (float)h[k];
                }
            }
            float *dh = b.dyn_head.alloc<float>(heads.size(), stream);
            HIPCHECK(hipMemcpyAsync(dh, heads.data(), heads.size() * sizeof(float), hipMemcpyHostToDevice, stream));
            HIPCHECK(hipStreamSynchronize(stream));  // `heads` is released at scope end
            a.dyn_head = dh;
        }
        a.coef = dc;
        b.n_units = n_items * b.ncols;
        a.renorm = factor_decay ? 1 : 0;
    } else if (p->dynamics && r.has_td) {
        // dense path, time-dependent H: one propagator per interval, built chunk by chunk
        // in advance(); here the static generator of every item and the time-dependent
        // operators Z_k = -i S_a (spin s) of the spins with time-dependent fields
        const int Dp = dense_padding(D);
        if (Dp == 0) throw ArgError{MSG_UNSUPPORTED, "dense path limited to D <= 512"};
        a.dpad = Dp;
        a.ustep = 1;
        b.dyn_dense = true;
        b.keep_state = true;
        const size_t DD = (size_t)D * D;
        std::vector<cd> H, all(DD * n_items);
        TimePoint static_part;
        static_part.skip = true;
        b.dyn_gnorm.assign(n_items, 0.0);
        for (int it = 0; it < n_items; ++it) {
            dense_hamiltonian(r, fields ? fields + 3 * it : nullptr, H, &static_part);
            cd *g = &all[(size_t)it * DD];
            for (size_t k = 0; k < DD; ++k) g[k] = cd(H[k].imag(), -H[k].real());  // -i H
            if (!factor_decay)
                for (int nu = 0; nu < D / 4; ++nu)
                    for (int i = 0; i < 4; ++i)
                        for (int j = 0; j < 4; ++j) g[(size_t)(4 * nu + i) * D + 4 * nu + j] -= r.K[i * 4 + j];
            b.dyn_gnorm[it] = norm1(g, D);
        }
        cd *dG = b.dynG.alloc<cd>(all.size(), stream);
        HIPCHECK(hipMemcpyAsync(dG, all.data(), all.size() * sizeof(cd), hipMemcpyHostToDevice, stream));
        std::vector<int> td_spins;
        for (const Term1 &t : r.t1)
            if (t.td && std::find(td_spins.begin(), td_spins.end(), t.s) == td_spins.end()) td_spins.push_back(t.s);
        b.dyn_terms.clear();
        std::vector<cd> Z(3 * td_spins.size() * DD);
        b.dyn_znorm.assign(3 * td_spins.size(), 0.0);
        for (size_t q = 0; q < td_spins.size(); ++q)
            for (int ax = 0; ax < 3; ++ax) {
                const size_t k = 3 * q + ax;
                b.dyn_terms.push_back({td_spins[q], ax});
                spin_operator_dense(r, td_spins[q], ax, H);
                cd *z = &Z[k * DD];
                for (size_t e = 0; e < DD; ++e) z[e] = cd(H[e].imag(), -H[e].real());
                b.dyn_znorm[k] = norm1(z, D);
            }
        cd *dZ = b.dynZ.alloc<cd>(Z.size(), stream);
        HIPCHECK(hipMemcpyAsync(dZ, Z.data(), Z.size() * sizeof(cd), hipMemcpyHostToDevice, stream));
        HIPCHECK(hipStreamSynchronize(stream));  // host staging is released at scope end
        b.field_given = fields != nullptr;
        if (fields) b.item_fields.assign(fields, fields + 3 * (size_t)n_items);
        // propagator planes of one chunk: at most ~1 GiB
        const size_t per = (size_t)n_items * kDensePlanes * Dp * Dp * sizeof(float);
        b.dyn_chunk = (int)std::max<size_t>(1, std::min<size_t>(ns - 1, ((size_t)1 << 30) / per));
        if (const char *env = std::getenv("MSG_DYN_CHUNK"))  // smaller chunks (tests of the resume)
            if (std::atoi(env) >= 1) b.dyn_chunk = std::min(b.dyn_chunk, std::atoi(env));
        a.ustride = b.dyn_chunk;
        a.renorm = factor_decay ? 1 : 0;
        b.n_units = n_items * ((b.ncols + 15) / 16);
    } else {
        const int Dp = dense_padding(D);
        if (Dp == 0) throw ArgError{MSG_UNSUPPORTED, "dense path limited to D <= 512"};
        a.dpad = Dp;
        const bool two = ns >= 2 && std::fabs(dts[ns - 2] - dts[0]) > 0.0;
        a.ulast = two ? 1 : 0;
        const int per = two ? 2 : 1;
        // exp(-i H_eff dt) on the device in FP64: Taylor(18) of X = A / 2^s evaluated by
        // Paterson-Stockmeyer in X^4 (7 complex GEMMs), then s squarings; items are
        // processed in chunks to bound the FP64 workspace
        float *dU = b.U.alloc<float>((size_t)n_items * per * kDensePlanes * Dp * Dp, stream);
        const int chunk = std::max(1, std::min(n_items, (int)((size_t)128 << 20) / (D * D * 16)));
        DBuf X, P, Q, X2, X3, X4, Bk;
        cd *dX = X.alloc<cd>((size_t)chunk * D * D, stream);
        cd *dP = P.alloc<cd>((size_t)chunk * D * D, stream);
        cd *dQ = Q.alloc<cd>((size_t)chunk * D * D, stream);
        cd *dX2 = X2.alloc<cd>((size_t)chunk * D * D, stream);
        cd *dX3 = X3.alloc<cd>((size_t)chunk * D * D, stream);
        cd *dX4 = X4.alloc<cd>((size_t)chunk * D * D, stream);
        cd *dB = Bk.alloc<cd>((size_t)chunk * D * D, stream);
        std::vector<cd> H, Hs;
        // a field sweep forms every item's generator on the device (SweepGenerators)
        const bool linear = fields != nullptr && n_items > 1;
        const size_t DD = (size_t)D * D;
        SweepGenerators sweep;
        if (linear)
            sweep.build(r, D, !factor_decay, chunk, stream);
        else
            Hs.resize((size_t)chunk * DD);
        for (int w = 0; w < per; ++w) {
            const double dt = w == 0 ? dts[0] : dts[ns - 2];
            for (int it0 = 0; it0 < n_items; it0 += chunk) {
                const int nb = std::min(chunk, n_items - it0);
                double nrm = 0.0;
                if (linear) {
                    for (int q = 0; q < nb; ++q) nrm = std::max(nrm, dt * sweep.bound(fields + 3 * (size_t)(it0 + q)));
                    int sq = 0;
                    const bool rk4 = p->method == MSG_METHOD_RK4;
                    while (!rk4 && nrm > 0.5) { nrm *= 0.5; ++sq; }
                    sweep.form(fields + 3 * (size_t)it0, nb, std::ldexp(dt, -sq), D, dX, stream);
                    taylor_batch(ctx, dX, dX2, dX3, dX4, dB, dP, dQ, D, nb, rk4, stream);
                    for (int q = 0; q < sq; ++q) {
                        HIPCHECK(zgemm(ctx, dP, dP, dQ, D, nb, 1.0, stream));
                        std::swap(dP, dQ);
                    }
                    HIPCHECK(launch_to_planes(dP, dU + ((size_t)it0 * per + w) * kDensePlanes * Dp * Dp, D, Dp,
                                              (size_t)per * kDensePlanes * Dp * Dp, nb, stream));
                    continue;
                }
                for (int q = 0; q < nb; ++q) {
                    dense_hamiltonian(r, fields ? fields + 3 * (it0 + q) : nullptr, H);
                    if (!factor_decay)  // H_eff = H - i K (x) 1 on the electron pair
                        for (int64_t nu = 0; nu < D / 4; ++nu)
                            for (int i = 0; i < 4; ++i)
                                for (int j = 0; j < 4; ++j)
                                    H[(size_t)(nu * 4 + i) * D + nu * 4 + j] += cd(0.0, -1.0) * r.K[i * 4 + j];
                    nrm = std::max(nrm, norm1(H.data(), D) * dt);  // 1-norm of A = -i dt H_eff
                    for (size_t k = 0; k < H.size(); ++k) Hs[(size_t)q * D * D + k] = cd(0.0, -dt) * H[k];
                }
                int sq = 0;
                const bool rk4 = p->method == MSG_METHOD_RK4;
                while (!rk4 && nrm > 0.5) { nrm *= 0.5; ++sq; }
                const double scale = std::ldexp(1.0, -sq);
                for (auto &x : Hs) x *= scale;
                HIPCHECK(hipMemcpyAsync(dX, Hs.data(), (size_t)nb * D * D * sizeof(cd), hipMemcpyHostToDevice, stream));
                taylor_batch(ctx, dX, dX2, dX3, dX4, dB, dP, dQ, D, nb, rk4, stream);
                for (int q = 0; q < sq; ++q) {
                    HIPCHECK(zgemm(ctx, dP, dP, dQ, D, nb, 1.0, stream));
                    std::swap(dP, dQ);
                }
                HIPCHECK(launch_to_planes(dP, dU + ((size_t)it0 * per + w) * kDensePlanes * Dp * Dp, D, Dp,
                                          (size_t)per * kDensePlanes * Dp * Dp, nb, stream));
            }
        }
        // X, P, Q are released (stream-ordered) after the conversion
        a.U = dU;
        a.ustride = per;
        a.renorm = factor_decay ? 1 : 0;  // U unitary: keep each column at its initial norm
        b.n_units = n_items * ((b.ncols + 15) / 16);
    }
    a.n_units = b.n_units;
    // device arrays
    double *ddts = b.dts.alloc<double>(dts.size(), stream);
    HIPCHECK(hipMemcpyAsync(ddts, dts.data(), dts.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    a.dts = ddts;
    double *dwy = b.wy.alloc<double>(wyh.size(), stream);
    HIPCHECK(hipMemcpyAsync(dwy, wyh.data(), wyh.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    a.wy = dwy;
    double *dwt = b.wt.alloc<double>(wth.size(), stream);
    HIPCHECK(hipMemcpyAsync(dwt, wth.data(), wth.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    a.wt = dwt;
    uint32_t *dthr = b.thr.alloc<uint32_t>(thr.size(), stream);
    HIPCHECK(hipMemcpyAsync(dthr, thr.data(), thr.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    a.thr = dthr;
    a.yacc = b.yacc.alloc<double>((size_t)b.n_units * b.nobs * 2, stream);
    a.trace = want_trace ? b.trace.alloc<double>((size_t)b.n_units * ns * b.nobs, stream) : nullptr;
    a.queue = b.queue.alloc<int>(1, stream);
    a.status = b.status.alloc<int>(1, stream);
    HIPCHECK(launch_zero_int(a.status, 1, stream));  // a fill kernel: no memset node under graph capture
    if (b.keep_state) {
        const size_t per_unit = b.path == Path::Lanczos ? (size_t)2 * D : (size_t)2 * a.dpad * 16;
        a.state = b.state.alloc<float>((size_t)b.n_units * per_unit, stream);
    }
    // launch geometry
    if (b.path == Path::Lanczos && b.uneven) {
        a.norm0 = b.norm0.alloc<double>(b.n_units, stream);
        if (a.renorm) HIPCHECK(launch_init_norms(a, push_args(b, a), stream));
    } else if (b.path == Path::Lanczos) {
        const size_t lds = b.G == 0 ? 0 : b.G == -1 ? (size_t)8 * D : b.G == -2 ? (size_t)32 * kTileGroups
                         : b.G == -3 ? (size_t)4 * kSplitSlab * 16
                         : a.mixm ? (size_t)(mix_dbuf(b.G, b.T) ? 16 : 8) * D + 4096
                         : (b.arnoldi || a.rk4) ? (size_t)8 * D + 4096 : (size_t)16 * D + 4096;
        const int occ = std::max(1, a.mixm  ? mix_occupancy(b.G, b.T, lds)
                                    : a.rk4 ? rk4_occupancy(b.G, b.T, lds)
                                            : lanczos_occupancy(b.G, b.T, lds, b.arnoldi));
        // split columns: a slot is S workgroups, all resident at once
        const int S = b.G == -3 ? (int)(D / 4 / kSplitSlab) : 1;
        if (ctx->cus * occ < S) throw ArgError{MSG_UNSUPPORTED, "split column wider than the device"};
        b.slots = std::max(1, std::min(b.n_units, ctx->cus * occ / S));
        if (ts && ts->slots > 0) b.slots = std::max(1, std::min(b.slots, ts->slots));
        if (b.share > 1) b.slots = std::max(1, std::min(b.slots, ctx->cus * occ / b.share));
        if (const char *env = std::getenv("MSG_LANCZOS_SLOTS")) {  // resident-workgroup override (tuning runs)
            const int sl = std::atoi(env);
            if (sl >= 1) b.slots = std::max(1, std::min(b.slots, sl));
        }
        if (b.G == -3) {  // whole groups of 8 slots (one per XCD); slots past the units stay idle
            const int most = ctx->cus * occ / S / 8 * 8;
            if (most < 8) throw ArgError{MSG_UNSUPPORTED, "split column wider than the device"};
            b.slots = std::min(most, (b.slots + 7) / 8 * 8);
        }
        const int vecs = b.G <= 0 && b.G != -3 ? a.m + 2 : a.m;  // streaming variants: + residual and state
        if (t_arena && b.G != -3 && !a.rk4) {  // as many resident columns as the scratch holds beside the rest of the run
            const size_t rest = (size_t)b.n_units * 8 + sizeof(StepArgs) + (size_t)n_items * (b.nobs + 1) * 8 +
                                (want_trace ? (size_t)n_items * ns * b.nobs * 8 : 0) + 4 * 256;
            const size_t per = (size_t)vecs * 2 * D * sizeof(float);
            const size_t room = t_arena->left() > rest ? t_arena->left() - rest : 0;
            if (room < per)
                throw ArgError{MSG_OOM_ADMISSION, "the run needs more than the " + std::to_string(t_arena->cap >> 20) +
                                                      " MiB of the allocation-free scratch"};
            b.slots = std::max(1, std::min<int>(b.slots, (int)(room / per)));
        }
        a.basis = a.rk4 ? nullptr : b.basis.alloc<float>((size_t)b.slots * vecs * 2 * D, stream);  // RK4: no basis
        if (b.G == -3) {
            a.xred = b.xred.alloc<double>((size_t)b.slots * 2 * S * 17, stream);
            a.bar = b.bar.alloc<int>(b.slots, stream);
        }
        a.norm0 = b.norm0.alloc<double>(b.n_units, stream);
        if (a.renorm) HIPCHECK(launch_init_norms(a, push_args(b, a), stream));
    } else if (!b.uneven) {
        const size_t lds = (size_t)4 * kDenseLdsPlanes * a.dpad * 16 + 4096;
        const int occ = std::max(1, dense_occupancy(a.dpad, lds));
        b.slots = std::max(1, std::min(b.n_units, ctx->cus * occ));
        if (b.share > 1) b.slots = std::max(1, std::min(b.slots, ctx->cus * occ / b.share));
    }
    b.footprint = b.coef.n + b.U.n + b.dts.n + b.wy.n + b.wt.n + b.thr.n + b.yacc.n + b.trace.n + b.state.n +
                  b.basis.n + b.given.n + b.norm0.n;
    b.s_next = 0;
}

// Copies the launch arguments to the device, ordered on the batch stream (a
// pageable source is staged before the call returns, so it may be reused).
const StepArgs *push_args(Batch &b, const StepArgs &a) {
    if (!b.args.p) b.args.alloc<StepArgs>(1, b.stream);
    b.args_host = a;
    HIPCHECK(hipMemcpyAsync(b.args.p, &b.args_host, sizeof(StepArgs), hipMemcpyHostToDevice, b.stream));
    return b.args.get<StepArgs>();
}

// Runs intervals [s_next, min(s_next + steps, nsamples-1)).
void run_chunk(Batch &b, int steps) {
    StepArgs a = b.a;
    const int last = b.nsamples - 1;
    a.s_begin = b.s_next;
    a.s_end = std::min(last, b.s_next + steps);
    if (b.s_next > 0) {
        if (!b.keep_state) throw DeviceError("chunked run without a state buffer");
        a.init_mode = 2;
    }
    HIPCHECK(launch_zero_int(a.queue, 1, b.stream));
    if (a.bar) HIPCHECK(launch_zero_int(a.bar, b.slots, b.stream));
    if (a.xred) HIPCHECK(launch_zero_int(reinterpret_cast<int *>(a.xred), (int)(b.xred.n / sizeof(int)), b.stream));
    const StepArgs *da = push_args(b, a);
    if (b.timed) {
        if (b.launches > 0) {  // the previous chunk's launch time, before the events are re-recorded
            HIPCHECK(hipEventSynchronize(b.ctx->ev1));
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, b.ctx->ev0, b.ctx->ev1));
            b.kernel_ms += ms;
        }
        HIPCHECK(hipEventRecord(b.ctx->ev0, b.stream));
    }
    if (b.path == Path::Lanczos) {
        const hipError_t e = launch_lanczos(a, da, b.G, b.T, b.slots, b.stream);
        if (e == hipErrorCooperativeLaunchTooLarge && b.G == -3)
            throw CoopRefused{"cooperative launch of the split-column kernel refused by the device"};
        HIPCHECK(e);
    } else {
        HIPCHECK(launch_dense(a, da, b.slots, b.stream));
    }
    if (b.timed) HIPCHECK(hipEventRecord(b.ctx->ev1, b.stream));
    ++b.launches;
    b.s_next = a.s_end;
}

// Device time of all propagation launches of a timed batch (call after collect()).
double batch_kernel_ms(Batch &b) {
    if (!b.timed || b.launches == 0) return 0.0;
    float ms = 0.f;
    HIPCHECK(hipEventSynchronize(b.ctx->ev1));
    HIPCHECK(hipEventElapsedTime(&ms, b.ctx->ev0, b.ctx->ev1));
    return b.kernel_ms + ms;
}

// Propagators of intervals [s0, s1) of a dynamic dense batch, into b.U with index
// item * (s1 - s0) + (s - s0): the midpoint exponential exp(dt G(t + dt/2)) (Taylor 18
// of dt G / 2^q, then q squarings; TaskHSGeneral.cpp:527-537) or, with RK4, the stage
// map I + (K1 + 2 K2 + 2 K3 + K4)/6 of G at t, t + dt/2, t + dt (HSPropagator.cpp:104-129),
// K1 = dt G0, K2 = dt Gm (1 + K1/2), K3 = dt Gm (1 + K2/2), K4 = dt G1 (1 + K3).
void build_dyn_chunk(Batch &b, int s0, int s1) {
    msg_ctx *ctx = b.ctx;
    const Resolved &r = b.sys->r;
    const int D = (int)r.D, Dp = b.a.dpad, nsc = s1 - s0;
    const size_t DD = (size_t)D * D;
    hipStream_t st = b.stream;
    const bool rk4 = b.prm.method == MSG_METHOD_RK4;
    const int nz = (int)b.dyn_terms.size(), stages = rk4 ? 3 : 1;
    float *dU = b.U.p ? b.U.get<float>() : b.U.alloc<float>((size_t)b.n_items * b.dyn_chunk * kDensePlanes * Dp * Dp, st);
    b.a.U = dU;  // the launch reads the chunk's propagators from here
    const int total = b.n_items * nsc;
    const int nbuf = rk4 ? 10 : 7;
    const int sub = (int)std::max<size_t>(1, std::min<size_t>(total, ((size_t)1 << 30) / (nbuf * DD * sizeof(cd))));
    DBuf bX, bX2, bX3, bX4, bB, bP, bQ, bc, bi, bd;
    cd *dX = bX.alloc<cd>((size_t)stages * sub * DD, st);
    cd *dX2 = bX2.alloc<cd>((size_t)sub * DD, st), *dX3 = bX3.alloc<cd>((size_t)sub * DD, st);
    cd *dX4 = bX4.alloc<cd>((size_t)sub * DD, st), *dB = bB.alloc<cd>((size_t)sub * DD, st);
    cd *dP = bP.alloc<cd>((size_t)sub * DD, st), *dQ = bQ.alloc<cd>((size_t)sub * DD, st);
    double *dcoef = bc.alloc<double>((size_t)stages * sub * std::max(nz, 1), st);
    int *ditem = bi.alloc<int>((size_t)stages * sub, st);
    double *ddt = bd.alloc<double>((size_t)stages * sub, st);
    std::vector<double> coef((size_t)stages * sub * std::max(nz, 1)), dts((size_t)stages * sub);
    std::vector<int> items((size_t)stages * sub);
    for (int q0 = 0; q0 < total; q0 += sub) {
        const int nb = std::min(sub, total - q0);
        double nrm = 0.0;
        std::fill(coef.begin(), coef.end(), 0.0);
        for (int q = 0; q < nb; ++q) {
            const int it = (q0 + q) / nsc, s = s0 + (q0 + q) % nsc;
            const double t0 = b.times[s], h = b.times[s + 1] - b.times[s];
            const double *f = b.field_given ? &b.item_fields[3 * (size_t)it] : nullptr;
            for (int g = 0; g < stages; ++g) {
                const size_t e = (size_t)g * nb + q;  // stage-major
                TimePoint tp;
                tp.t = rk4 ? t0 + 0.5 * h * g : t0 + 0.5 * h;
                double *c = &coef[e * std::max(nz, 1)];
                for (const Term1 &t : r.t1) {
                    if (!t.td) continue;
                    double v[3];
                    term1_value(t, f, &tp, v);
                    for (int k = 0; k < nz; ++k)
                        if (b.dyn_terms[k].first == t.s) c[k] += v[b.dyn_terms[k].second];
                }
                double bound = b.dyn_gnorm[it];
                for (int k = 0; k < nz; ++k) bound += std::fabs(c[k]) * b.dyn_znorm[k];
                nrm = std::max(nrm, h * bound);
                items[e] = it;
                dts[e] = h;
            }
        }
        int sq = 0;
        if (!rk4) {
            while (nrm > 0.5) { nrm *= 0.5; ++sq; }
            for (int q = 0; q < nb; ++q) dts[q] = std::ldexp(dts[q], -sq);
        }
        const size_t ne = (size_t)stages * nb;
        HIPCHECK(hipMemcpyAsync(dcoef, coef.data(), ne * std::max(nz, 1) * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(ditem, items.data(), ne * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(ddt, dts.data(), ne * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHECK(launch_dyn_generator(b.dynG.p, b.dynZ.p, dcoef, ditem, ddt, nz, D, (int)ne, dX, st));
        cd *res = nullptr;
        if (!rk4) {
            cd *p = dP, *qq = dQ;
            taylor_batch(ctx, dX, dX2, dX3, dX4, dB, p, qq, D, nb, false, st);
            for (int k = 0; k < sq; ++k) {
                HIPCHECK(zgemm(ctx, p, p, qq, D, nb, 1.0, st));
                std::swap(p, qq);
            }
            res = p;
        } else {
            const cd *a0 = dX, *am = dX + (size_t)nb * DD, *a1 = dX + (size_t)2 * nb * DD;
            HIPCHECK(zgemm(ctx, am, a0, dX2, D, nb, 0.5, st, false, am));   // K2 = am + am K1 / 2
            HIPCHECK(zgemm(ctx, am, dX2, dX3, D, nb, 0.5, st, false, am));  // K3 = am + am K2 / 2
            HIPCHECK(zgemm(ctx, a1, dX3, dX4, D, nb, 1.0, st, false, a1));  // K4 = a1 + a1 K3
            const double c1[4] = {1.0, 1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0};
            HIPCHECK(launch_zlincomb(a0, dX2, dX3, dB, D, nb, c1, st));            // 1 + K1/6 + K2/3 + K3/3
            const double c2[4] = {0.0, 1.0, 1.0 / 6.0, 0.0};
            HIPCHECK(launch_zlincomb(dB, dX4, dX4, dP, D, nb, c2, st));            // + K4/6
            res = dP;
        }
        HIPCHECK(launch_to_planes(res, dU + (size_t)q0 * kDensePlanes * Dp * Dp, D, Dp, (size_t)kDensePlanes * Dp * Dp,
                                  nb, st));
        HIPCHECK(hipStreamSynchronize(st));  // host staging reused by the next sub-batch
    }
}

// Runs `steps` intervals; a dynamic dense batch builds its propagators chunk by chunk.
void advance(Batch &b, int steps) {
    if (!b.dyn_dense || b.nsamples <= 1) {
        run_chunk(b, steps);
        return;
    }
    const int last = b.nsamples - 1;
    int left = steps;
    while (left > 0 && b.s_next < last) {
        const int s0 = b.s_next, s1 = std::min({last, s0 + b.dyn_chunk, s0 + left});
        build_dyn_chunk(b, s0, s1);
        b.a.ustride = s1 - s0;
        run_chunk(b, s1 - s0);
        left -= s1 - s0;
    }
}

// Copies results; per-item sums over columns in fixed order.
void collect(Batch &b, double *yields, double *trace) {
    const int cols_per_item = b.n_units / b.n_items;
    if (yields) {
        double *out = b.yout.alloc<double>((size_t)b.n_items * b.nobs, b.stream);
        HIPCHECK(launch_reduce_columns(b.a.yacc, out, b.n_items, cols_per_item, b.nobs, b.nobs * 2, 2, b.stream));
        HIPCHECK(hipMemcpyAsync(yields, out, (size_t)b.n_items * b.nobs * sizeof(double), hipMemcpyDeviceToHost, b.stream));
    }
    if (trace && b.a.trace) {
        const int nk = b.nsamples * b.nobs;
        double *out = b.tout.alloc<double>((size_t)b.n_items * nk, b.stream);
        HIPCHECK(launch_reduce_columns(b.a.trace, out, b.n_items, cols_per_item, nk, nk, 1, b.stream));
        HIPCHECK(hipMemcpyAsync(trace, out, (size_t)b.n_items * nk * sizeof(double), hipMemcpyDeviceToHost, b.stream));
    }
    int st = 0;
    HIPCHECK(hipMemcpyAsync(&st, b.a.status, sizeof(int), hipMemcpyDeviceToHost, b.stream));
    HIPCHECK(hipStreamSynchronize(b.stream));
    if (st & 1) throw ArgError{MSG_NUMERIC_BREAKDOWN, "non-finite state during propagation"};
    if (st & 2) throw ArgError{MSG_NUMERIC_BREAKDOWN, "time step too large for the Krylov exponential"};
    if (st & 4) throw ArgError{MSG_DEVICE_ERROR, "the workgroups of a split column lost synchronisation"};
    if (yields && b.prm.calculation == MSG_CALC_YIELDS && b.prm.yield_corrections) {
        const double den = 1.0 - std::exp(-b.prm.total_time * b.sys->r.kmax);
        if (!(den > 0.0) || !std::isfinite(den))
            throw ArgError{MSG_INVALID_ARG, "yieldcorrections=true needs a positive total time and rate"};
        for (size_t i = 0; i < (size_t)b.n_items * b.nobs; ++i) yields[i] /= den;
    }
}

template <class F>
msg_status guarded(msg_ctx *ctx, F f) {
    try {
        f();
        if (ctx) ctx->err.clear();
        return MSG_OK;
    } catch (const ArgError &e) {
        return fail(ctx, e.st, e.what);
    } catch (const DeviceError &e) {
        return fail(ctx, MSG_DEVICE_ERROR, e.what);
    } catch (const CoopRefused &e) {
        return fail(ctx, MSG_DEVICE_ERROR, e.what);
    } catch (const std::bad_alloc &) {
        return fail(ctx, MSG_OOM_ADMISSION, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(ctx, MSG_DEVICE_ERROR, e.what());
    }
}

// Class and packing of an entry in the uneven launch (the table of k_uneven).
struct UnevenShape {
    int cls = 0, per_pack = 1;
    size_t basis_floats = 0;  // per workgroup
};

UnevenShape uneven_shape(const Batch &b) {
    UnevenShape s;
    const int D = b.a.dim;
    if (b.path == Path::Lanczos) {
        const int ng = D / 4;
        int lg = 0;
        while ((1 << lg) < ng) ++lg;
        if ((1 << lg) != ng || lg < 6 || lg > 12)
            throw ArgError{MSG_UNSUPPORTED, "uneven batch: matrix-free entries need 256 <= D <= 16384"};
        if (lg == 12 && b.a.npairs > 0)
            throw ArgError{MSG_UNSUPPORTED, "uneven batch: nucleus-nucleus couplings need D <= 8192"};
        s.cls = UC_LZ64 + (lg - 6);
        s.per_pack = uneven_pack_width(s.cls);
        s.basis_floats = s.cls == UC_LZ4096 ? (size_t)(b.a.m + 2) * 2 * D : (size_t)s.per_pack * b.a.m * 2 * D;
    } else {
        static const int nts[10] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32};
        const int nt = b.a.dpad / 16;
        const int i = (int)(std::find(nts, nts + 10, nt) - nts);
        if (i == 10) throw ArgError{MSG_UNSUPPORTED, "uneven batch: no dense class for this dimension"};
        s.cls = UC_DENSE + i;
        s.per_pack = uneven_pack_width(s.cls);
    }
    return s;
}

// One uneven launch in flight: the entries' batches, the launch's own buffers and
// the host copies its asynchronous uploads read.
struct UnevenRun {
    std::vector<std::unique_ptr<Batch>> bs;
    std::vector<StepArgs> ents;
    std::vector<UnevenPack> packs;
    UnevenArgs ua{};
    DBuf dents, dpacks, dqueue, dbasis, dargs;
};

// Whether a request of `sys` under `p` can go into an uneven launch (the checks of
// uneven_begin made before anything is allocated).
bool uneven_eligible(const msg_system *sys, const msg_run_params *p) {
    const Resolved &r = sys->r;
    if (p->method == MSG_METHOD_UNCOUPLED || p->calculation != MSG_CALC_YIELDS) return false;
    if (p->dynamics && r.has_td) return false;
    Path path;
    try {
        path = choose_path(r, p);
    } catch (const ArgError &) {
        return false;
    }
    if (path == Path::Dense) return r.D <= 512;
    const int64_t ng = r.D / 4;
    return sym_for(r, p) && p->method != MSG_METHOD_RK4 && !(p->krylov_tol > 0.0) && r.mix_mult == 0 && ng >= 64 &&
           ng <= (r.nn_pairs.empty() ? 4096 : 2048) && (ng & (ng - 1)) == 0;
}

// Prepares every entry on stream st and queues one persistent k_uneven launch over all
// of their (item, column) units, packed per class and ordered largest first.  Nothing
// is synchronised: the results are valid once the stream has passed the launch
// (uneven_finish).  `timed`: HIP events of the context around the launch.
void uneven_begin(UnevenRun &run, msg_ctx *ctx, int n_entries, const msg_uneven_entry *entries, hipStream_t st,
                  bool timed) {
    run.bs.resize(n_entries);
    run.ents.resize(n_entries);
    struct Ordered {
        UnevenPack p;
        double cost;
    };
    std::vector<Ordered> order;
    size_t basis_floats = 0;
    for (int e = 0; e < n_entries; ++e) {
        const msg_uneven_entry &en = entries[e];
        const std::string at = "entry " + std::to_string(e) + ": ";
        if (!en.system || !en.params || !en.yields) throw ArgError{MSG_INVALID_ARG, at + "null system, parameters or yields"};
        if (en.params->method == MSG_METHOD_UNCOUPLED)
            throw ArgError{MSG_UNSUPPORTED, at + "propagationmethod=uncoupled runs through msg_propagate"};
        if (en.params->dynamics && en.system->r.has_td)
            throw ArgError{MSG_UNSUPPORTED, at + "the uneven batch takes static Hamiltonians"};
        run.bs[e].reset(new Batch());
        Batch &b = *run.bs[e];
        b.uneven = true;
        try {
            prepare(b, ctx, en.system, en.params, en.n_items, en.fields, en.columns, en.n_columns, en.trace != nullptr,
                    false, st);
        } catch (const ArgError &x) {
            throw ArgError{x.st, at + x.what};
        }
        if (b.arnoldi)
            throw ArgError{MSG_UNSUPPORTED, at + "unequal recombination rates above D = 512 run through msg_propagate"};
        if (b.a.ktol > 0.0) throw ArgError{MSG_UNSUPPORTED, at + "krylovtol basis reuse runs through msg_propagate"};
        if (b.a.rk4) throw ArgError{MSG_UNSUPPORTED, at + "matrix-free RK4 runs through msg_propagate"};
        if (b.a.mixm)
            throw ArgError{MSG_UNSUPPORTED, at + "a nucleus with spin > 1/2 above D = 512 runs through msg_propagate"};
        const UnevenShape sh = uneven_shape(b);
        basis_floats = std::max(basis_floats, sh.basis_floats);
        StepArgs a = b.a;
        a.s_begin = 0;
        a.s_end = b.nsamples - 1;
        run.ents[e] = a;
        // packs of consecutive units of one item, cost ~ device time of the pack
        const int upi = b.n_units / b.n_items;
        const double D = a.dim, steps = b.nsamples - 1;
        for (int it = 0; it < b.n_items; ++it)
            for (int c = 0; c < upi; c += sh.per_pack) {
                const int nact = std::min(sh.per_pack, upi - c);
                const double cost = b.path == Path::Lanczos ? nact * D * steps * a.m * (a.nnuc + 4)
                                                            : nact * 16.0 * a.dpad * a.dpad * steps / 16.0;
                order.push_back({UnevenPack{e, it * upi + c, nact, sh.cls}, cost});
            }
    }
    std::stable_sort(order.begin(), order.end(), [](const Ordered &x, const Ordered &y) { return x.cost > y.cost; });
    run.packs.resize(order.size());
    for (size_t k = 0; k < order.size(); ++k) run.packs[k] = order[k].p;
    const int n_packs = (int)run.packs.size();
    const int occ = std::max(1, uneven_occupancy());
    const int wgs = std::max(1, std::min(n_packs, ctx->cus * occ));
    StepArgs *de = run.dents.alloc<StepArgs>(run.ents.size(), st);
    HIPCHECK(hipMemcpyAsync(de, run.ents.data(), run.ents.size() * sizeof(StepArgs), hipMemcpyHostToDevice, st));
    UnevenPack *dp = run.dpacks.alloc<UnevenPack>(run.packs.size(), st);
    HIPCHECK(hipMemcpyAsync(dp, run.packs.data(), run.packs.size() * sizeof(UnevenPack), hipMemcpyHostToDevice, st));
    UnevenArgs &ua = run.ua;
    ua.ents = de;
    ua.packs = dp;
    ua.n_packs = n_packs;
    ua.queue = run.dqueue.alloc<int>(1, st);
    HIPCHECK(launch_zero_int(ua.queue, 1, st));
    ua.basis_stride = (basis_floats + 63) / 64 * 64;
    ua.basis = ua.basis_stride ? run.dbasis.alloc<float>((size_t)wgs * ua.basis_stride, st) : nullptr;
    UnevenArgs *du = run.dargs.alloc<UnevenArgs>(1, st);
    HIPCHECK(hipMemcpyAsync(du, &ua, sizeof(UnevenArgs), hipMemcpyHostToDevice, st));
    if (timed) HIPCHECK(hipEventRecord(ctx->ev0, st));
    HIPCHECK(launch_uneven(du, wgs, st));
    if (timed) HIPCHECK(hipEventRecord(ctx->ev1, st));
}

// Status and results of a launch the stream has passed (uneven_begin).
void uneven_finish(UnevenRun &run, const msg_uneven_entry *entries) {
    const int n = (int)run.bs.size();
    int bad_class = 0;
    for (int e = 0; e < n; ++e) {
        int sflag = 0;
        HIPCHECK(hipMemcpy(&sflag, run.bs[e]->a.status, sizeof(int), hipMemcpyDeviceToHost));
        bad_class |= sflag & 8;
    }
    if (bad_class) throw DeviceError("uneven launch: a pack named no class or did not fit its class");
    for (int e = 0; e < n; ++e) {
        try {
            collect(*run.bs[e], entries[e].yields, entries[e].trace);
        } catch (const ArgError &x) {
            throw ArgError{x.st, "entry " + std::to_string(e) + ": " + x.what};
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Server (continuous batching of yield requests)
// ---------------------------------------------------------------------------
struct Request {
    int64_t id = 0;
    const msg_system *sys = nullptr;
    msg_run_params prm{};
    std::vector<double> fields;
    int n_fields = 0;
    std::unique_ptr<Batch> batch;
    std::vector<double> result;
    bool failed = false;
    std::string error;
    int64_t bytes = 0;
    int stream = -1;             // index of the server stream the request owns
    bool whole = false;          // launched to completion in persistent launches of field chunks
    hipEvent_t done_ev = nullptr;
    double work = 0.0;           // state-steps x D: admission order (shortest first)
    int f_begin = 0, f_count = 0;  // fields of the launch in flight (whole mode)
    int nobs = 0;
    bool no_group = false;       // admitted on its own (its shared launch was refused)
};

// Requests that share one uneven launch (whole mode): small requests of different
// systems and dimensions go into one persistent launch together instead of one
// launch each, which alone could not fill the device.
struct RequestGroup {
    std::vector<std::unique_ptr<Request>> members;
    std::vector<msg_uneven_entry> ents;
    std::unique_ptr<UnevenRun> run;
    hipEvent_t done = nullptr;
    int stream = -1;
    int64_t bytes = 0;
};

// Work (D x columns x samples) of one whole-mode launch of the server: about 50 ms of
// device time at D = 16384 (one field of 4096 columns over 6 samples), so one large
// request holds the device for at most that long before smaller requests' launches,
// queued on the other streams, get their turn.
constexpr double kServerChunkWork = 4.0e8;
// Work of one shared launch of small requests (about 50 ms of device time, as one
// whole-mode launch): queued requests below it that the uneven launch takes are
// admitted together, smallest first.  On the bench's 40-request stream larger groups
// raised the median latency (0.95 s at 1e9, 1.5 s at 4e9 against 0.71 s here and
// 0.73 s without groups) at the same throughput.
constexpr double kServerGroupWork = 4.0e8;

struct msg_server {
    msg_ctx *ctx = nullptr;
    int max_slots = 0;
    int64_t budget = 0, in_use = 0;
    int64_t next_id = 1;
    std::deque<std::unique_ptr<Request>> queued;
    std::vector<std::unique_ptr<Request>> active;
    std::map<int64_t, std::unique_ptr<Request>> finished;
    std::vector<std::unique_ptr<RequestGroup>> groups;
    int64_t shared_launches = 0, shared_requests = 0;
    double group_work = kServerGroupWork;  // MSG_SERVER_GROUP_WORK overrides (tuning runs; 0 disables)
    std::vector<hipStream_t> streams;
    std::vector<char> stream_busy;
};

extern "C" {

const char *msg_last_error(const msg_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

msg_status msg_ctx_create(int device, msg_ctx **out) {
    if (!out) return fail(nullptr, MSG_INVALID_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return fail(nullptr, MSG_DEVICE_ERROR, "no HIP device");
    if (device < 0 || device >= n) return fail(nullptr, MSG_INVALID_ARG, "device index out of range");
    std::unique_ptr<msg_ctx> c(new (std::nothrow) msg_ctx());
    if (!c) return fail(nullptr, MSG_OOM_ADMISSION, "host allocation failed");
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, MSG_DEVICE_ERROR, "hipSetDevice failed");
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(nullptr, MSG_DEVICE_ERROR, "hipStreamCreate failed");
    if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c->cus < 1)
        c->cus = 1;
    if (set_lanczos_lds_limits() != hipSuccess) return fail(nullptr, MSG_DEVICE_ERROR, "cannot raise the LDS limit");
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return fail(nullptr, MSG_DEVICE_ERROR, "hipEventCreate failed");
    if (scratch_address(device, &c->scratch) != hipSuccess)
        return fail(nullptr, MSG_DEVICE_ERROR, "cannot locate the allocation-free scratch on this device");
    *out = c.release();
    return MSG_OK;
}

msg_status msg_ctx_destroy(msg_ctx *ctx) {
    if (!ctx) return MSG_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    // finish the context's work first; the rocBLAS handle is bound to the context's
    // stream, so it goes before the stream does
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    delete ctx;
    return MSG_OK;
}

msg_status msg_system_upload(msg_ctx *ctx, const msg_system_desc *desc, msg_system **out) {
    if (!ctx || !desc || !out) return fail(ctx, MSG_INVALID_ARG, "null argument");
    return guarded(ctx, [&] {
        std::unique_ptr<msg_system> s(new msg_system());
        s->ctx = ctx;
        const std::string e = resolve_system(desc, s->r);
        if (!e.empty()) throw ArgError{MSG_INVALID_ARG, e};
        *out = s.release();
    });
}

msg_status msg_system_release(msg_ctx *ctx, msg_system *sys) {
    if (!sys) return fail(ctx, MSG_INVALID_ARG, "null system");
    delete sys;
    return MSG_OK;
}

msg_status msg_system_dimension(const msg_system *sys, int64_t *dim) {
    if (!sys || !dim) return MSG_INVALID_ARG;
    *dim = sys->r.D;
    return MSG_OK;
}

msg_status msg_direct_columns(const msg_system *sys, int32_t *n) {
    if (!sys || !n) return MSG_INVALID_ARG;
    *n = (int32_t)(sys->r.zdim * (int64_t)sys->r.evec.size());
    return MSG_OK;
}

msg_status msg_trace_samples(const msg_system *sys, int32_t state, int32_t method, uint64_t seed, int32_t m,
                             double *columns) {
    if (!sys || !columns || m < 1 || (method != MSG_TRACE_COHERENT && method != MSG_TRACE_SUZ))
        return fail(sys ? sys->ctx : nullptr, MSG_INVALID_ARG, "bad trace-sample arguments");
    try {
        const std::string e = trace_samples(sys->r, state, method, seed, m, columns);
        if (!e.empty()) return fail(sys->ctx, MSG_INVALID_ARG, e);
    } catch (const std::bad_alloc &) {
        return fail(sys->ctx, MSG_OOM_ADMISSION, "host allocation failed");
    }
    return MSG_OK;
}

msg_status msg_observable_count(const msg_system *sys, const msg_run_params *p, int32_t *n) {
    if (!sys || !p || !n) return MSG_INVALID_ARG;
    *n = observable_count(sys->r, p);
    return MSG_OK;
}

msg_status msg_time_grid(double total, double dt, int64_t *n, double *times) {
    if (!n) return MSG_INVALID_ARG;
    std::vector<double> t;
    if (!free_times(total, dt, t).empty()) return MSG_INVALID_ARG;
    *n = (int64_t)t.size();
    if (times) std::copy(t.begin(), t.end(), times);
    return MSG_OK;
}

msg_status msg_query_footprint(const msg_system *sys, const msg_run_params *p, int32_t n_items, int32_t n_columns,
                               int64_t *bytes) {
    if (!sys || !p || !bytes || n_items < 1) return MSG_INVALID_ARG;
    const Resolved &r = sys->r;
    std::vector<double> t;
    if (!free_times(p->total_time, p->time_step, t).empty()) return MSG_INVALID_ARG;
    const int ncols = n_columns > 0 ? n_columns : (int)(r.zdim * (int64_t)std::max<size_t>(1, r.evec.size()));
    const int nobs = std::max(1, observable_count(r, p));
    const int64_t D = r.D;
    int64_t b = (int64_t)t.size() * nobs * 16 + (int64_t)t.size() * 12;
    if (p->method == MSG_METHOD_UNCOUPLED) {  // per radical: 5 n^2 FP64 complex per item of a chunk + correlations
        std::vector<int> SA, SB;
        if (!uncoupled_split(r, SA, SB).empty()) return MSG_UNSUPPORTED;
        int64_t nA = 1, nB = 1;
        for (int sp : SA) nA *= r.mult[sp];
        for (int sp : SB) nB *= r.mult[sp];
        const int64_t nmax = std::max(nA, nB);
        const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n_items, ((int64_t)512 << 20) / (5 * nmax * nmax * 16)));
        b += chunk * 5 * (nA * nA + nB * nB) * 16 + chunk * 2 * (int64_t)t.size() * 9 * 8 +
             (int64_t)n_items * nobs * 8;
        *bytes = b;
        return MSG_OK;
    }
    Path path;
    try {
        path = choose_path(r, p);
    } catch (...) {
        return MSG_UNSUPPORTED;
    }
    const int m = p->krylov_size > 0 ? p->krylov_size : 16;
    if (path == Path::Lanczos) {
        const int64_t units = (int64_t)n_items * ncols;
        b += (int64_t)n_items * (kCoefHead + (int64_t)r.nuc.size() * kNucStride + (int64_t)r.nn_pairs.size() * kPairStride) * 4;
        b += units * (nobs * 16 + 8);
        const int vecs = D > 8192 ? m + 2 : m;  // streaming variants keep residual and state beside the basis
        b += std::min<int64_t>(units, 2048) * vecs * 8 * D;  // Krylov bases of the resident workgroups
    } else {
        const int64_t Dp = dense_padding((int)D);
        const int64_t units = (int64_t)n_items * ((ncols + 15) / 16);
        b += (int64_t)n_items * 2 * 4 * kDensePlanes * Dp * Dp;
        b += units * nobs * 16;
    }
    if (p->sampling == MSG_SAMPLING_GIVEN) b += (int64_t)ncols * 8 * D;
    *bytes = b;
    return MSG_OK;
}

namespace {

// Pauli coefficients Tr[X (sigma_mu (x) sigma_nu)] of a 4x4 electron-pair matrix
// (pair index 2a + b, electron 1 first; sigma_0 = 1, x, y, z), real for Hermitian X.
void pauli_coefficients(const Mat4 &X, double out[16]) {
    const cd I(0.0, 1.0);
    const cd sg[4][4] = {{1.0, 0.0, 0.0, 1.0}, {0.0, 1.0, 1.0, 0.0}, {0.0, -I, I, 0.0}, {1.0, 0.0, 0.0, -1.0}};
    for (int m = 0; m < 4; ++m)
        for (int n = 0; n < 4; ++n) {
            cd tr = 0.0;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    // (sigma_m (x) sigma_n)[j][i] with j = 2 a + b, i = 2 a' + b'
                    const cd e = sg[m][(j >> 1) * 2 + (i >> 1)] * sg[n][(j & 1) * 2 + (i & 1)];
                    tr += X[i * 4 + j] * e;
                }
            out[m * 4 + n] = tr.real();
        }
}

// msg_propagate for MSG_METHOD_UNCOUPLED.
void propagate_uncoupled(msg_ctx *ctx, const msg_system *sys, const msg_run_params *p, int n_items,
                         const double *fields, double *yields, double *trace) {
    const Resolved &r = sys->r;
    check_params(sys, p);
    std::vector<int> SA, SB;
    const std::string why = uncoupled_split(r, SA, SB);
    if (!why.empty()) throw ArgError{MSG_UNSUPPORTED, "uncoupled propagation: " + why};
    if (p->sampling != MSG_SAMPLING_DIRECT) throw ArgError{MSG_UNSUPPORTED, "uncoupled propagation uses direct sampling"};
    if (r.relax) throw ArgError{MSG_UNSUPPORTED, "relaxation couples the radicals"};
    if (p->dynamics && r.has_td) throw ArgError{MSG_UNSUPPORTED, "uncoupled propagation takes a static Hamiltonian"};
    if (!sym_for(r, p)) throw ArgError{MSG_UNSUPPORTED, "uncoupled propagation needs symmetric recombination"};
    if (r.evec.empty()) throw ArgError{MSG_INVALID_ARG, "the system defines no initial state"};
    int nA = 1, nB = 1;
    for (int sp : SA) nA *= r.mult[sp];
    for (int sp : SB) nB *= r.mult[sp];
    if (nA > 512 || nB > 512) throw ArgError{MSG_UNSUPPORTED, "uncoupled propagation: a radical above dimension 512"};
    const int nobs = observable_count(r, p);
    if (nobs < 1) throw ArgError{MSG_INVALID_ARG, "no observables (no transitions / States)"};
    if (nobs > kObsMax) throw ArgError{MSG_UNSUPPORTED, "more than 8 observables"};
    // time grid and weights, as for the other paths (decay factored out)
    const bool legacy = p->task == MSG_TASK_LEGACY_DIRECT_YIELDS || p->task == MSG_TASK_LEGACY_STOCH_YIELDS;
    if (!(p->task == MSG_TASK_HSGENERAL || legacy)) throw ArgError{MSG_INVALID_ARG, "unknown task"};
    std::vector<double> times;
    const std::string te = legacy ? legacy_times(p->total_time, p->time_step, times) : free_times(p->total_time, p->time_step, times);
    if (!te.empty()) throw ArgError{MSG_INVALID_ARG, te};
    const int ns = (int)times.size();
    const double kappa = kappa_for(r, p);
    const std::vector<Mat4> &P = p->observables == MSG_OBS_TRANSITION_YIELDS ? r.P_trans : r.P_states;
    std::vector<double> scale(nobs, 1.0);
    if (p->observables == MSG_OBS_TRANSITION_YIELDS) scale = r.rates;
    std::vector<double> wyh((size_t)ns * nobs), wth((size_t)ns * nobs);
    for (int s = 0; s < ns; ++s) {
        double w = 0.0;
        if (s > 0) w += 0.5 * (times[s] - times[s - 1]);
        if (s + 1 < ns) w += 0.5 * (times[s + 1] - times[s]);
        const double decay = std::exp(-2.0 * kappa * times[s]);
        for (int c = 0; c < nobs; ++c) {
            wyh[(size_t)s * nobs + c] = w * decay * scale[c];
            wth[(size_t)s * nobs + c] = decay * scale[c];
        }
    }
    // Q_c[mu alpha][nu beta] = p^c_{mu nu} r_{alpha beta} / (16 Z_A Z_B), rho_e = sum_i evec_i evec_i^dagger
    Mat4 rho{};
    for (const auto &v : r.evec)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) rho[i * 4 + j] += v[i] * std::conj(v[j]);
    double rc[16];
    pauli_coefficients(rho, rc);
    const double norm = 1.0 / (16.0 * (nA / 2) * (nB / 2));
    std::vector<double> Q((size_t)nobs * 256);
    for (int c = 0; c < nobs; ++c) {
        double pc[16];
        pauli_coefficients(P[c], pc);
        for (int mu = 0; mu < 4; ++mu)
            for (int al = 0; al < 4; ++al)
                for (int nu = 0; nu < 4; ++nu)
                    for (int be = 0; be < 4; ++be)
                        Q[(size_t)c * 256 + (mu * 4 + al) * 16 + nu * 4 + be] = pc[mu * 4 + nu] * rc[al * 4 + be] * norm;
    }
    hipStream_t st = ctx->stream;
    (void)blas_on(ctx, st);
    DBuf dQ, dwy, dwt, dtimes, dy, dtr;
    double *q = dQ.alloc<double>(Q.size(), st);
    double *wy = dwy.alloc<double>(wyh.size(), st);
    double *wt = dwt.alloc<double>(wth.size(), st);
    double *tm = dtimes.alloc<double>(times.size(), st);
    HIPCHECK(hipMemcpyAsync(q, Q.data(), Q.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(wy, wyh.data(), wyh.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(wt, wth.data(), wth.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(tm, times.data(), times.size() * sizeof(double), hipMemcpyHostToDevice, st));
    double *yo = dy.alloc<double>((size_t)n_items * nobs, st);
    double *to = trace ? dtr.alloc<double>((size_t)n_items * ns * nobs, st) : nullptr;
    // items in chunks bounding the FP64 workspace (5 n^2 complex per item and radical)
    const int nmax = std::max(nA, nB);
    const int chunk = std::max(1, std::min(n_items, (int)(((size_t)512 << 20) / ((size_t)5 * nmax * nmax * 16))));
    DBuf hA, hB, yA, yB, sA, sB, lA, lB, eA, eB, iA, iB, cA, cB;
    cd *dHA = hA.alloc<cd>((size_t)chunk * nA * nA, st), *dHB = hB.alloc<cd>((size_t)chunk * nB * nB, st);
    cd *dYA = yA.alloc<cd>((size_t)chunk * nA * nA, st), *dYB = yB.alloc<cd>((size_t)chunk * nB * nB, st);
    cd *dSA = sA.alloc<cd>((size_t)chunk * 3 * nA * nA, st), *dSB = sB.alloc<cd>((size_t)chunk * 3 * nB * nB, st);
    double *dlA = lA.alloc<double>((size_t)chunk * nA, st), *dlB = lB.alloc<double>((size_t)chunk * nB, st);
    double *deA = eA.alloc<double>((size_t)chunk * nA, st), *deB = eB.alloc<double>((size_t)chunk * nB, st);
    rocblas_int *diA = iA.alloc<rocblas_int>(n_items, st), *diB = iB.alloc<rocblas_int>(n_items, st);
    // k_corr splits each radical's pairs into kCorrParts ranges (more workgroups in
    // flight); the partial correlations are summed in order by k_unc_combine
    const size_t pstride = (size_t)chunk * ns * 9;
    double *dCA = cA.alloc<double>(kCorrParts * pstride, st), *dCB = cB.alloc<double>(kCorrParts * pstride, st);
    // The radicals' Hamiltonians are linear in the swept field, H(f) = H(0) + sum_a f_a Z_a:
    // the host builds H(0) and Z_a = H(e_a) - H(0) once per radical, and the device forms
    // every item's H(f) (k_dyn_generator); without fields every item has the system's own
    // field, H = H(that field).
    const int nz = fields ? 3 : 0;
    DBuf h0A, h0B, zA, zB, dfl, ditem, done;
    auto unit_parts = [&](const std::vector<int> &spins, int n, DBuf &h0, DBuf &z) {
        const size_t nn = (size_t)n * n;
        std::vector<cd> H0, H1, Z((size_t)3 * nn);
        const double zero[3] = {0.0, 0.0, 0.0};
        subsystem_hamiltonian(r, spins, fields ? zero : nullptr, H0);
        for (int ax = 0; ax < nz; ++ax) {
            double e[3] = {0.0, 0.0, 0.0};
            e[ax] = 1.0;
            subsystem_hamiltonian(r, spins, e, H1);
            for (size_t k = 0; k < nn; ++k) Z[(size_t)ax * nn + k] = H1[k] - H0[k];
        }
        HIPCHECK(hipMemcpyAsync(h0.alloc<cd>(nn, st), H0.data(), nn * sizeof(cd), hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(z.alloc<cd>(3 * nn, st), Z.data(), 3 * nn * sizeof(cd), hipMemcpyHostToDevice, st));
        HIPCHECK(hipStreamSynchronize(st));  // H0 and Z are released at scope end
    };
    unit_parts(SA, nA, h0A, zA);
    unit_parts(SB, nB, h0B, zB);
    {
        std::vector<double> fh((size_t)3 * n_items, 0.0), one(chunk, 1.0);
        if (fields) std::copy(fields, fields + (size_t)3 * n_items, fh.begin());
        std::vector<int> zeros(chunk, 0);
        HIPCHECK(hipMemcpyAsync(dfl.alloc<double>(fh.size(), st), fh.data(), fh.size() * sizeof(double),
                                hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(done.alloc<double>(chunk, st), one.data(), chunk * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(ditem.alloc<int>(chunk, st), zeros.data(), chunk * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    std::vector<rocblas_int> info(2 * (size_t)n_items);
    HIPCHECK(hipEventRecord(ctx->ev0, st));
    for (int it0 = 0; it0 < n_items; it0 += chunk) {
        const int nb = std::min(chunk, n_items - it0);
        const double *fc = dfl.get<double>() + (size_t)3 * it0;
        HIPCHECK(launch_dyn_generator(h0A.p, zA.p, fc, ditem.get<int>(), done.get<double>(), nz, nA, nb, dHA, st));
        HIPCHECK(launch_dyn_generator(h0B.p, zB.p, fc, ditem.get<int>(), done.get<double>(), nz, nB, nb, dHB, st));
        // eigenpairs (the row-major view of the solver's output is M = V^dagger, as in
        // msg_symmetric_decay), then S_mu = M (sigma_mu (x) 1) M^dagger and the correlations
        struct Rad { int n; cd *Hm, *Y, *S; double *lam, *off, *C; rocblas_int *inf; };
        for (const Rad &rd : {Rad{nA, dHA, dYA, dSA, dlA, deA, dCA, diA + it0}, Rad{nB, dHB, dYB, dSB, dlB, deB, dCB, diB + it0}}) {
            if (rocsolver_zheevd_strided_batched(ctx->blas, rocblas_evect_original, rocblas_fill_upper, rd.n,
                                                 reinterpret_cast<rocblas_double_complex *>(rd.Hm), rd.n,
                                                 (rocblas_stride)rd.n * rd.n, rd.lam, rd.n, rd.off, rd.n, rd.inf,
                                                 nb) != rocblas_status_success)
                throw ArgError{MSG_DEVICE_ERROR, "rocsolver_zheevd_strided_batched failed"};
            const size_t plane = (size_t)nb * rd.n * rd.n;  // S_mu of all items of the chunk
            for (int mu = 0; mu < 3; ++mu) {
                HIPCHECK(launch_pauli_right(rd.Hm, rd.Y, rd.n, mu, nb, st));
                HIPCHECK(zgemm(ctx, rd.Y, rd.Hm, rd.S + mu * plane, rd.n, nb, 1.0, st, true));
            }
            HIPCHECK(launch_corr(rd.S, plane, rd.lam, tm, rd.n, ns, nb, kCorrParts, pstride, rd.C, st));
        }
        HIPCHECK(launch_unc_combine(dCA, dCB, kCorrParts, pstride, q, wy, wt, nb, ns, nobs, (double)nA, (double)nB,
                                    yo + (size_t)it0 * nobs, to ? to + (size_t)it0 * ns * nobs : nullptr, st));
    }
    HIPCHECK(hipEventRecord(ctx->ev1, st));
    HIPCHECK(hipMemcpyAsync(info.data(), diA, n_items * sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(info.data() + n_items, diB, n_items * sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    for (int q0 = 0; q0 < 2 * n_items; ++q0)
        if (info[q0] != 0) throw ArgError{MSG_NUMERIC_BREAKDOWN, "eigensolver did not converge"};
    if (yields)
        HIPCHECK(hipMemcpyAsync(yields, yo, (size_t)n_items * nobs * sizeof(double), hipMemcpyDeviceToHost, st));
    if (trace)
        HIPCHECK(hipMemcpyAsync(trace, to, (size_t)n_items * ns * nobs * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (yields && p->calculation == MSG_CALC_YIELDS && p->yield_corrections) {
        const double den = 1.0 - std::exp(-p->total_time * r.kmax);
        if (!(den > 0.0) || !std::isfinite(den))
            throw ArgError{MSG_INVALID_ARG, "yieldcorrections=true needs a positive total time and rate"};
        for (size_t i = 0; i < (size_t)n_items * nobs; ++i) yields[i] /= den;
    }
}

}  // namespace

msg_status msg_propagate(msg_ctx *ctx, const msg_system *sys, const msg_run_params *params, int32_t n_items,
                         const double *fields, const double *columns, int32_t n_columns, double *yields,
                         double *trace) {
    if (!ctx || !sys) return fail(ctx, MSG_INVALID_ARG, "null context or system");
    return guarded(ctx, [&] {
        HIPCHECK(hipSetDevice(ctx->device));
        if (params && params->method == MSG_METHOD_UNCOUPLED) {
            if (n_items < 1) throw ArgError{MSG_INVALID_ARG, "n_items must be positive"};
            propagate_uncoupled(ctx, sys, params, n_items, fields, yields, trace);
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            ctx->kernel_ms = ms;
            ctx->launches = 1;
            return;
        }
        Batch b;
        prepare(b, ctx, sys, params, n_items, fields, columns, n_columns, trace != nullptr, false, ctx->stream);
        b.timed = true;
        advance(b, b.nsamples);
        collect(b, yields, trace);
        ctx->kernel_ms = batch_kernel_ms(b);
        ctx->launches = b.launches;
    });
}

msg_status msg_propagate_noalloc(msg_ctx *ctx, const msg_system *sys, const msg_run_params *params, int32_t n_items,
                                 const double *fields, const double *columns, int32_t n_columns, double *yields,
                                 double *trace) {
    if (!ctx || !sys) return fail(ctx, MSG_INVALID_ARG, "null context or system");
    return guarded(ctx, [&] {
        HIPCHECK(hipSetDevice(ctx->device));
        if (!ctx->scratch) throw DeviceError("no allocation-free scratch on this device");
        check_params(sys, params);
        if (params->method == MSG_METHOD_UNCOUPLED || choose_path(sys->r, params) != Path::Lanczos)
            throw ArgError{MSG_UNSUPPORTED, "the allocation-free entry point runs the matrix-free path only (the "
                                            "dense and uncoupled paths build their operators with rocBLAS / "
                                            "rocSOLVER, which allocate)"};
        Arena arena{static_cast<char *>(ctx->scratch), (size_t)kScratchBytes, 0};
        ArenaScope scope(&arena);
        Batch b;
        b.no_split = true;  // the split-column shape needs its cooperative launch's own sizing
        prepare(b, ctx, sys, params, n_items, fields, columns, n_columns, trace != nullptr, false, ctx->stream);
        b.timed = true;
        advance(b, b.nsamples);
        collect(b, yields, trace);
        ctx->kernel_ms = batch_kernel_ms(b);
        ctx->launches = b.launches;
    });
}

// Relative margin by which a later autotune candidate must beat the best earlier one.
constexpr double kTuneMargin = 0.015;

msg_status msg_autotune(msg_ctx *ctx, const msg_system *sys, const msg_run_params *params, int32_t n_items,
                        const double *fields, const double *columns, int32_t n_columns, int32_t sample_steps,
                        msg_launch_shape *best) {
    if (!ctx || !sys || !params || n_items < 1 || sample_steps < 1)
        return fail(ctx, MSG_INVALID_ARG, "null argument, no items or no sample steps");
    return guarded(ctx, [&] {
        HIPCHECK(hipSetDevice(ctx->device));
        const Resolved &r = sys->r;
        msg_run_params p = *params;
        p.calculation = MSG_CALC_YIELDS;
        p.yield_corrections = 0;
        p.total_time = sample_steps * p.time_step;
        const Path path = choose_path(r, &p);
        msg_launch_shape out{};
        out.path = path == Path::Lanczos ? 1 : 2;
        std::vector<msg_ctx::Shape> cand;
        if (path == Path::Lanczos && (p.method == MSG_METHOD_RK4 || r.mix_mult > 0)) {
            cand.push_back({0, 0, 0});  // the RK4 kernel has a single launch shape per dimension
        } else if (path == Path::Lanczos && !sym_for(r, &p)) {
            // Arnoldi (unequal recombination rates): 1 or 2 groups per thread for D = 1024-4096,
            // 2 at D = 8192, 4 at D = 16384
            const int ng = (int)(r.D / 4);
            const int g0 = ng <= 1024 ? 1 : ng <= 2048 ? 2 : 4;
            cand.push_back({g0, ng / g0, 0});
            if (ng >= 256 && ng <= 1024) cand.push_back({2, ng / 2, 0});
        } else if (path == Path::Lanczos) {
            const int ng = (int)(r.D / 4);
            std::vector<msg_ctx::Shape> base;
            if (ng <= 512) base.push_back({1, ng, 0});
            if (ng == 256 || ng == 512) base.push_back({2, ng / 2, 0});
            // D = 4096: four groups per thread first, the usual winner over a whole launch (486
            // against 523 ms on the orientation batch, where the two-step trials put it only
            // 1-6 % ahead of two groups per thread)
            if (ng == 1024) { base.push_back({4, 256, 0}); base.push_back({2, 512, 0}); base.push_back({1, 1024, 0}); }
            if (ng == 2048) base.push_back({2, 1024, 0});
            if (ng == 4096) { base.push_back({-1, 1024, 0}); base.push_back({0, 1024, 0}); }
            if (ng > 4096) { base.push_back({-2, 1024, 0}); base.push_back({0, 1024, 0}); }
            if (ng >= 4096) {  // split columns (cooperative launch), 2 or 4 groups per thread
                base.push_back({-3, 1024, 0});
                base.push_back({-3, 512, 0});
            }
            for (const auto &c : base) {
                cand.push_back(c);  // every resident workgroup the occupancy allows
                msg_ctx::Shape h = c;
                h.slots = -2;       // half of them (fewer columns in flight, more cache per column)
                cand.push_back(h);
            }
        } else {
            cand.push_back({0, 0, 0});  // the dense path has a single launch shape
        }
        int nobs = 0;
        {
            int32_t n = 0;
            if (msg_observable_count(sys, &p, &n) != MSG_OK) throw ArgError{MSG_INVALID_ARG, "observables"};
            nobs = n;
        }
        std::vector<double> y((size_t)n_items * std::max(1, nobs));
        double best_ms = 1e300;
        msg_ctx::Shape best_shape{};
        for (auto c : cand) {
            double ms_min = 1e300;
            int slots_used = 0;
            for (int rep = 0; rep < 2; ++rep) {
                msg_ctx::Shape trial = c;
                Batch b;
                if (path == Path::Lanczos && c.slots == -2) {
                    // resolve "half the resident workgroups" against this launch's occupancy
                    trial.slots = 0;
                    ctx->force = &trial;
                    Batch probe;
                    prepare(probe, ctx, sys, &p, n_items, fields, columns, n_columns, false, false, ctx->stream);
                    trial.slots = std::max(1, probe.slots / 2);
                }
                ctx->force = path == Path::Lanczos ? &trial : nullptr;
                try {
                    prepare(b, ctx, sys, &p, n_items, fields, columns, n_columns, false, false, ctx->stream);
                    ctx->force = nullptr;
                    b.timed = true;
                    advance(b, b.nsamples);
                    collect(b, y.data(), nullptr);
                } catch (const CoopRefused &) {
                    // a split-column candidate whose cooperative launch the device refuses
                    // (nothing was launched) drops out; every other failure is reported
                    ctx->force = nullptr;
                    ms_min = 1e300;
                    break;
                }
                ms_min = std::min(ms_min, batch_kernel_ms(b));
                slots_used = b.slots;
                if (path == Path::Lanczos && c.slots == -2) c.slots = trial.slots;
            }
            // a later candidate has to beat the earlier ones by more than the trials' noise
            // (about 1 %): a short trial can rank shapes within that margin wrongly for the
            // full launch, and the candidates are listed with the usual winner first
            if (ms_min < best_ms * (1.0 - kTuneMargin)) {
                best_ms = ms_min;
                best_shape = c;
                out.groups_per_thread = c.G;
                out.threads = path == Path::Lanczos ? c.T : dense_threads(dense_padding((int)r.D));
                out.slots = slots_used;
            }
        }
        ctx->force = nullptr;
        if (path == Path::Lanczos && p.method != MSG_METHOD_RK4 && r.mix_mult == 0) ctx->tuned[r.D] = best_shape;
        out.kernel_ms = best_ms;
        if (best) *best = out;
    });
}

msg_status msg_propagate_uneven(msg_ctx *ctx, int32_t n_entries, const msg_uneven_entry *entries) {
    if (!ctx || n_entries < 1 || !entries) return fail(ctx, MSG_INVALID_ARG, "null context or no entries");
    return guarded(ctx, [&] {
        HIPCHECK(hipSetDevice(ctx->device));
        UnevenRun run;
        uneven_begin(run, ctx, n_entries, entries, ctx->stream, true);
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        uneven_finish(run, entries);
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->kernel_ms = ms;
        ctx->launches = 1;
    });
}

msg_status msg_step(msg_ctx *ctx, const msg_system *sys, const msg_run_params *params, const double *field, double dt,
                    int64_t step_index, double *factors, int32_t n_columns) {
    if (!ctx || !sys || !params || !factors || n_columns < 1) return fail(ctx, MSG_INVALID_ARG, "null argument");
    return guarded(ctx, [&] {
        HIPCHECK(hipSetDevice(ctx->device));
        if (params->dynamics && sys->r.has_td)
            throw ArgError{MSG_UNSUPPORTED, "msg_step takes a static Hamiltonian (dynamics=static)"};
        // one interval [0, dt] with the relaxation draws of step `step_index`
        msg_run_params p = *params;
        p.sampling = MSG_SAMPLING_GIVEN;
        p.calculation = MSG_CALC_TIMEEVOLUTION;
        p.total_time = dt;
        p.time_step = dt;
        p.yield_corrections = 0;
        Batch b;
        prepare(b, ctx, sys, &p, 1, field, factors, n_columns, false, true, ctx->stream);
        if (b.nsamples != 2) throw ArgError{MSG_INVALID_ARG, "dt must be positive"};
        if (step_index < 0 || step_index > 0x7fffffff) throw ArgError{MSG_INVALID_ARG, "step index out of range"};
        b.a.rng_step0 = (int)step_index;  // relaxation draws of interval `step_index`
        run_chunk(b, 1);
        // read the state back (device order) into MolSpin order
        const Resolved &r = sys->r;
        const int D = (int)r.D;
        std::vector<float> st(b.state.n / sizeof(float));
        HIPCHECK(hipMemcpyAsync(st.data(), b.a.state, b.state.n, hipMemcpyDeviceToHost, b.stream));
        int status = 0;
        HIPCHECK(hipMemcpyAsync(&status, b.a.status, sizeof(int), hipMemcpyDeviceToHost, b.stream));
        HIPCHECK(hipStreamSynchronize(b.stream));
        if (status) throw ArgError{MSG_NUMERIC_BREAKDOWN, "non-finite state during the step"};
        // the Lanczos path propagates the unitary part; apply the symmetric decay of K
        const double decay = r.symmetric ? std::exp(-r.kappa * dt) : 1.0;
        for (int j = 0; j < n_columns; ++j)
            for (int i = 0; i < D; ++i) {
                const int64_t dev = r.perm[i];
                const int64_t nu = dev >> 2, c = dev & 3;
                float re, im;
                if (b.path == Path::Lanczos) {
                    const size_t off = (size_t)j * 2 * D + (c < 2 ? 4 * nu + 2 * c : D + 4 * nu + 2 * (c - 2));
                    re = st[off];
                    im = st[off + 1];
                } else {
                    const int Dp = b.a.dpad, tile = j / 16, tc = j % 16;
                    const float *u = &st[(size_t)tile * 2 * Dp * 16];
                    re = u[dev * 16 + tc];
                    im = u[Dp * 16 + dev * 16 + tc];
                }
                factors[2 * ((size_t)j * D + i)] = decay * re;
                factors[2 * ((size_t)j * D + i) + 1] = decay * im;
            }
    });
}

msg_status msg_symmetric_decay(msg_ctx *ctx, const msg_system *sys, int32_t n_items, const double *fields,
                               double rate, double *yields) {
    if (!ctx || !sys || !yields || n_items < 1) return fail(ctx, MSG_INVALID_ARG, "null argument or no items");
    return guarded(ctx, [&] {
        const Resolved &r = sys->r;
        if (!r.pair_ok)
            throw ArgError{MSG_UNSUPPORTED, "symmetric decay on the device needs States on the radical-pair electrons"};
        if (r.initial.empty()) throw ArgError{MSG_INVALID_ARG, "no initial state"};
        if (r.D > 4096) throw ArgError{MSG_UNSUPPORTED, "symmetric decay on the device is limited to D <= 4096"};
        if (rate < 0.0 || std::isnan(rate)) throw ArgError{MSG_INVALID_ARG, "rate must be positive (0: from the system)"};
        // the reference's rule when no rate is given: first Transition's rate, else 1e-3
        const double k = rate > 0.0 ? rate : (r.first_rate >= 0.0 ? r.first_rate : 1e-3);
        if (!(k > 0.0)) throw ArgError{MSG_INVALID_ARG, "decay rate must be positive"};
        const int D = (int)r.D;
        const int S = (int)r.P_states.size();
        hipStream_t st = ctx->stream;
        (void)blas_on(ctx, st);
        // rho0 = (sum_i P_i) (x) 1 / tr(...)  (TaskStaticHSSymmetricDecay.cpp:100-112)
        double rho[32] = {0.0};
        double tr = 0.0;
        for (int s0 : r.initial) {
            const Mat4 &P = r.P_states[s0];
            for (int i = 0; i < 16; ++i) {
                rho[2 * i] += P[i].real();
                rho[2 * i + 1] += P[i].imag();
            }
            tr += (P[0] + P[5] + P[10] + P[15]).real();
        }
        if (!(tr > 0.0)) throw ArgError{MSG_INVALID_ARG, "initial state has zero trace"};
        const double rho_scale = 1.0 / (tr * (double)(r.D / 4));
        DBuf A, Y, Rt, Pt, lam, off, info, rows, res;
        cd *dA = A.alloc<cd>((size_t)D * D, st);
        cd *dY = Y.alloc<cd>((size_t)D * D, st);
        cd *dR = Rt.alloc<cd>((size_t)D * D, st);
        cd *dP = Pt.alloc<cd>((size_t)D * D, st);
        double *dl = lam.alloc<double>(D, st);
        double *de = off.alloc<double>(D, st);
        rocblas_int *di = info.alloc<rocblas_int>(1, st);
        double *drows = rows.alloc<double>(D, st);
        double *dres = res.alloc<double>((size_t)n_items * S, st);
        std::vector<cd> H;
        for (int it = 0; it < n_items; ++it) {
            dense_hamiltonian(r, fields ? fields + 3 * it : nullptr, H);  // device order, row-major
            HIPCHECK(hipMemcpyAsync(dA, H.data(), (size_t)D * D * sizeof(cd), hipMemcpyHostToDevice, st));
            // The solver reads the row-major H as column-major, i.e. conj(H) = W L W^dagger; the
            // row-major view of W is then M = W^T = V^dagger with H = V L V^dagger.
            if (rocsolver_zheevd(ctx->blas, rocblas_evect_original, rocblas_fill_upper, D,
                                 reinterpret_cast<rocblas_double_complex *>(dA), D, dl, de, di) !=
                rocblas_status_success)
                throw ArgError{MSG_DEVICE_ERROR, "rocsolver_zheevd failed"};
            rocblas_int hinfo = 0;
            HIPCHECK(hipMemcpyAsync(&hinfo, di, sizeof(hinfo), hipMemcpyDeviceToHost, st));
            HIPCHECK(launch_blockdiag_right(dA, rho, rho_scale, dY, D, st));       // Y = M rho0
            HIPCHECK(zgemm(ctx, dY, dA, dR, D, 1, 1.0, st, true));          // Rt = M rho0 M^dagger
            for (int s0 = 0; s0 < S; ++s0) {
                double P[32];
                for (int i = 0; i < 16; ++i) {
                    P[2 * i] = r.P_states[s0][i].real();
                    P[2 * i + 1] = r.P_states[s0][i].imag();
                }
                HIPCHECK(launch_blockdiag_right(dA, P, 1.0, dY, D, st));           // Y = M P
                HIPCHECK(zgemm(ctx, dY, dA, dP, D, 1, 1.0, st, true));      // Pt = M P M^dagger
                HIPCHECK(launch_lorentz_sum(dP, dR, dl, k, D, drows, dres + (size_t)it * S + s0, st));
            }
            HIPCHECK(hipStreamSynchronize(st));
            if (hinfo != 0) throw ArgError{MSG_NUMERIC_BREAKDOWN, "eigensolver did not converge"};
        }
        HIPCHECK(hipMemcpyAsync(yields, dres, (size_t)n_items * S * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    });
}

msg_status msg_time_infinity(msg_ctx *ctx, const msg_system *sys, const msg_run_params *params, int32_t n_items,
                             const double *fields, const double *columns, int32_t n_columns, double *yields) {
    if (!ctx || !sys || !params || !yields || n_items < 1) return fail(ctx, MSG_INVALID_ARG, "null argument or no items");
    return guarded(ctx, [&] {
        const msg_run_params *p = params;
        check_params(sys, p);
        const Resolved &r = sys->r;
        if (p->calculation != MSG_CALC_YIELDS) throw ArgError{MSG_INVALID_ARG, "method=timeinf needs calculation=yields"};
        if (p->task != MSG_TASK_HSGENERAL) throw ArgError{MSG_INVALID_ARG, "time-infinity yields are an HSGeneral mode"};
        if (p->dynamics)
            throw ArgError{MSG_INVALID_ARG, "dynamic yields require finite-time propagation; method=timeinf is only "
                                            "defined for static Hamiltonians"};
        if (r.relax)
            throw ArgError{MSG_UNSUPPORTED, "time-infinity yields with relaxation operators need the Liouville-space solve"};
        const int nobs = observable_count(r, p);
        if (nobs < 1) throw ArgError{MSG_INVALID_ARG, "no observables (no transitions / States)"};
        const std::vector<Mat4> &Pobs = p->observables == MSG_OBS_TRANSITION_YIELDS ? r.P_trans : r.P_states;
        std::vector<double> sc(nobs, 1.0);
        if (p->observables == MSG_OBS_TRANSITION_YIELDS) sc = r.rates;
        const int D = (int)r.D;
        const bool given = p->sampling == MSG_SAMPLING_GIVEN;
        if (given && (!columns || n_columns < 1)) throw ArgError{MSG_INVALID_ARG, "sampling=given needs columns"};
        if (given && n_columns > D) throw ArgError{MSG_UNSUPPORTED, "time-infinity yields take at most D columns"};
        const bool eig = r.symmetric && D <= 4096 && p->method != MSG_METHOD_EXPONENTIAL;
        if (!eig && D > 2048)
            throw ArgError{MSG_UNSUPPORTED, r.symmetric ? "time-infinity yields are limited to D <= 4096"
                                                        : "time-infinity yields with unequal recombination need D <= 2048"};
        const char *diverges = "failed to solve the static Hilbert time-integral equation: part of the state never decays";
        if (r.symmetric && !(r.kappa > 0.0)) throw ArgError{MSG_NUMERIC_BREAKDOWN, diverges};
        hipStream_t st = ctx->stream;
        HIPCHECK(hipSetDevice(ctx->device));
        const size_t DD = (size_t)D * D;
        // rho0 pieces: rho_e = sum_i evec_i evec_i^dagger with 1/Z, or the caller's columns
        // as a zero-padded D x D row-major matrix in device order
        double rho_e[32] = {0.0};
        for (const auto &v : r.evec)
            for (int a = 0; a < 4; ++a)
                for (int c = 0; c < 4; ++c) {
                    const cd x = v[a] * std::conj(v[c]);
                    rho_e[2 * (a * 4 + c)] += x.real();
                    rho_e[2 * (a * 4 + c) + 1] += x.imag();
                }
        const double inv_z = 4.0 / (double)D;
        std::vector<cd> Cpad;
        if (given) {
            Cpad.assign(DD, cd(0.0));
            for (int j = 0; j < n_columns; ++j)
                for (int64_t i = 0; i < D; ++i)
                    Cpad[(size_t)r.perm[i] * D + j] = cd(columns[2 * ((size_t)j * D + i)], columns[2 * ((size_t)j * D + i) + 1]);
        }
        std::vector<cd> H;
        std::vector<double> out((size_t)n_items * nobs, 0.0);
        if (eig) {
            // eigenbasis closed form: Re Tr[P X] = sum_gh Re(conj(Pt_gh) Rt_gh / (2 kappa + i w_gh))
            (void)blas_on(ctx, st);
            DBuf A, Y, Rt, Pt, lam, off, info, rows, res, Cb;
            cd *dA = A.alloc<cd>(DD, st), *dY = Y.alloc<cd>(DD, st), *dR = Rt.alloc<cd>(DD, st);
            cd *dP = Pt.alloc<cd>(DD, st);
            double *dl = lam.alloc<double>(D, st), *de = off.alloc<double>(D, st);
            rocblas_int *di = info.alloc<rocblas_int>(1, st);
            double *drows = rows.alloc<double>(D, st);
            double *dres = res.alloc<double>((size_t)n_items * nobs, st);
            cd *dC = nullptr;
            if (given) {
                dC = Cb.alloc<cd>(DD, st);
                HIPCHECK(hipMemcpyAsync(dC, Cpad.data(), DD * sizeof(cd), hipMemcpyHostToDevice, st));
            }
            for (int it = 0; it < n_items; ++it) {
                dense_hamiltonian(r, fields ? fields + 3 * it : nullptr, H);
                HIPCHECK(hipMemcpyAsync(dA, H.data(), DD * sizeof(cd), hipMemcpyHostToDevice, st));
                // row-major H read as column-major: the row-major view of the eigenvectors is M = V^dagger
                if (rocsolver_zheevd(ctx->blas, rocblas_evect_original, rocblas_fill_upper, D,
                                     reinterpret_cast<rocblas_double_complex *>(dA), D, dl, de, di) !=
                    rocblas_status_success)
                    throw ArgError{MSG_DEVICE_ERROR, "rocsolver_zheevd failed"};
                rocblas_int hinfo = 0;
                HIPCHECK(hipMemcpyAsync(&hinfo, di, sizeof(hinfo), hipMemcpyDeviceToHost, st));
                if (given) {
                    HIPCHECK(zgemm(ctx, dA, dC, dY, D, 1, 1.0, st));         // Y = M C
                    HIPCHECK(zgemm(ctx, dY, dY, dR, D, 1, 1.0, st, true));   // Rt = Y Y^dagger
                } else {
                    HIPCHECK(launch_blockdiag_right(dA, rho_e, inv_z, dY, D, st));  // Y = M rho0
                    HIPCHECK(zgemm(ctx, dY, dA, dR, D, 1, 1.0, st, true));   // Rt = M rho0 M^dagger
                }
                for (int c = 0; c < nobs; ++c) {
                    double P[32];
                    for (int i = 0; i < 16; ++i) {
                        P[2 * i] = Pobs[c][i].real();
                        P[2 * i + 1] = Pobs[c][i].imag();
                    }
                    HIPCHECK(launch_blockdiag_right(dA, P, 1.0, dY, D, st));           // Y = M P
                    HIPCHECK(zgemm(ctx, dY, dA, dP, D, 1, 1.0, st, true));      // Pt = M P M^dagger
                    HIPCHECK(launch_timeinf_sum(dP, dR, dl, 2.0 * r.kappa, D, drows, dres + (size_t)it * nobs + c, st));
                }
                HIPCHECK(hipStreamSynchronize(st));
                if (hinfo != 0) throw ArgError{MSG_NUMERIC_BREAKDOWN, "eigensolver did not converge"};
            }
            HIPCHECK(hipMemcpyAsync(out.data(), dres, out.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
        } else {
            // doubling on X = int_0^inf e^{Gt} rho0 e^{G^dagger t} dt, G = -iH - K (x) 1
            // ten D x D FP64 complex buffers per item of a chunk, at most ~12 GiB in all
            const int chunk = std::max(1, std::min(n_items, (int)(((size_t)12 << 30) / (10 * DD * sizeof(cd)))));
            DBuf bG, bR, bX, bX2, bX3, bX4, bB, bP, bQ, bn, bt;
            cd *dG = bG.alloc<cd>((size_t)chunk * DD, st), *dR = bR.alloc<cd>((size_t)chunk * DD, st);
            cd *dX = bX.alloc<cd>((size_t)chunk * DD, st), *dX2 = bX2.alloc<cd>((size_t)chunk * DD, st);
            cd *dX3 = bX3.alloc<cd>((size_t)chunk * DD, st), *dX4 = bX4.alloc<cd>((size_t)chunk * DD, st);
            cd *dB = bB.alloc<cd>((size_t)chunk * DD, st), *dP = bP.alloc<cd>((size_t)chunk * DD, st);
            cd *dQ = bQ.alloc<cd>((size_t)chunk * DD, st);
            const int parts = 64;  // slabs per matrix of the fused accumulation and norms
            double *dn = bn.alloc<double>(2 * (size_t)chunk * parts, st);
            cd *dtr = bt.alloc<cd>((size_t)chunk * 16, st);
            // rho0 (device order, row-major), replicated over the chunk
            if (given) {
                HIPCHECK(hipMemcpyAsync(dX, Cpad.data(), DD * sizeof(cd), hipMemcpyHostToDevice, st));
                HIPCHECK(zgemm(ctx, dX, dX, dR, D, 1, 1.0, st, true));  // C C^dagger
            } else {
                std::vector<cd> rho0(DD, cd(0.0));
                for (int nu = 0; nu < D / 4; ++nu)
                    for (int a = 0; a < 4; ++a)
                        for (int c = 0; c < 4; ++c)
                            rho0[(size_t)(4 * nu + a) * D + 4 * nu + c] =
                                inv_z * cd(rho_e[2 * (a * 4 + c)], rho_e[2 * (a * 4 + c) + 1]);
                HIPCHECK(hipMemcpyAsync(dR, rho0.data(), DD * sizeof(cd), hipMemcpyHostToDevice, st));
                HIPCHECK(hipStreamSynchronize(st));  // rho0 is released at scope end
            }
            for (int q = 1; q < chunk; ++q)
                HIPCHECK(hipMemcpyAsync(dR + (size_t)q * DD, dR, DD * sizeof(cd), hipMemcpyDeviceToDevice, st));
            // generators: a field sweep forms G(f) = G(0) + sum_a f_a Z_a on the device
            // (SweepGenerators); a single item or no fields builds G on the host
            const bool linear = fields != nullptr && n_items > 1;
            SweepGenerators sweep;
            std::vector<cd> Gh;
            if (linear)
                sweep.build(r, D, true, chunk, st);
            else
                Gh.resize((size_t)chunk * DD);
            std::vector<double> fro(2 * (size_t)chunk * parts);
            std::vector<cd> tr((size_t)chunk * 16);
            for (int it0 = 0; it0 < n_items; it0 += chunk) {
                const int nb = std::min(chunk, n_items - it0);
                double nrm = 0.0;
                if (linear) {
                    for (int q = 0; q < nb; ++q) nrm = std::max(nrm, sweep.bound(fields + 3 * (size_t)(it0 + q)));
                    sweep.form(fields + 3 * (size_t)it0, nb, 1.0, D, dG, st);
                } else {
                    for (int q = 0; q < nb; ++q) {
                        dense_hamiltonian(r, fields ? fields + 3 * (it0 + q) : nullptr, H);
                        cd *g = &Gh[(size_t)q * DD];
                        for (size_t k = 0; k < DD; ++k) g[k] = cd(H[k].imag(), -H[k].real());  // -i H
                        for (int nu = 0; nu < D / 4; ++nu)
                            for (int a = 0; a < 4; ++a)
                                for (int c = 0; c < 4; ++c) g[(size_t)(4 * nu + a) * D + 4 * nu + c] -= r.K[a * 4 + c];
                        nrm = std::max(nrm, norm1(g, D));
                    }
                    HIPCHECK(hipMemcpyAsync(dG, Gh.data(), (size_t)nb * DD * sizeof(cd), hipMemcpyHostToDevice, st));
                }
                if (!(nrm > 0.0)) throw ArgError{MSG_NUMERIC_BREAKDOWN, diverges};
                const double tau = 0.25 / nrm;  // ||tau G||_1 <= 1/4
                // U = e^{tau G}: Taylor(18) of tau G (no squaring needed at norm 1/4)
                const double cs[4] = {0.0, tau, 0.0, 0.0};
                HIPCHECK(launch_zlincomb(dG, dG, dG, dX, D, nb, cs, st));
                cd *dU = dP, *dU2 = dQ;
                taylor_batch(ctx, dX, dX2, dX3, dX4, dB, dU, dU2, D, nb, false, st);
                // X_tau = tau S_0: S_N = rho0, S_n = rho0 + tau/(n+2) (G S + (G S)^dagger), N = 16
                const int N = 16;
                cd *dS = dX3, *dY = dX4;
                const cd *src = dR;
                for (int n = N - 1; n >= 0; --n) {
                    HIPCHECK(zgemm(ctx, dG, src, dY, D, nb, 1.0, st));
                    HIPCHECK(launch_herm_update(dY, dR, dS, D, nb, tau / (n + 2), n == 0 ? tau : 1.0, st));
                    src = dS;
                }
                // doubling from T to 2T: the integral over [T, 2T] is Delta = U_T X_T U_T^dagger;
                // X <- X + Delta, U <- U^2, until ||Delta||_F <= 1e-10 ||X||_F for every item
                // (the tail left out is then ~1e-10 of X while it decays).  A part of the
                // space that never decays but is never populated adds nothing beyond its
                // rounding-level seed, which grows only linearly in T (the reference's
                // singular solve falls back to the minimum-norm solution there); a populated
                // one makes X grow without bound (Delta ~ X).  Repeated squaring drifts the
                // modulus of U's undamped part by ~2^d eps, so the doubling stops at
                // T = tau 2^36 (drift < 1e-5) and reports divergence there.
                const int kMaxDoublings = 36;
                cd *dXc = dS, *dXn = dB, *dW = dY;
                bool done = false;
                for (int d = 0; d < kMaxDoublings && !done; ++d) {
                    HIPCHECK(zgemm(ctx, dU, dXc, dW, D, nb, 1.0, st));              // W = U X
                    HIPCHECK(zgemm(ctx, dW, dU, dXn, D, nb, 1.0, st, true));        // Delta = W U^dagger
                    HIPCHECK(launch_accumulate_fro(dXc, dXn, dW, D, nb, parts, dn, st));  // X + Delta, norms
                    std::swap(dXc, dW);
                    HIPCHECK(zgemm(ctx, dU, dU, dU2, D, nb, 1.0, st));              // U^2
                    std::swap(dU, dU2);
                    HIPCHECK(hipMemcpyAsync(fro.data(), dn, 2 * (size_t)nb * parts * sizeof(double), hipMemcpyDeviceToHost,
                                            st));
                    HIPCHECK(hipStreamSynchronize(st));
                    done = true;
                    for (int q = 0; q < nb; ++q) {
                        double fd = 0.0, fx = 0.0;  // ||Delta||_F^2, ||X||_F^2
                        for (int k = 0; k < parts; ++k) {
                            fd += fro[2 * ((size_t)q * parts + k)];
                            fx += fro[2 * ((size_t)q * parts + k) + 1];
                        }
                        if (!std::isfinite(fd) || !std::isfinite(fx))
                            throw ArgError{MSG_NUMERIC_BREAKDOWN, "non-finite value in the time-integral doubling"};
                        done &= fd <= 1e-20 * fx;
                    }
                }
                if (!done) {
                    char msg[200];
                    std::snprintf(msg, sizeof msg, "%s within %.3g ns", diverges, tau * std::ldexp(1.0, kMaxDoublings));
                    throw ArgError{MSG_NUMERIC_BREAKDOWN, msg};
                }
                HIPCHECK(launch_pair_block_trace(dXc, D, nb, dtr, st));
                HIPCHECK(hipMemcpyAsync(tr.data(), dtr, (size_t)nb * 16 * sizeof(cd), hipMemcpyDeviceToHost, st));
                HIPCHECK(hipStreamSynchronize(st));
                // Tr[(P (x) 1) X] = sum_ab P_ab r_ba with r_ba = sum_nu X[4 nu + b][4 nu + a]
                for (int q = 0; q < nb; ++q)
                    for (int c = 0; c < nobs; ++c) {
                        cd acc = 0.0;
                        for (int a = 0; a < 4; ++a)
                            for (int b = 0; b < 4; ++b) acc += Pobs[c][a * 4 + b] * tr[(size_t)q * 16 + b * 4 + a];
                        out[(size_t)(it0 + q) * nobs + c] = acc.real();
                    }
            }
        }
        for (int it = 0; it < n_items; ++it)
            for (int c = 0; c < nobs; ++c) yields[(size_t)it * nobs + c] = sc[c] * out[(size_t)it * nobs + c];
    });
}

msg_status msg_last_kernel_time(const msg_ctx *ctx, double *ms, int32_t *launches) {
    if (!ctx || !ms) return MSG_INVALID_ARG;
    *ms = ctx->kernel_ms;
    if (launches) *launches = ctx->launches;
    return MSG_OK;
}

msg_status msg_server_create(msg_ctx *ctx, int32_t max_slots, int64_t hbm_budget_bytes, msg_server **out) {
    if (!ctx || !out || max_slots < 1 || hbm_budget_bytes < 1) return fail(ctx, MSG_INVALID_ARG, "bad server arguments");
    return guarded(ctx, [&] {
        std::unique_ptr<msg_server> s(new msg_server());
        s->ctx = ctx;
        s->max_slots = max_slots;
        s->budget = hbm_budget_bytes;
        for (int i = 0; i < max_slots; ++i) {
            hipStream_t st;
            HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            s->streams.push_back(st);
        }
        s->stream_busy.assign(max_slots, 0);
        if (const char *env = std::getenv("MSG_SERVER_GROUP_WORK")) s->group_work = std::max(0.0, std::atof(env));
        *out = s.release();
    });
}

msg_status msg_server_destroy(msg_server *srv) {
    if (!srv) return MSG_INVALID_ARG;
    for (hipStream_t st : srv->streams) (void)hipStreamSynchronize(st);
    for (auto &q : srv->active)
        if (q->done_ev) (void)hipEventDestroy(q->done_ev);
    for (auto &g : srv->groups)
        if (g->done) (void)hipEventDestroy(g->done);
    srv->groups.clear();
    srv->active.clear();
    srv->finished.clear();
    // requests may have left the context's rocBLAS handle bound to a server stream
    if (srv->ctx->blas) (void)rocblas_set_stream(srv->ctx->blas, srv->ctx->stream);
    for (hipStream_t st : srv->streams) (void)hipStreamDestroy(st);
    delete srv;
    return MSG_OK;
}

msg_status msg_server_submit(msg_server *srv, const msg_system *sys, const msg_run_params *params, int32_t n_fields,
                             const double *fields, int64_t *request_id) {
    if (!srv || !sys || !params || n_fields < 1 || !fields || !request_id)
        return fail(srv ? srv->ctx : nullptr, MSG_INVALID_ARG, "bad request");
    return guarded(srv->ctx, [&] {
        check_params(sys, params);
        if (params->calculation != MSG_CALC_YIELDS) throw ArgError{MSG_INVALID_ARG, "the server serves yield requests"};
        if (params->sampling != MSG_SAMPLING_DIRECT) throw ArgError{MSG_UNSUPPORTED, "the server uses direct sampling"};
        if (params->dynamics && sys->r.has_td && choose_path(sys->r, params) == Path::Dense)
            throw ArgError{MSG_UNSUPPORTED, "time-dependent requests on the dense path run through msg_propagate"};
        std::unique_ptr<Request> q(new Request());
        q->id = srv->next_id++;
        q->sys = sys;
        q->prm = *params;
        q->n_fields = n_fields;
        q->fields.assign(fields, fields + 3 * (size_t)n_fields);
        int64_t bytes = 0;
        if (msg_query_footprint(sys, params, n_fields, 0, &bytes) != MSG_OK)
            throw ArgError{MSG_UNSUPPORTED, "request not representable on the device"};
        q->bytes = bytes;
        if (bytes > srv->budget) throw ArgError{MSG_OOM_ADMISSION, "request exceeds the HBM budget"};
        std::vector<double> t;
        (void)free_times(params->total_time, params->time_step, t);
        q->work = (double)sys->r.D * (double)sys->r.zdim * std::max<size_t>(1, sys->r.evec.size()) * n_fields *
                  (double)t.size();
        *request_id = q->id;
        srv->queued.push_back(std::move(q));
    });
}

msg_status msg_server_pump(msg_server *srv, int32_t max_steps, int32_t *n_active) {
    if (!srv) return MSG_INVALID_ARG;
    return guarded(srv->ctx, [&] {
        msg_ctx *ctx = srv->ctx;
        HIPCHECK(hipSetDevice(ctx->device));
        // yields of the launch in flight into the request's rows [f_begin, f_begin + f_count)
        auto collect_part = [&](Request &q) {
            if (q.result.empty()) q.result.assign((size_t)q.n_fields * q.batch->nobs, 0.0);
            try {
                collect(*q.batch, q.result.data() + (size_t)q.f_begin * q.batch->nobs, nullptr);
            } catch (const ArgError &e) {
                q.failed = true;
                q.error = e.what;
            }
        };
        // one persistent launch over the whole time grid for fields [f_begin, f_begin + f_count)
        auto launch_part = [&](Request &q) {
            q.batch.reset(new Batch());
            q.batch->no_split = true;  // requests share the device
            prepare(*q.batch, ctx, q.sys, &q.prm, q.f_count, q.fields.data() + 3 * (size_t)q.f_begin, nullptr, 0, false,
                    false, srv->streams[q.stream]);
            run_chunk(*q.batch, q.batch->nsamples);
            if (!q.done_ev) HIPCHECK(hipEventCreateWithFlags(&q.done_ev, hipEventDisableTiming));
            HIPCHECK(hipEventRecord(q.done_ev, srv->streams[q.stream]));
        };
        auto finish = [&](std::unique_ptr<Request> &q) {
            if (!q->whole) {
                q->f_begin = 0;
                collect_part(*q);
            }
            srv->in_use -= q->bytes;
            q->batch.reset();
            if (q->done_ev) (void)hipEventDestroy(q->done_ev);
            q->done_ev = nullptr;
            srv->stream_busy[q->stream] = 0;
            const int64_t id = q->id;
            srv->finished[id] = std::move(q);
        };
        // 1. completions (non-blocking): shared launches, then requests of their own
        for (auto it = srv->groups.begin(); it != srv->groups.end();) {
            RequestGroup &g = **it;
            if (hipEventQuery(g.done) != hipSuccess) {
                ++it;
                continue;
            }
            std::string err;
            try {
                uneven_finish(*g.run, g.ents.data());
            } catch (const ArgError &e) {
                err = e.what;
            }
            g.run.reset();  // releases the launch's buffers (stream-ordered)
            (void)hipEventDestroy(g.done);
            srv->stream_busy[g.stream] = 0;
            srv->in_use -= g.bytes;
            for (auto &q : g.members) {
                if (!err.empty()) {
                    q->failed = true;
                    q->error = err;
                }
                const int64_t id = q->id;
                srv->finished[id] = std::move(q);
            }
            it = srv->groups.erase(it);
        }
        for (auto it = srv->active.begin(); it != srv->active.end();) {
            Request &q = **it;
            const bool done = q.whole ? hipEventQuery(q.done_ev) == hipSuccess : q.batch->s_next >= q.batch->nsamples - 1;
            if (done && q.whole) {
                collect_part(q);
                q.f_begin += q.f_count;
                if (!q.failed && q.f_begin < q.n_fields) {  // next field chunk on the same stream
                    q.f_count = std::min(q.f_count, q.n_fields - q.f_begin);
                    try {
                        launch_part(q);
                    } catch (const ArgError &e) {
                        q.failed = true;
                        q.error = e.what;
                    }
                    if (!q.failed) {
                        ++it;
                        continue;
                    }
                }
            }
            if (done) {
                finish(*it);
                it = srv->active.erase(it);
            } else {
                ++it;
            }
        }
        // 2a. whole mode: the queued requests below kServerGroupWork that the uneven
        //     launch takes share one launch on a free stream, smallest first, while their
        //     summed work stays below kServerGroupWork and their footprint fits
        while (max_steps <= 0) {
            int slot = -1;
            for (int i = 0; i < (int)srv->streams.size(); ++i)
                if (!srv->stream_busy[i]) { slot = i; break; }
            if (slot < 0) break;
            std::vector<size_t> cand;
            for (size_t i = 0; i < srv->queued.size(); ++i) {
                const Request &q = *srv->queued[i];
                if (!q.no_group && q.work <= srv->group_work && uneven_eligible(q.sys, &q.prm)) cand.push_back(i);
            }
            std::stable_sort(cand.begin(), cand.end(),
                             [&](size_t x, size_t y) { return srv->queued[x]->work < srv->queued[y]->work; });
            std::vector<size_t> take;
            double work = 0.0;
            int64_t bytes = 0;
            for (size_t i : cand) {
                const Request &q = *srv->queued[i];
                if (work + q.work > srv->group_work || srv->in_use + bytes + q.bytes > srv->budget) continue;
                take.push_back(i);
                work += q.work;
                bytes += q.bytes;
            }
            if (take.size() < 2) break;  // a single request runs on its own below
            std::sort(take.begin(), take.end());
            std::unique_ptr<RequestGroup> g(new RequestGroup());
            for (auto k = take.rbegin(); k != take.rend(); ++k) {
                g->members.push_back(std::move(srv->queued[*k]));
                srv->queued.erase(srv->queued.begin() + (std::ptrdiff_t)*k);
            }
            std::reverse(g->members.begin(), g->members.end());
            for (auto &q : g->members) {
                q->stream = slot;
                q->whole = true;
                q->result.assign((size_t)q->n_fields * std::max(1, observable_count(q->sys->r, &q->prm)), 0.0);
                msg_uneven_entry en{};
                en.system = q->sys;
                en.params = &q->prm;
                en.n_items = q->n_fields;
                en.fields = q->fields.data();
                en.yields = q->result.data();
                g->ents.push_back(en);
            }
            g->run.reset(new UnevenRun());
            try {
                uneven_begin(*g->run, ctx, (int)g->ents.size(), g->ents.data(), srv->streams[slot], false);
            } catch (const ArgError &) {
                // the shared launch refused a member: every member goes back to the
                // queue and runs on its own (nothing of the refused launch ran)
                g->run.reset();
                for (auto &q : g->members) {
                    q->no_group = true;
                    srv->queued.push_back(std::move(q));
                }
                continue;
            }
            HIPCHECK(hipEventCreateWithFlags(&g->done, hipEventDisableTiming));
            HIPCHECK(hipEventRecord(g->done, srv->streams[slot]));
            g->stream = slot;
            g->bytes = bytes;
            srv->stream_busy[slot] = 1;
            srv->in_use += bytes;
            ++srv->shared_launches;
            srv->shared_requests += (int64_t)g->members.size();
            srv->groups.push_back(std::move(g));
        }
        // 2b. admission by free stream and HBM footprint, shortest job first (arrival order
        //     among equals).  (Capping each launch's resident workgroups while others wait
        //     was measured slower: a persistent launch cannot grow once the others finish.)
        for (;;) {
            int slot = -1;
            for (int i = 0; i < (int)srv->streams.size(); ++i)
                if (!srv->stream_busy[i]) { slot = i; break; }
            if (slot < 0) break;
            auto pick = srv->queued.end();
            for (auto it = srv->queued.begin(); it != srv->queued.end(); ++it)
                if (srv->in_use + (*it)->bytes <= srv->budget && (pick == srv->queued.end() || (*it)->work < (*pick)->work))
                    pick = it;
            if (pick == srv->queued.end()) break;
            std::unique_ptr<Request> q = std::move(*pick);
            srv->queued.erase(pick);
            if (q->prm.method == MSG_METHOD_UNCOUPLED) {
                // exact uncoupled-radical propagation: a few short launches, completed here
                q->result.assign((size_t)q->n_fields * std::max(1, observable_count(q->sys->r, &q->prm)), 0.0);
                try {
                    propagate_uncoupled(ctx, q->sys, &q->prm, q->n_fields, q->fields.data(), q->result.data(), nullptr);
                } catch (const ArgError &e) {
                    q->failed = true;
                    q->error = e.what;
                }
                const int64_t id = q->id;
                srv->finished[id] = std::move(q);
                break;  // one synchronous request per call, so callers collect between them
            }
            q->stream = slot;
            q->whole = max_steps <= 0;
            q->f_begin = 0;
            q->f_count = q->whole ? (int)std::max(1.0, std::min((double)q->n_fields,
                                                                 std::floor(kServerChunkWork * q->n_fields / q->work)))
                                  : q->n_fields;
            try {
                if (q->whole) {
                    launch_part(*q);
                } else {
                    q->batch.reset(new Batch());
                    q->batch->no_split = true;  // requests share the device
                    prepare(*q->batch, ctx, q->sys, &q->prm, q->n_fields, q->fields.data(), nullptr, 0, false, true,
                            srv->streams[slot]);
                }
            } catch (const ArgError &e) {
                q->failed = true;
                q->error = e.what;
                q->batch.reset();
                if (q->done_ev) (void)hipEventDestroy(q->done_ev);
                q->done_ev = nullptr;
                const int64_t id = q->id;
                srv->finished[id] = std::move(q);
                continue;
            }
            srv->stream_busy[slot] = 1;
            srv->in_use += q->bytes;
            srv->active.push_back(std::move(q));
        }
        // 3. chunked mode: one chunk of every resident request, overlapping on the streams
        if (max_steps > 0) {
            for (auto &q : srv->active)
                if (!q->whole) run_chunk(*q->batch, max_steps);
            for (auto it = srv->active.begin(); it != srv->active.end();) {
                if (!(*it)->whole && (*it)->batch->s_next >= (*it)->batch->nsamples - 1) {
                    finish(*it);
                    it = srv->active.erase(it);
                } else {
                    ++it;
                }
            }
        }
        size_t grouped = 0;
        for (auto &g : srv->groups) grouped += g->members.size();
        if (n_active) *n_active = (int32_t)(srv->queued.size() + srv->active.size() + grouped);
    });
}

msg_status msg_server_stats(const msg_server *srv, int64_t *shared_launches, int64_t *shared_requests) {
    if (!srv) return MSG_INVALID_ARG;
    if (shared_launches) *shared_launches = srv->shared_launches;
    if (shared_requests) *shared_requests = srv->shared_requests;
    return MSG_OK;
}

msg_status msg_server_collect(msg_server *srv, int64_t request_id, double *yields, int32_t *done) {
    if (!srv || !done) return MSG_INVALID_ARG;
    auto it = srv->finished.find(request_id);
    if (it == srv->finished.end()) {
        for (auto &q : srv->active)
            if (q->id == request_id) { *done = 0; return MSG_OK; }
        for (auto &q : srv->queued)
            if (q->id == request_id) { *done = 0; return MSG_OK; }
        for (auto &g : srv->groups)
            for (auto &q : g->members)
                if (q->id == request_id) { *done = 0; return MSG_OK; }
        return fail(srv->ctx, MSG_INVALID_ARG, "unknown request id");
    }
    Request &q = *it->second;
    if (q.failed) {
        const std::string e = q.error;
        srv->finished.erase(it);
        return fail(srv->ctx, MSG_NUMERIC_BREAKDOWN, e);
    }
    if (yields) std::copy(q.result.begin(), q.result.end(), yields);
    *done = 1;
    srv->finished.erase(it);
    return MSG_OK;
}

}  // extern "C"
