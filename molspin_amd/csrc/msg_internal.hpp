// Shared host/device definitions of the MI355X propagation backend.
//
// Device basis order ("group-major"): index = nu * 4 + c, where c = 2*a + b is
// the electron-pair index (a, b = 0 for spin up, 1 for down, electron 1 first)
// and nu enumerates the nuclear configuration, nucleus 0 least significant
// (mixed radix; for spin-1/2 nuclei bit k of nu is nucleus k, 0 = up).
// A state column of the matrix-free path is stored as two planes of float4:
//   plane A[nu] = (re c0, im c0, re c1, im c1), plane B[nu] = (re c2, im c2, re c3, im c3),
// so one group is 32 bytes and a column of dimension D occupies 2*D floats.
#pragma once
#include <stdint.h>

namespace msg {

constexpr int kMaxKrylov = 32;
constexpr int kMaxObs = 8;
constexpr int kCorrParts = 4;     // uncoupled method: pair partitions of the correlation sums
constexpr int kMaxNuclei = 16;     // matrix-free path: D = 4 * 2^N <= 262144
constexpr int kMaxPairs = 16;      // matrix-free path: nucleus-nucleus couplings
constexpr int kMaxInitial = 4;
constexpr int kChebMax = 96;
constexpr int kTileGroups = 4096;  // groups per LDS tile of the tiled streaming Lanczos (bits 0-11)
constexpr int kSplitSlab = 2048;   // groups per workgroup slab of the split-column Lanczos (bits 0-10)
constexpr int kDensePlanes = 2;    // dense-path propagator planes: re, im
constexpr int kDenseLdsPlanes = 3; // dense-path state tile in LDS: re, im, re + im
// Device scratch of the allocation-free entry point, per GPU (msg_noalloc.hip)
constexpr unsigned long long kScratchBytes = 64ull << 20;

// Per-item coefficient block of the matrix-free operator (floats):
//   [0..2]  w1: field on electron 1 (H += w1 . S1), [3..5] w2: field on electron 2
//   [6]     scalar shift (identity), [7] pad
//   [8..39] E: 4x4 complex electron-pair matrix (row-major, interleaved), present if flag
//   then per nucleus n (kNucStride floats):
//     [0..2] h1, [3..5] h2, [6] r         diagonal: s_n (h1.S1 + h2.S2 + r), s_n = +1 up / -1 down
//     [7..14] C1 (2x2 complex on e1), [15..22] C2 (2x2 complex on e2), [23..24] t (complex)
//     off-diagonal <up_n| H |down_n> = C1 (x) 1 + 1 (x) C2 + t 1 on the electron pair
//   then per nucleus-nucleus coupling c (kPairStride floats): the coupling as a 4x4 complex
//     matrix on the two nuclei's states (row-major, interleaved; index 2 b1 + b2 with b = 0
//     up, 1 down, b1 the nucleus pair_bits[2 c]), times the identity on the electron pair
constexpr int kCoefHead = 40;
constexpr int kNucStride = 28;
constexpr int kPairStride = 32;
// LDS block table of the matrix-free kernels (float4 entries): per nucleus 8 (its 2x2
// blocks on both electrons in both orientations), then per nucleus-nucleus coupling 8
// (for each of the 4 rows: the diagonal, and the entries of the three partners)
constexpr int kBlkTable = 8 * kMaxNuclei + 8 * kMaxPairs;
// NUC_A1 / NUC_A2: the nucleus's coupling to electron 1 / 2 is axial about z (isotropic
// hyperfine among others): h1 (h2) has no x and y part, and C1 (C2) = [[0, 0], [b, 0]]
// with b = C[2] real.  OP_AXIAL: every electron-nucleus coupling of the system is, and
// the kernels apply the blocks' two real entries instead of the full 2x2 blocks.
enum NucFlags { NUC_E1 = 1, NUC_E2 = 2, NUC_T = 4, NUC_A1 = 8, NUC_A2 = 16 };
enum OpFlags { OP_EREST = 1, OP_AXIAL = 2 };

struct StepArgs {
    // geometry
    int nnuc;          // number of spin-1/2 nuclei (matrix-free path)
    int dim;           // Hilbert dimension D
    int m;             // Krylov order
    int nobs;
    int n_units;       // (item, column) pairs
    int ncols;         // columns per item
    int op_flags;
    const int *nuc_flags;  // [nnuc] device
    int npairs;            // nucleus-nucleus couplings (0: none)
    const int *pair_bits;  // [npairs][2] device: the two nuclei (index bits) of each coupling
    const float *coef; // [n_items][coef_stride]
    int coef_stride;
    int arnoldi;            // 1: Arnoldi on G = -iH - K (unequal recombination rates)
    int rk4;                // 1: RK4 stages of -iH - K instead of the Krylov step (k_rk4)
    int mixm;               // > 0: the last nucleus has this multiplicity (3 or 4) and is the
                            // group index of k_lanczos_mix; the others are spin-1/2
    float kmat[32];         // K on the electron pair, 4x4 complex row-major (Arnoldi)
    const float *dyn_head;  // [n_items][nsamples-1][8] head (w1, w2, shift) of each interval of a
                            // time-dependent operator, or NULL (static: the head of `coef`)
    // dense path
    const float *U;    // [n_items][ustride][kDensePlanes][Dp*Dp] planes re, im (Dp = padded dim)
    int dpad;
    int ustride;       // propagators per item (2 when the last interval is shorter)
    int ulast;         // use the item's second propagator for interval nsamples-2
    int ustep;         // one propagator per interval of the launch: item * ustride + (s - s_begin)
    // observables: electron-pair 4x4 complex matrices (row-major interleaved), device
    const float *obs;        // [nobs][32]
    // time grid
    int nsamples;
    int s_begin, s_end;      // intervals processed by this launch
    const double *dts;       // [nsamples-1]
    const double *wy;        // [nsamples*nobs] yield weights
    const double *wt;        // [nsamples*nobs] trace factors (NULL: no trace)
    double *yacc;            // [n_units][nobs][2] Kahan sum / compensation
    double *trace;           // [n_units][nsamples][nobs] or NULL
    // relaxation (singlet-triplet dephasing of the electron pair)
    int relax;
    uint64_t seed;
    const uint32_t *thr;     // [nsamples-1] flip thresholds of one half-step
    int rng_step0;           // RNG counter of interval s is s + rng_step0
    double ktol;             // > 0: keep a column's Krylov basis while h_{m+1,m} |c_{m-1}| <= ktol
                             // (legacy direct yields' krylovtol; one column per workgroup only)
    // state source / sink
    int init_mode;           // 0 direct (generated), 1 given columns, 2 resume from state
    int n_init;              // direct: number of electron vectors
    const float *evec;       // [n_init][8] device
    int zdim;                // direct: columns per electron vector (nuclear dimension)
    const float *given;      // [ncols][2D] device order
    float *state;            // [n_units][2D] saved state (NULL: not saved)
    double *norm0;           // [n_units] norms for renormalisation (NULL: none)
    int renorm;              // 1: rescale the Krylov result to norm0 (unitary part)
    // scheduling
    int *queue;              // work counter (zeroed before launch)
    float *basis;            // [slots][m][2D]
    int *status;             // device error flags (bit 0: non-finite, bit 1: dt too large,
                             // bit 2: workgroups of a split column lost each other,
                             // bit 3: a pack of the uneven launch names no class or
                             // holds more units than its class packs)
    double *xred;            // split columns: [slots][2][S][17] partial sums
    int *bar;                // split columns: [slots] barrier counters (zeroed before launch)
};

// One persistent launch over radical pairs of uneven Hilbert dimension (k_uneven):
// each pack is up to P columns (matrix-free classes) or 16-column tiles (dense
// classes) of one item of one entry; the class fixes the body and its packing.
enum UnevenClass {
    UC_LZ64 = 0, UC_LZ128, UC_LZ256, UC_LZ512, UC_LZ1024, UC_LZ2048, UC_LZ4096,  // groups per column D / 4
    UC_DENSE = 16,  // + index of the row-tile count in {1, 2, 3, 4, 6, 8, 12, 16, 24, 32}
};
// Columns (matrix-free classes) or 16-column tiles (dense classes) per pack of each
// class: what the kernel's body for the class packs into one workgroup.  Shared by the
// host, which forms the packs, and the kernel, which checks them.
constexpr int uneven_pack_width(int cls) {
    return cls >= UC_LZ64 && cls <= UC_LZ4096 ? (cls <= UC_LZ64 ? 16 : cls == UC_LZ128 ? 8 : cls == UC_LZ256 ? 4
                                                                    : cls <= UC_LZ1024 ? 2 : 1)
         : cls >= UC_DENSE && cls < UC_DENSE + 10 ? (cls < UC_DENSE + 4 ? 4 : cls < UC_DENSE + 6 ? 2 : 1)
                                                  : 0;
}
constexpr int kUnevenThreads = 1024;
constexpr int kUnevenPlanesBytes = 131072;       // per workgroup: the body's LDS planes
constexpr int kUnevenScratchDoubles = 2 * 17 * 16;  // reductions, double-buffered
struct UnevenPack {
    int entry;  // index into UnevenArgs::ents
    int u0;     // first unit of the pack (entry's unit numbering)
    int nact;   // live units in the pack
    int cls;    // UnevenClass
};
struct UnevenArgs {
    const StepArgs *ents;     // [n_entries] device
    const UnevenPack *packs;  // [n_packs] device, in launch order
    int n_packs;
    int *queue;               // pack counter (zeroed before launch)
    float *basis;             // [workgroups][basis_stride] Krylov bases
    size_t basis_stride;      // floats per workgroup
};

}  // namespace msg
