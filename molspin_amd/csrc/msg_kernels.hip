// HIP kernels of the MI355X (gfx950) Hilbert-space propagation backend.
//
//   k_lanczos<G>   matrix-free Krylov propagation of radical-pair states with
//                  spin-1/2 nuclei.  One workgroup owns one state column for a
//                  whole range of time steps (persistent, work queue over
//                  (item, column) units).  The Krylov vector lives in LDS as two
//                  float4 planes; each thread owns G electron-pair groups
//                  nu = g*T + t; the operator is applied matrix-free from
//                  per-item Kronecker factors (electron fields, per-nucleus 2x2
//                  blocks); inner products, the tridiagonal exponential and the
//                  yield accumulation run in FP64.
//   k_dense        D <= 512: U = exp(-i H dt) applied as a complex GEMM on the
//                  FP32 matrix cores (v_mfma_f32_16x16x4f32), one workgroup per
//                  (item, 16-column tile), persistent over the time steps.
//   k_reduce_*     fixed-order reductions of per-column results.
//
// Replaces the serial loops of TaskHSGeneral.cpp:470-576 (observable sampling,
// HSPropagator::StepStochastic HSPropagator.cpp:131-140, KrylovExpmGeneral
// SpinSpace_operators.cpp:1257-1408) for whole batches of items.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "msg_internal.hpp"

namespace msg {

struct C4 {
    float2 c[4];
};

template <int N>
struct Log2 {
    static constexpr int value = 1 + Log2<N / 2>::value;
};
template <>
struct Log2<1> {
    static constexpr int value = 0;
};

__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 acc) {
    acc.x = fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x));
    acc.y = fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y));
    return acc;
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

__device__ __forceinline__ C4 load_group(const float4 *A, const float4 *B, int nu) {
    const float4 p = A[nu], q = B[nu];
    C4 r;
    r.c[0] = make_float2(p.x, p.y);
    r.c[1] = make_float2(p.z, p.w);
    r.c[2] = make_float2(q.x, q.y);
    r.c[3] = make_float2(q.z, q.w);
    return r;
}
__device__ __forceinline__ void store_group(float4 *A, float4 *B, int nu, const C4 &x) {
    A[nu] = make_float4(x.c[0].x, x.c[0].y, x.c[1].x, x.c[1].y);
    B[nu] = make_float4(x.c[2].x, x.c[2].y, x.c[3].x, x.c[3].y);
}
// The same on planes in device memory: the pointers are cast to the global address
// space at the access, so the accesses are global instead of flat instructions (a flat
// access also counts against lgkmcnt, so every wait for an LDS result would wait for
// it as well, and it needs a 64-bit vector address).
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) f4v gf4v;
__device__ __forceinline__ C4 load_group_g(const float4 *A, const float4 *B, int nu) {
    const f4v p = ((const gf4v *)A)[nu], q = ((const gf4v *)B)[nu];
    C4 r;
    r.c[0] = make_float2(p.x, p.y);
    r.c[1] = make_float2(p.z, p.w);
    r.c[2] = make_float2(q.x, q.y);
    r.c[3] = make_float2(q.z, q.w);
    return r;
}
__device__ __forceinline__ float4 load_f4_g(const float *p) {
    const f4v v = *(const gf4v *)p;
    return float4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ void store_group_g(float4 *A, float4 *B, int nu, const C4 &x) {
    ((gf4v *)A)[nu] = f4v{x.c[0].x, x.c[0].y, x.c[1].x, x.c[1].y};
    ((gf4v *)B)[nu] = f4v{x.c[2].x, x.c[2].y, x.c[3].x, x.c[3].y};
}

// y += M (x) 1 on electron 1 (pairs (0,2), (1,3)); M row-major 2x2 complex.
__device__ __forceinline__ void apply_e1(const float2 m[4], const C4 &x, C4 &y) {
    y.c[0] = cfma(m[0], x.c[0], cfma(m[1], x.c[2], y.c[0]));
    y.c[2] = cfma(m[2], x.c[0], cfma(m[3], x.c[2], y.c[2]));
    y.c[1] = cfma(m[0], x.c[1], cfma(m[1], x.c[3], y.c[1]));
    y.c[3] = cfma(m[2], x.c[1], cfma(m[3], x.c[3], y.c[3]));
}
// y += 1 (x) M on electron 2 (pairs (0,1), (2,3)).
__device__ __forceinline__ void apply_e2(const float2 m[4], const C4 &x, C4 &y) {
    y.c[0] = cfma(m[0], x.c[0], cfma(m[1], x.c[1], y.c[0]));
    y.c[1] = cfma(m[2], x.c[0], cfma(m[3], x.c[1], y.c[1]));
    y.c[2] = cfma(m[0], x.c[2], cfma(m[1], x.c[3], y.c[2]));
    y.c[3] = cfma(m[2], x.c[2], cfma(m[3], x.c[3], y.c[3]));
}
// y += (f.S) x on one electron for a real field f (S = sigma/2), the electron's two
// states being x.c[i] and x.c[i + d] for i in {i0, i0 + e}:
// f.S = [[hz, hx - i hy], [hx + i hy, -hz]] with h = f/2 (12 FMAs per pair).
template <int D_, int E_>
__device__ __forceinline__ void apply_field(float fx, float fy, float fz, const C4 &x, C4 &y) {
    const float hx = 0.5f * fx, hy = 0.5f * fy, hz = 0.5f * fz;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int i = k * E_, j = i + D_;
        const float2 a = x.c[i], b = x.c[j];
        y.c[i].x = fmaf(hz, a.x, fmaf(hx, b.x, fmaf(hy, b.y, y.c[i].x)));
        y.c[i].y = fmaf(hz, a.y, fmaf(hx, b.y, fmaf(-hy, b.x, y.c[i].y)));
        y.c[j].x = fmaf(hx, a.x, fmaf(-hy, a.y, fmaf(-hz, b.x, y.c[j].x)));
        y.c[j].y = fmaf(hx, a.y, fmaf(hy, a.x, fmaf(-hz, b.y, y.c[j].y)));
    }
}
// electron 1: pairs (0,2), (1,3); electron 2: pairs (0,1), (2,3)
__device__ __forceinline__ void apply_field_e1(float fx, float fy, float fz, const C4 &x, C4 &y) {
    apply_field<2, 1>(fx, fy, fz, x, y);
}
__device__ __forceinline__ void apply_field_e2(float fx, float fy, float fz, const C4 &x, C4 &y) {
    apply_field<1, 2>(fx, fy, fz, x, y);
}
// y += [[0, pd], [pu, 0]] on one electron (pairs as in apply_field), pu and pd real: the
// flip block of an axial coupling, [[0, 0], [b, 0]] for the nucleus up (pu = b, pd = 0)
// and its transpose for the nucleus down (8 FMAs instead of the full block's 32).
template <int D_, int E_>
__device__ __forceinline__ void apply_axial(float pu, float pd, const C4 &x, C4 &y) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int i = k * E_, j = i + D_;
        y.c[j].x = fmaf(pu, x.c[i].x, y.c[j].x);
        y.c[j].y = fmaf(pu, x.c[i].y, y.c[j].y);
        y.c[i].x = fmaf(pd, x.c[j].x, y.c[i].x);
        y.c[i].y = fmaf(pd, x.c[j].y, y.c[i].y);
    }
}
// Read-only data that is the same for a whole wavefront (an item's operator
// coefficients, the nucleus flags) is read through the constant address space: with
// a wave-uniform address the compiler then loads it with scalar loads into SGPRs,
// which the FMAs take as operands, instead of vector loads into VGPRs.
template <typename T>
using ConstPtr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ ConstPtr<T> as_const(const T *p) {
    return (ConstPtr<T>)p;
}

// A wave-uniform pointer the compiler must treat as unknown at this point (an empty
// asm statement on its scalar registers): the coefficient loads through it are issued
// where they are used, once per Krylov iteration, instead of being hoisted out of the
// iteration loop and held in SGPRs, which spilled SGPRs into VGPRs and those to
// scratch (k_arnoldi<2,512>: 440 -> 172 bytes of scratch per lane, 6 % faster;
// k_lanczos<2,512>: SGPR spills 115 -> 37, 2 % faster).
__device__ __forceinline__ const float *opaque_uniform(const float *p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    asm volatile("" : "+s"(lo), "+s"(hi));
    return reinterpret_cast<const float *>(((uint64_t)hi << 32) | lo);
}

// y += (E + f1.S1 + f2.S2 + r) x on the electron pair, E the item's 4x4 complex
// electron-pair block (row-major, interleaved; index 2 b1 + b2 with b1, b2 the
// electrons' states), f1, f2 real fields: the fields and the shift are added into the
// twelve entries of E they touch and the sum is applied once (64 FMAs instead of 64
// for E plus 56 for the fields and the shift).
template <typename P>
__device__ __forceinline__ void apply_pair_block(P E, float r, float f1x, float f1y, float f1z, float f2x, float f2y,
                                                 float f2z, const C4 &x, C4 &y) {
    const float h1x = 0.5f * f1x, h1y = 0.5f * f1y, h1z = 0.5f * f1z;
    const float h2x = 0.5f * f2x, h2y = 0.5f * f2y, h2z = 0.5f * f2z;
    float2 M[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) M[i][j] = make_float2(E[8 * i + 2 * j], E[8 * i + 2 * j + 1]);
    M[0][0].x += r + h1z + h2z;
    M[1][1].x += r + h1z - h2z;
    M[2][2].x += r - h1z + h2z;
    M[3][3].x += r - h1z - h2z;
    // electron 1 flips (b1), electron 2 spectator: <up|f.S|down> = (hx - i hy)
    M[0][2].x += h1x; M[0][2].y -= h1y; M[1][3].x += h1x; M[1][3].y -= h1y;
    M[2][0].x += h1x; M[2][0].y += h1y; M[3][1].x += h1x; M[3][1].y += h1y;
    // electron 2 flips (b2)
    M[0][1].x += h2x; M[0][1].y -= h2y; M[2][3].x += h2x; M[2][3].y -= h2y;
    M[1][0].x += h2x; M[1][0].y += h2y; M[3][2].x += h2x; M[3][2].y += h2y;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) y.c[i] = cfma(M[i][j], x.c[j], y.c[i]);
}

// Loads a 2x2 complex block, conjugate-transposed if `herm` (lane-selected).
template <typename P>
__device__ __forceinline__ void load_block(P p, bool herm, float2 m[4]) {
    const float2 a = make_float2(p[0], p[1]), b = make_float2(p[2], p[3]);
    const float2 c = make_float2(p[4], p[5]), d = make_float2(p[6], p[7]);
    m[0] = herm ? conjf2(a) : a;
    m[1] = herm ? conjf2(c) : b;
    m[2] = herm ? conjf2(b) : c;
    m[3] = herm ? conjf2(d) : d;
}

// Table in LDS of every nucleus's coupling blocks in both orientations:
// blk[((n 2 + e) 2 + down) 2 + 0..1] holds the block of nucleus n on electron e (e1, e2),
// conjugate-transposed for down = 1.  A lane then reads the block of its own index bit
// instead of selecting it entry by entry.  Filled by the first 4 nnuc threads; the caller
// synchronises the workgroup before the first read.
// The rows of the nucleus-nucleus couplings follow at blk[8 kMaxNuclei + ((c 4 + r) 2 + 0..1)]:
// for row r = 2 b1 + b2 of coupling c, (Re H_rr, H_{r,r^2}, H_{r,r^1}, H_{r,r^3}) with
// the three partners flipping b1, b2 and both.
__device__ __forceinline__ void fill_block_table(const StepArgs &a, const float *co, int t, float4 *blk) {
    if (t < a.nnuc * 4) {
        const int n = t >> 2, e = (t >> 1) & 1;
        float2 m[4];
        load_block(co + kCoefHead + n * kNucStride + (e ? 15 : 7), (t & 1) != 0, m);
        blk[2 * t] = make_float4(m[0].x, m[0].y, m[1].x, m[1].y);
        blk[2 * t + 1] = make_float4(m[2].x, m[2].y, m[3].x, m[3].y);
    }
    if (t < a.npairs * 4) {
        const int c = t >> 2, r = t & 3;
        const float *h = co + kCoefHead + a.nnuc * kNucStride + c * kPairStride + 8 * r;
        blk[8 * kMaxNuclei + 2 * t] = make_float4(h[2 * r], h[2 * (r ^ 2)], h[2 * (r ^ 2) + 1], h[2 * (r ^ 1)]);
        blk[8 * kMaxNuclei + 2 * t + 1] = make_float4(h[2 * (r ^ 1) + 1], h[2 * (r ^ 3)], h[2 * (r ^ 3) + 1], 0.f);
    }
}
__device__ __forceinline__ void table_block(const float4 *blk, int n, int e, bool down, float2 m[4]) {
    const float4 *b = blk + ((n * 2 + e) * 2 + (down ? 1 : 0)) * 2;
    const float4 u = b[0], v = b[1];
    m[0] = make_float2(u.x, u.y);
    m[1] = make_float2(u.z, u.w);
    m[2] = make_float2(v.x, v.y);
    m[3] = make_float2(v.z, v.w);
}

// H x for the group nu whose value is x; neighbours come from the LDS planes.
// H = w1.S1 + w2.S2 + shift + [E] + sum_n (diagonal s_n(...) + flip-flop blocks);
// w1, w2 and the shift come from the head `hd` (the item's block, or the interval's
// head of a time-dependent operator), the rest from the item's block `co`.
// Groups per LDS tile of the tiled streaming kernel: index bits 0-11 of nu.
constexpr int kTileBits = 12;
constexpr int kTileG = 1 << kTileBits;
static_assert(kTileG == kTileGroups, "tile size shared with the host");

// NN >= 0: the number of nuclei is a compile-time constant (the loop unrolls);
// NN < 0: a.nnuc at run time.  TILED: neighbours across index bits below
// kTileBits come from the LDS tile (LA/LB indexed by nu mod kTileG), the others
// from the vector in global memory (GA/GB).
// TBL: the coupling blocks come from the LDS table `blk` (fill_block_table).
// AX: every electron-nucleus coupling is axial (OP_AXIAL), applied in its short form.
template <int NN, bool TILED, bool TBL, bool AX>
__device__ __forceinline__ C4 apply_h_k(const StepArgs &a, const float *__restrict__ co, const float *__restrict__ hd,
                                        int nu, const C4 &x, const float4 *LA, const float4 *LB, const float4 *GA,
                                        const float4 *GB, const float4 *blk) {
    const auto hc = as_const(hd), cc = as_const(co);
    const auto flags = as_const(a.nuc_flags);
    float f1x = hc[0], f1y = hc[1], f1z = hc[2];
    float f2x = hc[3], f2y = hc[4], f2z = hc[5];
    float rr = hc[6];
    C4 y;
#pragma unroll
    for (int c = 0; c < 4; ++c) y.c[c] = make_float2(0.f, 0.f);
    auto nucleus = [&](int n) {
        const auto cn = cc + kCoefHead + n * kNucStride;
        const int fl = flags[n];
        const bool down = (nu >> n) & 1;
        const float s = down ? -1.f : 1.f;
        const C4 nb = !TILED ? load_group(LA, LB, nu ^ (1 << n))
                             : (n < kTileBits ? load_group(LA, LB, (nu ^ (1 << n)) & (kTileG - 1))
                                              : load_group_g(GA, GB, nu ^ (1 << n)));
        float2 m[4];
        if (fl & NUC_E1) {
            if constexpr (AX) {
                f1z = fmaf(s, cn[2], f1z);
                apply_axial<2, 1>(down ? 0.f : cn[11], down ? cn[11] : 0.f, nb, y);
            } else {
                f1x = fmaf(s, cn[0], f1x); f1y = fmaf(s, cn[1], f1y); f1z = fmaf(s, cn[2], f1z);
                if constexpr (TBL) table_block(blk, n, 0, down, m);
                else load_block(cn + 7, down, m);
                apply_e1(m, nb, y);
            }
        }
        if (fl & NUC_E2) {
            if constexpr (AX) {
                f2z = fmaf(s, cn[5], f2z);
                apply_axial<1, 2>(down ? 0.f : cn[19], down ? cn[19] : 0.f, nb, y);
            } else {
                f2x = fmaf(s, cn[3], f2x); f2y = fmaf(s, cn[4], f2y); f2z = fmaf(s, cn[5], f2z);
                if constexpr (TBL) table_block(blk, n, 1, down, m);
                else load_block(cn + 15, down, m);
                apply_e2(m, nb, y);
            }
        }
        if (fl & NUC_T) {
            rr = fmaf(s, cn[6], rr);
            const float2 tt = make_float2(cn[23], down ? -cn[24] : cn[24]);
#pragma unroll
            for (int c = 0; c < 4; ++c) y.c[c] = cfma(tt, nb.c[c], y.c[c]);
        }
    };
    if constexpr (NN >= 0) {
#pragma unroll
        for (int n = 0; n < NN; ++n) nucleus(n);
    } else {
        for (int n = 0; n < a.nnuc; ++n) nucleus(n);
    }
    if (a.op_flags & OP_EREST) {
        apply_pair_block(cc + 8, rr, f1x, f1y, f1z, f2x, f2y, f2z, x, y);
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            y.c[c].x = fmaf(rr, x.c[c].x, y.c[c].x);
            y.c[c].y = fmaf(rr, x.c[c].y, y.c[c].y);
        }
        apply_field_e1(f1x, f1y, f1z, x, y);
        apply_field_e2(f2x, f2y, f2z, x, y);
    }
    return y;
}
template <int NN = -1, bool TILED = false, bool TBL = false>
__device__ __forceinline__ C4 apply_h(const StepArgs &a, const float *__restrict__ co, const float *__restrict__ hd,
                                      int nu, const C4 &x, const float4 *LA, const float4 *LB,
                                      const float4 *GA = nullptr, const float4 *GB = nullptr,
                                      const float4 *blk = nullptr) {
    if (a.op_flags & OP_AXIAL) return apply_h_k<NN, TILED, TBL, true>(a, co, hd, nu, x, LA, LB, GA, GB, blk);
    return apply_h_k<NN, TILED, TBL, false>(a, co, hd, nu, x, LA, LB, GA, GB, blk);
}

// H x for the G groups nu_g = base + g 2^LOGT + t of one thread (t < 2^LOGT, NN
// nuclei).  The index bits below LOGT are the thread's and shared by its G groups, so
// those nuclei's blocks, field sums and shift are formed once per thread; the bits
// from LOGT up are g's and base's, the same for the whole workgroup, so their blocks
// need no lane select.  The workgroup's slab of the vector (SB = LOGT + log2 G index
// bits, local index g 2^LOGT + t) is in LDS (LA, LB); nuclei on bits from SB up
// (a column split over several workgroups, base = slab 2^SB) read their neighbours
// from the whole vector in global memory (GA, GB).
// MIXM > 0: the last nucleus (n = LOGT, NN = LOGT + 1) has multiplicity MIXM = G and is
// the thread's group index g, with magnetic quantum number m_g = (G - 1)/2 - g; its
// diagonal terms scale by 2 m_g and its blocks between g and g + 1 by
// <m_g|I+|m_g - 1> = sqrt(I(I+1) - m_g (m_g - 1)) (Spin.cpp:330-410), applied within
// the thread's registers.
// PART 1 forms y from the workgroup's slab alone: every diagonal term and the flips of
// the nuclei on bits below SB; PART 2 adds the flips of the nuclei on bits from SB up
// (neighbours from GA, GB) to y.  PART 0 does both.
// AX: every electron-nucleus coupling is axial (OP_AXIAL), applied in its short form.
// The neighbours of a nucleus on the group bits are loaded for all G groups before they
// are applied, so their load latencies overlap.
template <int NN, int G, int LOGT, bool TBL, int MIXM, int PART, bool AX>
__device__ __forceinline__ void apply_h_groups_k(const StepArgs &a, const float *__restrict__ co,
                                                 const float *__restrict__ hd, int t, const C4 (&x)[G],
                                                 const float4 *LA, const float4 *LB, C4 (&y)[G], int base,
                                                 const float4 *GA, const float4 *GB, const float4 *blk) {
    constexpr int SB = LOGT + Log2<G>::value;
    const auto hc = as_const(hd), cc = as_const(co);
    const auto flags = as_const(a.nuc_flags);
    // the thread's coupling block of nucleus n on electron e (TBL: from the LDS table)
    auto block = [&](int n, int e, bool down, ConstPtr<float> cb, float2 m[4]) {
        if constexpr (TBL) table_block(blk, n, e, down, m);
        else load_block(cb, down, m);
    };
    float f1x = hc[0], f1y = hc[1], f1z = hc[2];
    float f2x = hc[3], f2y = hc[4], f2z = hc[5];
    float rr = hc[6];
    float e1[G][3], e2[G][3], er[G];  // contributions of the nuclei on g's bits
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if constexpr (PART != 2)
#pragma unroll
            for (int c = 0; c < 4; ++c) y[g].c[c] = make_float2(0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 3; ++k) e1[g][k] = e2[g][k] = 0.f;
        er[g] = 0.f;
    }
#pragma unroll
    for (int n = 0; n < NN; ++n) {
        if (PART == 2 && n < SB) continue;
        const auto cn = cc + kCoefHead + n * kNucStride;
        const int fl = flags[n];
        float2 m[4];
        if (n < LOGT) {
            const bool down = (t >> n) & 1;
            const float s = down ? -1.f : 1.f;
            C4 nb[G];
#pragma unroll
            for (int g = 0; g < G; ++g) nb[g] = load_group(LA, LB, (g << LOGT) + (t ^ (1 << n)));
            if (fl & NUC_E1) {
                if constexpr (AX) {
                    f1z = fmaf(s, cn[2], f1z);
                    const float pu = down ? 0.f : cn[11], pd = down ? cn[11] : 0.f;
#pragma unroll
                    for (int g = 0; g < G; ++g) apply_axial<2, 1>(pu, pd, nb[g], y[g]);
                } else {
                    f1x = fmaf(s, cn[0], f1x); f1y = fmaf(s, cn[1], f1y); f1z = fmaf(s, cn[2], f1z);
                    block(n, 0, down, cn + 7, m);
#pragma unroll
                    for (int g = 0; g < G; ++g) apply_e1(m, nb[g], y[g]);
                }
            }
            if (fl & NUC_E2) {
                if constexpr (AX) {
                    f2z = fmaf(s, cn[5], f2z);
                    const float pu = down ? 0.f : cn[19], pd = down ? cn[19] : 0.f;
#pragma unroll
                    for (int g = 0; g < G; ++g) apply_axial<1, 2>(pu, pd, nb[g], y[g]);
                } else {
                    f2x = fmaf(s, cn[3], f2x); f2y = fmaf(s, cn[4], f2y); f2z = fmaf(s, cn[5], f2z);
                    block(n, 1, down, cn + 15, m);
#pragma unroll
                    for (int g = 0; g < G; ++g) apply_e2(m, nb[g], y[g]);
                }
            }
            if (fl & NUC_T) {
                rr = fmaf(s, cn[6], rr);
                const float2 tt = make_float2(cn[23], down ? -cn[24] : cn[24]);
#pragma unroll
                for (int g = 0; g < G; ++g)
#pragma unroll
                    for (int c = 0; c < 4; ++c) y[g].c[c] = cfma(tt, nb[g].c[c], y[g].c[c]);
            }
        } else if constexpr (MIXM > 0) {
            static_assert(MIXM == G && NN == LOGT + 1, "the nucleus of multiplicity G is the group index");
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const float sg = (float)(G - 1 - 2 * g);  // 2 m_g
                if (fl & NUC_E1) {
                    e1[g][0] = fmaf(sg, cn[0], e1[g][0]); e1[g][1] = fmaf(sg, cn[1], e1[g][1]);
                    e1[g][2] = fmaf(sg, cn[2], e1[g][2]);
                }
                if (fl & NUC_E2) {
                    e2[g][0] = fmaf(sg, cn[3], e2[g][0]); e2[g][1] = fmaf(sg, cn[4], e2[g][1]);
                    e2[g][2] = fmaf(sg, cn[5], e2[g][2]);
                }
                if (fl & NUC_T) er[g] = fmaf(sg, cn[6], er[g]);
            }
#pragma unroll
            for (int g = 0; g + 1 < G; ++g) {
                // <m_g| I+ |m_g - 1> with m_g = (G - 1)/2 - g: sqrt(I(I+1) - m_g (m_g - 1))
                constexpr float ii = 0.25f * (G * G - 1);
                const float mg = 0.5f * (G - 1) - g;
                const float cp = sqrtf(ii - mg * (mg - 1.f));
                float2 up[4], dn[4];
                auto scale = [&](float2 (&q)[4]) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) q[k] = make_float2(cp * q[k].x, cp * q[k].y);
                };
                if (fl & NUC_E1) {
                    load_block(cn + 7, false, up); load_block(cn + 7, true, dn);
                    scale(up); scale(dn);
                    apply_e1(up, x[g + 1], y[g]);
                    apply_e1(dn, x[g], y[g + 1]);
                }
                if (fl & NUC_E2) {
                    load_block(cn + 15, false, up); load_block(cn + 15, true, dn);
                    scale(up); scale(dn);
                    apply_e2(up, x[g + 1], y[g]);
                    apply_e2(dn, x[g], y[g + 1]);
                }
                if (fl & NUC_T) {
                    const float2 tu = make_float2(cp * cn[23], cp * cn[24]), td = make_float2(cp * cn[23], -cp * cn[24]);
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        y[g].c[c] = cfma(tu, x[g + 1].c[c], y[g].c[c]);
                        y[g + 1].c[c] = cfma(td, x[g].c[c], y[g + 1].c[c]);
                    }
                }
            }
        } else {
            const bool diag = PART != 2, flips = PART == 0 || (PART == 1) == (n < SB);
            C4 nbg[G];
            if (flips)
#pragma unroll
                for (int g = 0; g < G; ++g)
                    nbg[g] = n < SB ? load_group(LA, LB, ((g << LOGT) ^ (1 << n)) + t)
                                    : load_group_g(GA, GB, ((base + (g << LOGT)) ^ (1 << n)) + t);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int hi = base + (g << LOGT);  // the group's index bits from LOGT up
                const bool down = (hi >> n) & 1;
                const float s = down ? -1.f : 1.f;
                const C4 &nb = nbg[g];
                if (fl & NUC_E1) {
                    if constexpr (AX) {
                        if (diag) e1[g][2] = fmaf(s, cn[2], e1[g][2]);
                        if (flips) apply_axial<2, 1>(down ? 0.f : cn[11], down ? cn[11] : 0.f, nb, y[g]);
                    } else {
                        if (diag) {
                            e1[g][0] = fmaf(s, cn[0], e1[g][0]); e1[g][1] = fmaf(s, cn[1], e1[g][1]);
                            e1[g][2] = fmaf(s, cn[2], e1[g][2]);
                        }
                        if (flips) {
                            load_block(cn + 7, down, m);
                            apply_e1(m, nb, y[g]);
                        }
                    }
                }
                if (fl & NUC_E2) {
                    if constexpr (AX) {
                        if (diag) e2[g][2] = fmaf(s, cn[5], e2[g][2]);
                        if (flips) apply_axial<1, 2>(down ? 0.f : cn[19], down ? cn[19] : 0.f, nb, y[g]);
                    } else {
                        if (diag) {
                            e2[g][0] = fmaf(s, cn[3], e2[g][0]); e2[g][1] = fmaf(s, cn[4], e2[g][1]);
                            e2[g][2] = fmaf(s, cn[5], e2[g][2]);
                        }
                        if (flips) {
                            load_block(cn + 15, down, m);
                            apply_e2(m, nb, y[g]);
                        }
                    }
                }
                if (fl & NUC_T) {
                    if (diag) er[g] = fmaf(s, cn[6], er[g]);
                    if (flips) {
                        const float2 tt = make_float2(cn[23], down ? -cn[24] : cn[24]);
#pragma unroll
                        for (int c = 0; c < 4; ++c) y[g].c[c] = cfma(tt, nb.c[c], y[g].c[c]);
                    }
                }
            }
        }
    }
    if constexpr (PART == 2) return;
    if constexpr (PART == 0) {
        // nucleus-nucleus couplings (every index bit of the vector is in LDS): row
        // 2 b1 + b2 of the coupling's 4x4 block on the group and its three partners
        const auto pb = as_const(a.pair_bits);
        const float4 *ptab = blk + 8 * kMaxNuclei;
        for (int c = 0; c < a.npairs; ++c) {
            const int k1 = pb[2 * c], k2 = pb[2 * c + 1];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int nl = (g << LOGT) + t, hi = base + nl;
                const int r = (((hi >> k1) & 1) << 1) | ((hi >> k2) & 1);
                const float4 q0 = ptab[(c * 4 + r) * 2], q1 = ptab[(c * 4 + r) * 2 + 1];
                const C4 x1 = load_group(LA, LB, nl ^ (1 << k1)), x2 = load_group(LA, LB, nl ^ (1 << k2)),
                         x12 = load_group(LA, LB, nl ^ (1 << k1) ^ (1 << k2));
                const float2 f1 = make_float2(q0.y, q0.z), f2 = make_float2(q0.w, q1.x), f12 = make_float2(q1.y, q1.z);
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    float2 v = cfma(f1, x1.c[cc], cfma(f2, x2.c[cc], cfma(f12, x12.c[cc], y[g].c[cc])));
                    v.x = fmaf(q0.x, x[g].c[cc].x, v.x);
                    v.y = fmaf(q0.x, x[g].c[cc].y, v.y);
                    y[g].c[cc] = v;
                }
            }
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const float rg = rr + er[g];
        if (a.op_flags & OP_EREST) {
            apply_pair_block(cc + 8, rg, f1x + e1[g][0], f1y + e1[g][1], f1z + e1[g][2], f2x + e2[g][0],
                             f2y + e2[g][1], f2z + e2[g][2], x[g], y[g]);
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                y[g].c[c].x = fmaf(rg, x[g].c[c].x, y[g].c[c].x);
                y[g].c[c].y = fmaf(rg, x[g].c[c].y, y[g].c[c].y);
            }
            apply_field_e1(f1x + e1[g][0], f1y + e1[g][1], f1z + e1[g][2], x[g], y[g]);
            apply_field_e2(f2x + e2[g][0], f2y + e2[g][1], f2z + e2[g][2], x[g], y[g]);
        }
    }
}

template <int NN, int G, int LOGT, bool TBL = false, int MIXM = 0, int PART = 0>
__device__ __forceinline__ void apply_h_groups(const StepArgs &a, const float *__restrict__ co,
                                               const float *__restrict__ hd, int t, const C4 (&x)[G],
                                               const float4 *LA, const float4 *LB, C4 (&y)[G], int base = 0,
                                               const float4 *GA = nullptr, const float4 *GB = nullptr,
                                               const float4 *blk = nullptr) {
    if (a.op_flags & OP_AXIAL)
        apply_h_groups_k<NN, G, LOGT, TBL, MIXM, PART, true>(a, co, hd, t, x, LA, LB, y, base, GA, GB, blk);
    else
        apply_h_groups_k<NN, G, LOGT, TBL, MIXM, PART, false>(a, co, hd, t, x, LA, LB, y, base, GA, GB, blk);
}

// Electron-pair density of a group, accumulated as 16 reals:
// R[0..3] = |x_r|^2, then for r < c the complex x_c conj(x_r) (6 pairs).
__device__ __forceinline__ void accum_rho(const C4 &x, float (&R)[16]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) R[r] = fmaf(x.c[r].x, x.c[r].x, fmaf(x.c[r].y, x.c[r].y, R[r]));
    int k = 4;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = r + 1; c < 4; ++c) {
            // x_c conj(x_r)
            R[k] = fmaf(x.c[c].x, x.c[r].x, fmaf(x.c[c].y, x.c[r].y, R[k]));
            R[k + 1] = fmaf(x.c[c].y, x.c[r].x, fmaf(-x.c[c].x, x.c[r].y, R[k + 1]));
            k += 2;
        }
}
// Re Tr[P rho] = sum_r Re P_rr d_r + 2 sum_{r<c} Re(P_rc z_rc), z_rc = sum x_c conj(x_r).
__device__ __forceinline__ double obs_from_rho(const float *P, const double *R) {
    double v = 0.0;
    for (int r = 0; r < 4; ++r) v += (double)P[8 * r + 2 * r] * R[r];
    int k = 4;
    for (int r = 0; r < 4; ++r)
        for (int c = r + 1; c < 4; ++c) {
            const double pr = P[8 * r + 2 * c], pi = P[8 * r + 2 * c + 1];
            v += 2.0 * (pr * R[k] - pi * R[k + 1]);
            k += 2;
        }
    return v;
}

// Samples every observable of one sample from the block-reduced density and
// accumulates the Kahan-compensated yield and the trace (thread 0 only).
__device__ __forceinline__ void record_sample(const StepArgs &a, int u, int s, const double *R, double *ys,
                                              double *yc) {
    for (int c = 0; c < a.nobs; ++c) {
        const double v = obs_from_rho(a.obs + 32 * c, R);
        const double term = a.wy[(size_t)s * a.nobs + c] * v - yc[c];
        const double nsum = ys[c] + term;
        yc[c] = (nsum - ys[c]) - term;
        ys[c] = nsum;
        if (a.trace) a.trace[((size_t)u * a.nsamples + s) * a.nobs + c] = a.wt[(size_t)s * a.nobs + c] * v;
    }
}

// Sum over the workgroup of K doubles per thread, fixed order; every thread
// receives the totals.  `scr` holds K*16 doubles.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *scr) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) scr[k * 16 + wave] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += scr[k * 16 + w];
        v[k] = s;
    }
}

// Philox4x32-10 (Salmon et al., SC'11), first output word.
__device__ __forceinline__ uint32_t philox_x0(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t key) {
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    return c0;
}

// Singlet-triplet dephasing event: psi -> (P_S - P_T) psi, i.e. the electron
// pair (uu, ud, du, dd) -> (-uu, -du, -ud, -dd) (SpinSpace_relaxation.cpp:1171-1199).
__device__ __forceinline__ void st_flip(C4 &x) {
    const float2 a = x.c[1];
    x.c[0] = make_float2(-x.c[0].x, -x.c[0].y);
    x.c[1] = make_float2(-x.c[2].x, -x.c[2].y);
    x.c[2] = make_float2(-a.x, -a.y);
    x.c[3] = make_float2(-x.c[3].x, -x.c[3].y);
}

// Whether the relaxation half-step `half` of interval s flips the column: the first
// Philox word x of counter (column, step, half, item) against the interval's two
// half-open ranges [t0, t1) and [t2, t3) (a[4 s .. 4 s + 3]): a single channel's
// Bernoulli draw is x < t1 with t0 = 0, an odd count of the superposed Poisson
// process of several channels on the electron pair is the union of two ranges.
__device__ __forceinline__ bool relax_flip(const StepArgs &a, int col, int s, uint32_t half, int item) {
    const uint32_t x = philox_x0((uint32_t)col, (uint32_t)(s + a.rng_step0), half, (uint32_t)item, a.seed);
    const uint32_t *t = a.thr + 4 * (size_t)s;
    return (x >= t[0] && x < t[1]) || (x >= t[2] && x < t[3]);
}

// exp(-i dt T) v for the real symmetric tridiagonal T of order meff and a complex
// vector v, in FP64, by a Chebyshev expansion on the Gershgorin interval of T, summed
// with Clenshaw's recurrence while Miller's backward recurrence generates the Bessel
// coefficients J_k(dt*rho) in the same order.  Every wavefront evaluates it
// redundantly: lane i holds ai = alpha_i, bi = beta_i (coupling of i and i+1) and
// v_i = (vr, vi), and receives component i of the result in (re, im).  Lanes >= meff
// are zero.
__device__ __forceinline__ void cheb_expm_vec(int meff, double dt, double ai_in, double bi_in, double vr_in,
                                              double vi_in, double &re, double &im, int *status) {
    const int i = threadIdx.x & 63;
    const bool on = i < meff;
    const double ai = on ? ai_in : 0.0;
    const double br = (i < meff - 1) ? bi_in : 0.0;
    const double bl = __shfl_up(br, 1, 64) * (i > 0 ? 1.0 : 0.0);
    const double er = on ? vr_in : 0.0, ei = on ? vi_in : 0.0;
    double lo = on ? ai - fabs(bl) - fabs(br) : 1e300;
    double hi = on ? ai + fabs(bl) + fabs(br) : -1e300;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64));
        hi = fmax(hi, __shfl_xor(hi, off, 64));
    }
    const double sigma = 0.5 * (lo + hi);
    const double rho = 0.5 * (hi - lo);
    const double z = dt * rho;
    // (T - sigma) u for a lane-distributed real vector u
    auto tmul = [&](double u) {
        const double ul = __shfl_up(u, 1, 64), ur = __shfl_down(u, 1, 64);
        return (on ? (ai - sigma) * u : 0.0) + (i > 0 ? bl * ul : 0.0) + br * ur;
    };
    double rr, ri;
    if (z < 1e-6) {
        // second order Taylor of exp(-i dt (T - sigma)) v; error < z^3/6
        const double tr = tmul(er), ti = tmul(ei);
        const double t2r = tmul(tr), t2i = tmul(ti);
        rr = er + dt * ti - 0.5 * dt * dt * t2r;
        ri = ei - dt * tr - 0.5 * dt * dt * t2i;
    } else {
        int K = (int)ceil(1.3 * z) + 18;
        if (K > kChebMax) {
            if (i == 0) atomicOr(status, 2);
            K = kChebMax;
        }
        const double xd = on ? (ai - sigma) / rho : 0.0;
        const double xl = bl / rho, xr = br / rho;
        double b1r = 0, b1i = 0, b2r = 0, b2i = 0;
        double jn1 = 0.0, jn = 1e-280, S = 0.0;
        const int N = K + 24;
        for (int k = N; k >= 0; --k) {
            if (k <= K) {
                const double mag = (k == 0 ? 1.0 : 2.0) * jn;  // c_k = (2 - delta_k0) (-i)^k J_k
                double cr = 0.0, ci = 0.0;
                switch (k & 3) {
                    case 0: cr = mag; break;
                    case 1: ci = -mag; break;
                    case 2: cr = -mag; break;
                    default: ci = mag; break;
                }
                const double cvr = cr * er - ci * ei, cvi = cr * ei + ci * er;  // c_k v
                const double lr = __shfl_up(b1r, 1, 64), li = __shfl_up(b1i, 1, 64);
                const double qr = __shfl_down(b1r, 1, 64), qi = __shfl_down(b1i, 1, 64);
                const double xbr = xd * b1r + (i > 0 ? xl * lr : 0.0) + xr * qr;
                const double xbi = xd * b1i + (i > 0 ? xl * li : 0.0) + xr * qi;
                if (k > 0) {  // b_k = c_k v + 2 x b_{k+1} - b_{k+2}
                    const double nr = cvr + 2.0 * xbr - b2r, ni = cvi + 2.0 * xbi - b2i;
                    b2r = b1r; b2i = b1i; b1r = nr; b1i = ni;
                } else {      // f = c_0 v + x b_1 - b_2
                    b1r = cvr + xbr - b2r;
                    b1i = cvi + xbi - b2i;
                }
            }
            S += (k == 0) ? jn : ((k & 1) ? 0.0 : 2.0 * jn);
            if (k > 0) {
                const double jm = (2.0 * k / z) * jn - jn1;  // J_{k-1}
                jn1 = jn;
                jn = jm;
                if (fabs(jn) > 1e200) {
                    jn *= 1e-200; jn1 *= 1e-200; S *= 1e-200;
                    b1r *= 1e-200; b1i *= 1e-200; b2r *= 1e-200; b2i *= 1e-200;
                }
            }
        }
        rr = b1r / S;
        ri = b1i / S;
    }
    // shift phase e^{-i dt sigma}
    const double pr = cos(dt * sigma), pi = -sin(dt * sigma);
    re = on ? rr * pr - ri * pi : 0.0;
    im = on ? rr * pi + ri * pr : 0.0;
}
// exp(-i dt T) e1 (the fresh-basis step of KrylovExpmGeneral / LanczosProcess).
__device__ __forceinline__ void cheb_expm_lanes(int meff, double dt, double ai_in, double bi_in, double &re,
                                                double &im, int *status) {
    cheb_expm_vec(meff, dt, ai_in, bi_in, (threadIdx.x & 63) == 0 ? 1.0 : 0.0, 0.0, re, im, status);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Workgroup-shared state of the propagation bodies, carved by the kernel that calls
// them (one allocation per kernel however many bodies it instantiates).  Per
// subgroup sg (a body may pack several columns or tiles of one item into a
// workgroup, one per subgroup of consecutive waves): Kahan yield sums ys/yc
// [sg * kMaxObs], Krylov combination coefficients coef [sg * kMaxKrylov]; one
// coupling-block table blk per workgroup (every packed column belongs to one item).
struct BodyShared {
    double *ys, *yc;
    float2 *coef;
    float4 *blk;
    double *colsum;  // dense bodies: per-wave column norms, 2 x 16 x 16 doubles
};

// Sum over the W waves of the calling thread's subgroup (waves w0 .. w0 + W - 1)
// of K doubles per thread, in fixed order; every thread of
// the subgroup receives its subgroup's totals.  `scr` holds K x 16 doubles.  With
// W equal to the workgroup's wave count this is block_sum.
template <int K>
__device__ __forceinline__ void group_sum(double (&v)[K], double *scr, int W, int w0) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) scr[k * 16 + wave] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        if (W == 1) {
            s = scr[k * 16 + w0];
        } else {  // W even: two partials per read (ds_read2_b64), summed in wave order
            const double *p = scr + k * 16 + w0;
#pragma unroll 2
            for (int w = 0; w < W; w += 2) {  // not unrolled further: the loads would all go live at once
                const double q0 = p[w], q1 = p[w + 1];
                s += q0;
                s += q1;
            }
        }
        v[k] = s;
    }
}

// After group_sum<2> of (alpha, |w|^2) into `scr`: whether the Lanczos cancellation
// guard fires in any of the P subgroups (each subgroup's sums formed in the order
// group_sum uses, so every subgroup takes the same decision).
__device__ __forceinline__ bool any_guard(const double *scr, int W, int P) {
    bool any = false;
#pragma unroll 1
    for (int q = 0; q < P; ++q) {
        double al = 0.0, ww = 0.0;
#pragma unroll 1
        for (int w = 0; w < W; ++w) {
            al += scr[q * W + w];
            ww += scr[16 + q * W + w];
        }
        any |= !(ww - al * al > 0.02 * ww);
    }
    return any;
}

// Matrix-free Lanczos propagation of P columns of one item, one per subgroup of T
// threads (G groups per thread, D = 4 G T), over the launch's intervals: the
// HSGeneral time loop of those columns (observables at every sample, ST flips,
// KrylovExpmGeneral, SpinSpace_operators.cpp:1257-1408, as symmetric Lanczos).
// Units u0 .. u0 + nact - 1 (unit = item * ncols + column) are live; subgroups
// from nact up repeat unit u0 and discard their results, so every subgroup meets
// every barrier.  LDS: `smem` holds per subgroup the planes of v_j (and v_{j-1}
// with DBUF), `scr` 2 x 256 doubles of reduction scratch.  `bas0`: the workgroup's
// basis store, m vectors per subgroup.
// DBUF: two LDS buffers hold v_j and v_{j-1} (swapped every iteration); without
// DBUF one buffer holds v_j and v_{j-1} is re-read from the basis store.
// DYN: the operator head (electron fields, shift) comes per interval from
// a.dyn_head (time-dependent Zeeman fields); otherwise from the item's block.
// Packed subgroups take every decision that guards a barrier together: a
// cancellation-guard reduction runs when any subgroup needs it (each uses it only
// for itself), and a subgroup whose Krylov space closed early (happy breakdown)
// idles through the remaining iterations.  A column's arithmetic is therefore the
// same whatever P and its neighbours are.
template <int G, int T, int P, bool DBUF, bool DYN, bool MIX = false>
__device__ __forceinline__ void lanczos_columns(const StepArgs &a, int u0, int nact, float *bas0, float4 *smem,
                                                double *scr, const BodyShared &sh) {
    constexpr int ng = G * T, W = T / 64;
    // nuclei: log2(D / 4), or with MIX the spin-1/2 nuclei on the thread index plus the
    // nucleus of multiplicity G on the group index
    constexpr int NNUC = MIX ? Log2<T>::value + 1 : Log2<ng>::value;
    static_assert(T % 64 == 0 && P * T <= 1024 && P <= 16, "subgroups of whole waves");
    const int sg = P == 1 ? 0 : (int)threadIdx.x / T;
    const int t = (int)threadIdx.x - sg * T, lane = threadIdx.x & 63, w0 = sg * W;
    const bool act = sg < nact;
    const int u = u0 + (act ? sg : 0);
    const int item = u / a.ncols, col = u - item * a.ncols;
    const float *co = a.coef + (size_t)item * a.coef_stride;
    float4 *pl = smem + (size_t)sg * (DBUF ? 4 : 2) * ng;
    auto LA = [&](int i) -> float4 * { return pl + ((DBUF && i) ? 2 * ng : 0); };
    auto LB = [&](int i) -> float4 * { return pl + ng + ((DBUF && i) ? 2 * ng : 0); };
    double *ys = sh.ys + sg * kMaxObs, *yc = sh.yc + sg * kMaxObs;
    float2 *s_coef = sh.coef + sg * kMaxKrylov;
    const int D = a.dim;
    const int D4 = D >> 2;
    float *bas = bas0 + (size_t)sg * a.m * 2 * D;
    const int meff0 = a.m < D ? a.m : D;
    int par = 0;
    auto SCR = [&]() { par ^= 1; return scr + par * 256; };
    fill_block_table(a, co, threadIdx.x, sh.blk);  // read after the first barrier below
    __syncthreads();

    // ---- initial state of this column
    C4 psi[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int nu = g * T + t;
        if (a.init_mode == 0) {
            const int which = col / a.zdim, ncol = col - which * a.zdim;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                psi[g].c[c] = (nu == ncol) ? make_float2(a.evec[which * 8 + 2 * c], a.evec[which * 8 + 2 * c + 1])
                                           : make_float2(0.f, 0.f);
        } else {
            const float *src = a.init_mode == 1 ? a.given + (size_t)col * 2 * D : a.state + (size_t)u * 2 * D;
            const float4 *sA = reinterpret_cast<const float4 *>(src);
            psi[g] = load_group_g(sA, sA + D4, nu);
        }
    }
    if (t == 0)
        for (int c = 0; c < a.nobs; ++c) {
            ys[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2] : 0.0;
            yc[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2 + 1] : 0.0;
        }
    const double n0 = a.norm0 ? a.norm0[u] : 0.0;
    // Basis reuse (one column per workgroup): the tridiagonal of the column's current
    // basis (lane j: alpha_j, beta_j), its order, h_{m+1,m}, and (wavefront 0) the
    // coefficients c of the state in that basis, advanced by exp(-i dt T) while kept.
    const bool reuse_on = P == 1 && a.ktol > 0.0 && a.renorm && !a.relax;
    double my_al = 0.0, my_be = 0.0, kc_r = 0.0, kc_i = 0.0, hlast = 0.0;
    int meff = meff0;
    bool have_basis = false;
    __shared__ int s_keep;

    for (int s = a.s_begin;; ++s) {
        const bool end = (s == a.s_end);
        // ---- observables at sample s (HSObservableCollector::Evaluate)
        if (!end || s == a.nsamples - 1) {
            float Rf[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) Rf[k] = 0.f;
#pragma unroll
            for (int g = 0; g < G; ++g) accum_rho(psi[g], Rf);
            double R[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) R[k] = Rf[k];
            group_sum<16>(R, SCR(), W, w0);
            if (t == 0 && act) record_sample(a, u, s, R, ys, yc);
        }
        if (end) break;
        const double dt = a.dts[s];
        const float *hd = DYN ? a.dyn_head + ((size_t)item * (a.nsamples - 1) + s) * 8 : co;
        // ---- relaxation half-step before (HSPropagator.cpp:137)
        if (a.relax && relax_flip(a, col, s, 0u, item))
#pragma unroll
            for (int g = 0; g < G; ++g) st_flip(psi[g]);

        // ---- keep the basis while h_{m+1,m} |c_{m-1}| <= krylovtol (TaskStaticHSDirectYields.cpp:
        // 569-652): psi = |psi| V exp(-i dt T) c, with |psi| = the column's initial norm
        if (reuse_on && have_basis) {
            if (t < 64) {
                const double cl = readlane_f64(sqrt(kc_r * kc_r + kc_i * kc_i), meff0 - 1);
                if (t == 0) s_keep = hlast * (meff0 - 1 < meff ? cl : 0.0) <= a.ktol;
            }
            __syncthreads();
            if (s_keep) {
                if (t < 64) {
                    double nr, ni;
                    cheb_expm_vec(meff, dt, my_al, my_be, kc_r, kc_i, nr, ni, a.status);
                    kc_r = nr;
                    kc_i = ni;
                    if (t < meff) s_coef[t] = make_float2((float)(n0 * nr), (float)(n0 * ni));
                }
                __syncthreads();
#pragma unroll
                for (int g = 0; g < G; ++g)
#pragma unroll
                    for (int c = 0; c < 4; ++c) psi[g].c[c] = make_float2(0.f, 0.f);
                for (int j = 0; j < meff; ++j) {
                    const float2 cj = s_coef[j];
                    const float4 *bA = reinterpret_cast<const float4 *>(bas + (size_t)j * 2 * D);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        const C4 vj = load_group_g(bA, bA + D4, g * T + t);
#pragma unroll
                        for (int c = 0; c < 4; ++c) psi[g].c[c] = cfma(cj, vj.c[c], psi[g].c[c]);
                    }
                }
                __syncthreads();  // s_coef is rewritten at the next step
                continue;
            }
        }
        // ---- Lanczos basis of H from psi (KrylovExpmGeneral, SpinSpace_operators.cpp:1257-1408)
        double nn[1] = {0.0};
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                nn[0] += (double)psi[g].c[c].x * psi[g].c[c].x + (double)psi[g].c[c].y * psi[g].c[c].y;
        group_sum<1>(nn, SCR(), W, w0);
        const double beta0 = sqrt(nn[0]);
        const bool live = beta0 > 0.0;
        bool anylive = live;
        if constexpr (P > 1) anylive = __syncthreads_or(live) != 0;
        if (anylive) {
            {
                const float inv = live ? (float)(1.0 / beta0) : 0.f;
                float4 *bA = reinterpret_cast<float4 *>(bas);
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    C4 v0;
#pragma unroll
                    for (int c = 0; c < 4; ++c) v0.c[c] = make_float2(psi[g].c[c].x * inv, psi[g].c[c].y * inv);
                    const int nu = g * T + t;
                    store_group(LA(0), LB(0), nu, v0);
                    store_group_g(bA, bA + D4, nu, v0);
                }
            }
            my_al = 0.0;  // lane j keeps alpha_j and beta_j
            my_be = 0.0;
            double bprev = 0.0;
            meff = live ? meff0 : 1;
            bool done = !live;
            int cur = 0;
            for (int j = 0; j < meff0; ++j) {
                __syncthreads();  // v_j complete in buffer `cur`
                const float4 *cA = LA(cur), *cB = LB(cur);
                float4 *pA = LA(cur ^ 1), *pB = LB(cur ^ 1);
                const float4 *gA = reinterpret_cast<const float4 *>(bas + (size_t)(j > 0 ? j - 1 : 0) * 2 * D);
                const float fb = (float)bprev;
                // w = H v_j - beta_{j-1} v_{j-1}; one reduction gives alpha and |w|^2
                C4 w[G], vjs[G];
                // the thread's partial <v_j, w> and |w|^2 over its 8 G values in FP32 (two
                // chains each), then FP64 across the workgroup (4 % faster than FP64 per term)
                float rf[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int g = 0; g < G; ++g) vjs[g] = load_group(cA, cB, g * T + t);
                const float *coj = opaque_uniform(co), *hdj = opaque_uniform(hd);
                apply_h_groups<NNUC, G, Log2<T>::value, true, MIX ? G : 0>(a, coj, hdj, t, vjs, cA, cB, w, 0, nullptr,
                                                                           nullptr, sh.blk);
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const int nu = g * T + t;
                    const C4 &vj = vjs[g];
                    if (j > 0) {
                        const C4 vp = DBUF ? load_group(pA, pB, nu) : load_group_g(gA, gA + D4, nu);
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            w[g].c[c].x = fmaf(-fb, vp.c[c].x, w[g].c[c].x);
                            w[g].c[c].y = fmaf(-fb, vp.c[c].y, w[g].c[c].y);
                        }
                    }
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const int h = c & 1;
                        rf[h] = fmaf(vj.c[c].x, w[g].c[c].x, fmaf(vj.c[c].y, w[g].c[c].y, rf[h]));
                        rf[2 + h] = fmaf(w[g].c[c].x, w[g].c[c].x, fmaf(w[g].c[c].y, w[g].c[c].y, rf[2 + h]));
                    }
                }
                double red[2] = {(double)rf[0] + (double)rf[1], (double)rf[2] + (double)rf[3]};
                double *rs = SCR();
                group_sum<2>(red, rs, W, w0);
                const double alpha = red[0];
                if (lane == j && !done) my_al = alpha;
                const float fa = (float)alpha;
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const C4 vj = load_group(cA, cB, g * T + t);
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        w[g].c[c].x = fmaf(-fa, vj.c[c].x, w[g].c[c].x);
                        w[g].c[c].y = fmaf(-fa, vj.c[c].y, w[g].c[c].y);
                    }
                }
                if (j == meff0 - 1) {
                    hlast = sqrt(fmax(red[1] - alpha * alpha, 0.0));  // h_{m+1,m} of the basis-reuse test
                    break;
                }
                // |w - alpha v|^2 = |w|^2 - alpha^2 for orthonormal v; recomputed when it cancels
                double beta2 = red[1] - alpha * alpha;
                const bool need = !(beta2 > 0.02 * red[1]);
                bool anyneed = need;
                if constexpr (P > 1) anyneed = any_guard(rs, W, P);
                if (anyneed) {
                    double bb[1] = {0.0};
#pragma unroll
                    for (int g = 0; g < G; ++g)
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            bb[0] += (double)w[g].c[c].x * w[g].c[c].x + (double)w[g].c[c].y * w[g].c[c].y;
                    group_sum<1>(bb, SCR(), W, w0);
                    if (need) beta2 = bb[0];
                }
                const double beta = sqrt(fmax(beta2, 0.0));
                if constexpr (P == 1) {
                    if (!(beta > 1e-30)) {  // invariant Krylov subspace (happy breakdown)
                        meff = j + 1;
                        hlast = 0.0;
                        break;
                    }
                } else if (!done && !(beta > 1e-30)) {
                    meff = j + 1;
                    done = true;
                }
                if (P == 1 || !done) {
                    if (lane == j) my_be = beta;
                    const float ib = (float)(1.0 / beta);
                    float4 *bA = reinterpret_cast<float4 *>(bas + (size_t)(j + 1) * 2 * D);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) w[g].c[c] = make_float2(w[g].c[c].x * ib, w[g].c[c].y * ib);
                        const int nu = g * T + t;
                        store_group(pA, pB, nu, w[g]);  // v_{j+1} replaces v_{j-1} (DBUF) or v_j (own elements only)
                        store_group_g(bA, bA + D4, nu, w[g]);
                    }
                }
                bprev = beta;
                if (DBUF) cur ^= 1;
            }
            // wavefront 0 of the subgroup evaluates the small exponential; psi = beta V c
            // (beta = norm0 when renormalising the unitary part)
            if (t < 64) {
                double cre, cim;
                cheb_expm_lanes(meff, dt, my_al, my_be, cre, cim, a.status);
                const double sc = live ? (a.renorm ? n0 : beta0) : 0.0;
                if (t < meff) s_coef[t] = make_float2((float)(sc * cre), (float)(sc * cim));
                kc_r = cre;
                kc_i = cim;
            }
            have_basis = live;
            __syncthreads();
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int c = 0; c < 4; ++c) psi[g].c[c] = make_float2(0.f, 0.f);
            for (int j = 0; j < meff; ++j) {
                const float2 cj = s_coef[j];
                const float4 *bA = reinterpret_cast<const float4 *>(bas + (size_t)j * 2 * D);
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const C4 vj = load_group_g(bA, bA + D4, g * T + t);
#pragma unroll
                    for (int c = 0; c < 4; ++c) psi[g].c[c] = cfma(cj, vj.c[c], psi[g].c[c]);
                }
            }
        }
        // ---- relaxation half-step after (HSPropagator.cpp:139)
        if (a.relax && relax_flip(a, col, s, 1u, item))
#pragma unroll
            for (int g = 0; g < G; ++g) st_flip(psi[g]);
    }
    // ---- results of this column
    if (t == 0 && act)
        for (int c = 0; c < a.nobs; ++c) {
            a.yacc[((size_t)u * a.nobs + c) * 2] = ys[c];
            a.yacc[((size_t)u * a.nobs + c) * 2 + 1] = yc[c];
        }
    if (a.state && act) {
        float4 *dA = reinterpret_cast<float4 *>(a.state + (size_t)u * 2 * D);
#pragma unroll
        for (int g = 0; g < G; ++g) store_group_g(dA, dA + D4, g * T + t, psi[g]);
    }
    bool bad = false;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int c = 0; c < 4; ++c) bad |= !isfinite(psi[g].c[c].x) || !isfinite(psi[g].c[c].y);
    if (bad && act) atomicOr(a.status, 1);
}

// One column per workgroup of TMAX threads, persistent over an atomic queue of
// (item, column) units.
template <int G, int TMAX, bool DBUF, bool DYN = false>
__global__ __launch_bounds__(TMAX, G >= 4 ? 2 : 4) void k_lanczos(const StepArgs *__restrict__ pa) {
    const StepArgs &a = *pa;
    extern __shared__ float4 smem[];
    constexpr int ng = G * TMAX;
    double *scr = reinterpret_cast<double *>(smem + (DBUF ? 4 : 2) * ng);  // 2 x 16 x 16 doubles
    __shared__ double ys[kMaxObs], yc[kMaxObs];
    __shared__ float2 s_coef[kMaxKrylov];
    __shared__ float4 s_blk[kBlkTable];
    __shared__ int s_unit;
    const BodyShared sh{ys, yc, s_coef, s_blk, nullptr};
    float *bas = a.basis + (size_t)blockIdx.x * a.m * 2 * a.dim;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(a.queue, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        __syncthreads();
        if (u >= a.n_units) break;
        lanczos_columns<G, TMAX, 1, DBUF, DYN>(a, u, 1, bas, smem, scr, sh);
    }
}

// One nucleus of spin above 1/2 (multiplicity M = 3 or 4) beside spin-1/2 nuclei:
// D = 4 M T, the spin-1/2 nuclei on the thread index (T = 2^n threads), the other
// nucleus on the group index (M groups per thread).
template <int M, int T, bool DBUF>
__global__ __launch_bounds__(T, 1) void k_lanczos_mix(const StepArgs *__restrict__ pa) {
    const StepArgs &a = *pa;
    extern __shared__ float4 smem[];
    constexpr int ng = M * T;
    double *scr = reinterpret_cast<double *>(smem + (DBUF ? 4 : 2) * ng);  // 2 x 16 x 16 doubles
    __shared__ double ys[kMaxObs], yc[kMaxObs];
    __shared__ float2 s_coef[kMaxKrylov];
    __shared__ float4 s_blk[kBlkTable];
    __shared__ int s_unit;
    const BodyShared sh{ys, yc, s_coef, s_blk, nullptr};
    float *bas = a.basis + (size_t)blockIdx.x * a.m * 2 * a.dim;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(a.queue, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        __syncthreads();
        if (u >= a.n_units) break;
        lanczos_columns<M, T, 1, DBUF, false, true>(a, u, 1, bas, smem, scr, sh);
    }
}

// Matrix-free RK4 (HSPropagator::Step with propagationmethod=rk4, HSPropagator.cpp:80-90)
// for one column per workgroup of T threads (G groups per thread, D = 4 G T):
// k1 = A y, k2 = A (y + dt/2 k1), k3 = A (y + dt/2 k2), k4 = A (y + dt k3),
// y <- y + dt/6 (k1 + 2 k2 + 2 k3 + k4) with A = -iH - K (x) 1 (K kept inside, as the
// reference's RK4 does).  Each stage's argument is published to LDS for the operator's
// neighbour reads; the state and the stage sum stay in registers.
template <int G, int T>
__global__ __launch_bounds__(T, G >= 4 ? 1 : 2) void k_rk4(const StepArgs *__restrict__ pa) {
    const StepArgs &a = *pa;
    extern __shared__ float4 smem[];
    constexpr int ng = G * T;
    float4 *LA = smem, *LB = smem + ng;
    double *scr = reinterpret_cast<double *>(smem + 2 * ng);  // 2 x 256 doubles
    __shared__ double ys[kMaxObs], yc[kMaxObs];
    __shared__ float4 s_blk[kBlkTable];
    __shared__ int s_unit;
    const int t = threadIdx.x;
    const int D = a.dim, D4 = D >> 2;
    int par = 0;
    auto SCR = [&]() { par ^= 1; return scr + par * 256; };
    for (;;) {
        if (t == 0) s_unit = atomicAdd(a.queue, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        __syncthreads();
        if (u >= a.n_units) break;
        const int item = u / a.ncols, col = u - item * a.ncols;
        const float *co = a.coef + (size_t)item * a.coef_stride;
        fill_block_table(a, co, t, s_blk);  // read after the first barrier below
        C4 psi[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int nu = g * T + t;
            if (a.init_mode == 0) {
                const int which = col / a.zdim, ncol = col - which * a.zdim;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    psi[g].c[c] = (nu == ncol) ? make_float2(a.evec[which * 8 + 2 * c], a.evec[which * 8 + 2 * c + 1])
                                               : make_float2(0.f, 0.f);
            } else {
                const float *src = a.init_mode == 1 ? a.given + (size_t)col * 2 * D : a.state + (size_t)u * 2 * D;
                const float4 *sA = reinterpret_cast<const float4 *>(src);
                psi[g] = load_group_g(sA, sA + D4, nu);
            }
        }
        if (t == 0)
            for (int c = 0; c < a.nobs; ++c) {
                ys[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2] : 0.0;
                yc[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2 + 1] : 0.0;
            }
        for (int s = a.s_begin;; ++s) {
            const bool end = (s == a.s_end);
            if (!end || s == a.nsamples - 1) {  // observables at sample s
                float Rf[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) Rf[k] = 0.f;
#pragma unroll
                for (int g = 0; g < G; ++g) accum_rho(psi[g], Rf);
                double R[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) R[k] = Rf[k];
                block_sum<16>(R, SCR());
                if (t == 0) record_sample(a, u, s, R, ys, yc);
            }
            if (end) break;
            const float dt = (float)a.dts[s];
            if (a.relax && relax_flip(a, col, s, 0u, item))
#pragma unroll
                for (int g = 0; g < G; ++g) st_flip(psi[g]);
            C4 acc[G], arg[G];
#pragma unroll
            for (int g = 0; g < G; ++g) arg[g] = psi[g];
            for (int st = 0; st < 4; ++st) {
                const float wst = (st == 0 || st == 3) ? 1.f : 2.f, ast = st == 2 ? 1.f : 0.5f;
                __syncthreads();  // the previous stage's neighbour reads are done
#pragma unroll
                for (int g = 0; g < G; ++g) store_group(LA, LB, g * T + t, arg[g]);
                __syncthreads();
                C4 hv[G];
                apply_h_groups<Log2<ng>::value, G, Log2<T>::value, true>(a, co, co, t, arg, LA, LB, hv, 0, nullptr,
                                                                         nullptr, s_blk);
#pragma unroll
                for (int g = 0; g < G; ++g)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {  // k = -i H x - K x
                        float2 kv = make_float2(0.f, 0.f);
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            kv = cfma(make_float2(a.kmat[8 * r + 2 * c], a.kmat[8 * r + 2 * c + 1]), arg[g].c[c], kv);
                        const float2 k = make_float2(hv[g].c[r].y - kv.x, -hv[g].c[r].x - kv.y);
                        acc[g].c[r] = st == 0 ? make_float2(k.x, k.y)
                                              : make_float2(fmaf(wst, k.x, acc[g].c[r].x),
                                                            fmaf(wst, k.y, acc[g].c[r].y));
                        if (st < 3)
                            arg[g].c[r] = make_float2(fmaf(ast * dt, k.x, psi[g].c[r].x),
                                                      fmaf(ast * dt, k.y, psi[g].c[r].y));
                    }
            }
            const float h6 = dt / 6.f;
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    psi[g].c[c] = make_float2(fmaf(h6, acc[g].c[c].x, psi[g].c[c].x), fmaf(h6, acc[g].c[c].y, psi[g].c[c].y));
            if (a.relax && relax_flip(a, col, s, 1u, item))
#pragma unroll
                for (int g = 0; g < G; ++g) st_flip(psi[g]);
        }
        if (t == 0)
            for (int c = 0; c < a.nobs; ++c) {
                a.yacc[((size_t)u * a.nobs + c) * 2] = ys[c];
                a.yacc[((size_t)u * a.nobs + c) * 2 + 1] = yc[c];
            }
        if (a.state) {
            float4 *dA = reinterpret_cast<float4 *>(a.state + (size_t)u * 2 * D);
#pragma unroll
            for (int g = 0; g < G; ++g) store_group_g(dA, dA + D4, g * T + t, psi[g]);
        }
        bool bad = false;
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int c = 0; c < 4; ++c) bad |= !isfinite(psi[g].c[c].x) || !isfinite(psi[g].c[c].y);
        if (bad) atomicOr(a.status, 1);
    }
}

// Matrix-free Arnoldi for the non-Hermitian generator G = -iH - K (x) 1 of unequal
// recombination rates (KrylovExpmGeneral / ArnoldiProcess, SpinSpace_operators.cpp:
// 1257-1408): psi <- beta V_m exp(dt H_m) e1 with the m x m upper Hessenberg H_m of
// modified Gram-Schmidt (FP64 inner products; happy breakdown below 1e-14).  One
// workgroup per (item, column); v_j in LDS for the operator's neighbour reads, the
// basis in the slot's global scratch, w in registers.  Wavefront 0 evaluates
// exp(dt H_m) e1 in FP64 as N applications of the Taylor series of exp(dt H_m / N)
// to the vector, N = ceil(2 ||dt H_m||_1).
template <int G, int TMAX>
__global__ __launch_bounds__(TMAX, 4) void k_arnoldi(const StepArgs *__restrict__ pa) {
    const StepArgs &a = *pa;
    extern __shared__ float4 smem[];
    constexpr int T = TMAX;
    constexpr int ng = G * T;
    const int t = threadIdx.x, lane = t & 63;
    float4 *LA = smem, *LB = smem + ng;
    double *scr = reinterpret_cast<double *>(smem + 2 * ng);  // 2 x 256 doubles
    __shared__ double2 s_h[kMaxKrylov * kMaxKrylov];          // H_m, row-major
    __shared__ float2 s_coef[kMaxKrylov];
    __shared__ double ys[kMaxObs], yc[kMaxObs];
    __shared__ int s_unit;
    __shared__ float4 s_blk[kBlkTable];  // the nucleus-nucleus rows (fill_block_table)
    const int D = a.dim, D4 = D >> 2;
    float *bas = a.basis + (size_t)blockIdx.x * a.m * 2 * D;
    const int meff0 = a.m < D ? a.m : D;
    int par = 0;
    auto SCR = [&]() { par ^= 1; return scr + par * 256; };
    for (;;) {
        if (t == 0) s_unit = atomicAdd(a.queue, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        __syncthreads();
        if (u >= a.n_units) break;
        const int item = u / a.ncols, col = u - item * a.ncols;
        const float *co = a.coef + (size_t)item * a.coef_stride;
        fill_block_table(a, co, t, s_blk);  // read after the workgroup's next barrier
        C4 psi[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int nu = g * T + t;
            if (a.init_mode == 0) {
                const int which = col / a.zdim, ncol = col - which * a.zdim;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    psi[g].c[c] = (nu == ncol) ? make_float2(a.evec[which * 8 + 2 * c], a.evec[which * 8 + 2 * c + 1])
                                               : make_float2(0.f, 0.f);
            } else {
                const float *src = a.init_mode == 1 ? a.given + (size_t)col * 2 * D : a.state + (size_t)u * 2 * D;
                const float4 *sA = reinterpret_cast<const float4 *>(src);
                psi[g] = load_group_g(sA, sA + D4, nu);
            }
        }
        if (t == 0)
            for (int c = 0; c < a.nobs; ++c) {
                ys[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2] : 0.0;
                yc[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2 + 1] : 0.0;
            }
        for (int s = a.s_begin;; ++s) {
            const bool end = (s == a.s_end);
            if (!end || s == a.nsamples - 1) {  // observables at sample s
                float Rf[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) Rf[k] = 0.f;
#pragma unroll
                for (int g = 0; g < G; ++g) accum_rho(psi[g], Rf);
                double R[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) R[k] = Rf[k];
                block_sum<16>(R, SCR());
                if (t == 0) record_sample(a, u, s, R, ys, yc);
            }
            if (end) break;
            const double dt = a.dts[s];
            const float *hd = a.dyn_head ? a.dyn_head + ((size_t)item * (a.nsamples - 1) + s) * 8 : co;
            if (a.relax && relax_flip(a, col, s, 0u, item))
#pragma unroll
                for (int g = 0; g < G; ++g) st_flip(psi[g]);
            double nn[1] = {0.0};
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    nn[0] += (double)psi[g].c[c].x * psi[g].c[c].x + (double)psi[g].c[c].y * psi[g].c[c].y;
            block_sum<1>(nn, SCR());
            const double beta0 = sqrt(nn[0]);
            if (beta0 > 0.0) {
                for (int k = t; k < kMaxKrylov * kMaxKrylov; k += T) s_h[k] = make_double2(0.0, 0.0);
                {
                    const float inv = (float)(1.0 / beta0);
                    float4 *bA = reinterpret_cast<float4 *>(bas);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        C4 v0;
#pragma unroll
                        for (int c = 0; c < 4; ++c) v0.c[c] = make_float2(psi[g].c[c].x * inv, psi[g].c[c].y * inv);
                        const int nu = g * T + t;
                        store_group(LA, LB, nu, v0);
                        store_group_g(bA, bA + D4, nu, v0);
                    }
                }
                int meff = meff0;
                for (int j = 0; j < meff0; ++j) {
                    __syncthreads();  // v_j complete in LDS
                    C4 w[G], vjs[G], hv[G];
#pragma unroll
                    for (int g = 0; g < G; ++g) vjs[g] = load_group(LA, LB, g * T + t);
                    const float *coj = opaque_uniform(co), *hdj = opaque_uniform(hd);
                    apply_h_groups<Log2<ng>::value, G, Log2<T>::value>(a, coj, hdj, t, vjs, LA, LB, hv, 0, nullptr, nullptr,
                                                                       s_blk);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {  // w = -i H v - K v
                            float2 kv = make_float2(0.f, 0.f);
#pragma unroll
                            for (int c = 0; c < 4; ++c)
                                kv = cfma(make_float2(a.kmat[8 * r + 2 * c], a.kmat[8 * r + 2 * c + 1]), vjs[g].c[c], kv);
                            w[g].c[r] = make_float2(hv[g].c[r].y - kv.x, -hv[g].c[r].x - kv.y);
                        }
                    }
                    // modified Gram-Schmidt: h_ij = <v_i, w>, w -= h_ij v_i.  v_j comes from
                    // LDS, earlier basis vectors from the slot's global scratch; each v_i is
                    // read once into registers.  (Requesting v_{i+1} before the reduction of
                    // h_ij spilled registers and measured 6 % slower.)
                    auto basis_group = [&](int i, int nu) {
                        const float4 *gA = reinterpret_cast<const float4 *>(bas + (size_t)i * 2 * D);
                        return i == j ? load_group(LA, LB, nu) : load_group_g(gA, gA + D4, nu);
                    };
                    C4 vi[G];
#pragma unroll
                    for (int g = 0; g < G; ++g) vi[g] = basis_group(0, g * T + t);
                    for (int i = 0; i <= j; ++i) {
                        double h[2] = {0.0, 0.0};
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            // the group's 8 products in FP32, accumulated over groups in FP64
                            float hr0 = 0.f, hi0 = 0.f;
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                hr0 = fmaf(vi[g].c[c].x, w[g].c[c].x, fmaf(vi[g].c[c].y, w[g].c[c].y, hr0));
                                hi0 = fmaf(vi[g].c[c].x, w[g].c[c].y, fmaf(-vi[g].c[c].y, w[g].c[c].x, hi0));
                            }
                            h[0] += (double)hr0;
                            h[1] += (double)hi0;
                        }
                        block_sum<2>(h, SCR());
                        if (t == 0) s_h[i * kMaxKrylov + j] = make_double2(h[0], h[1]);
                        const float hr = (float)h[0], hi = (float)h[1];
#pragma unroll
                        for (int g = 0; g < G; ++g)
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                w[g].c[c].x -= hr * vi[g].c[c].x - hi * vi[g].c[c].y;
                                w[g].c[c].y -= hr * vi[g].c[c].y + hi * vi[g].c[c].x;
                            }
                        if (i + 1 < j) {
                            const float4 *nA = reinterpret_cast<const float4 *>(bas + (size_t)(i + 1) * 2 * D);
#pragma unroll
                            for (int g = 0; g < G; ++g) vi[g] = load_group_g(nA, nA + D4, g * T + t);
                        } else if (i + 1 == j) {
#pragma unroll
                            for (int g = 0; g < G; ++g) vi[g] = load_group(LA, LB, g * T + t);
                        }
                    }
                    if (j == meff0 - 1) break;
                    double wn[1] = {0.0};
#pragma unroll
                    for (int g = 0; g < G; ++g)
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            wn[0] += (double)w[g].c[c].x * w[g].c[c].x + (double)w[g].c[c].y * w[g].c[c].y;
                    block_sum<1>(wn, SCR());
                    const double hn = sqrt(wn[0]);
                    if (!(hn >= 1e-14)) {  // invariant Krylov subspace (happy breakdown)
                        meff = j + 1;
                        break;
                    }
                    if (t == 0) s_h[(j + 1) * kMaxKrylov + j] = make_double2(hn, 0.0);
                    // v_{j+1}: the operator's neighbour reads of v_j finished before the
                    // first reduction above, so the LDS copy can be overwritten
                    const float ib = (float)(1.0 / hn);
                    float4 *bA = reinterpret_cast<float4 *>(bas + (size_t)(j + 1) * 2 * D);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) w[g].c[c] = make_float2(w[g].c[c].x * ib, w[g].c[c].y * ib);
                        const int nu = g * T + t;
                        store_group(LA, LB, nu, w[g]);
                        store_group_g(bA, bA + D4, nu, w[g]);
                    }
                }
                __syncthreads();  // H_m complete
                if (t < 64) {
                    // exp(dt H_m) e1 in FP64: lane r owns component r
                    const int mm = meff;
                    double nrm = 0.0;  // 1-norm of dt H_m (column sums)
                    if (lane < mm)
                        for (int r = 0; r < mm; ++r) {
                            const double2 h = s_h[r * kMaxKrylov + lane];
                            nrm += dt * sqrt(h.x * h.x + h.y * h.y);
                        }
                    for (int off = 32; off > 0; off >>= 1) nrm = fmax(nrm, __shfl_xor(nrm, off, 64));
                    int nrep = (int)ceil(2.0 * nrm);
                    if (nrep < 1) nrep = 1;
                    if (nrep > 4096) {
                        if (lane == 0) atomicOr(a.status, 2);
                        nrep = 4096;
                    }
                    const double sc = dt / nrep;
                    double xr = lane == 0 ? 1.0 : 0.0, xi = 0.0;  // x = e1
                    for (int rep = 0; rep < nrep; ++rep) {
                        double sr = xr, si = xi, tr = xr, ti = xi;  // sum and current term
                        for (int k = 1; k <= 40; ++k) {
                            double nr = 0.0, ni = 0.0;  // (sc H_m term)_lane / k
                            for (int c = 0; c < mm; ++c) {
                                const double cr = __shfl(tr, c, 64), ci = __shfl(ti, c, 64);
                                if (lane < mm) {
                                    const double2 h = s_h[lane * kMaxKrylov + c];
                                    nr += h.x * cr - h.y * ci;
                                    ni += h.x * ci + h.y * cr;
                                }
                            }
                            tr = lane < mm ? nr * sc / k : 0.0;
                            ti = lane < mm ? ni * sc / k : 0.0;
                            sr += tr;
                            si += ti;
                            double tn = fabs(tr) + fabs(ti), sn = fabs(sr) + fabs(si);
                            for (int off = 32; off > 0; off >>= 1) {
                                tn = fmax(tn, __shfl_xor(tn, off, 64));
                                sn = fmax(sn, __shfl_xor(sn, off, 64));
                            }
                            if (tn <= 1e-18 * sn) break;
                        }
                        xr = sr;
                        xi = si;
                    }
                    if (lane < mm) s_coef[lane] = make_float2((float)(beta0 * xr), (float)(beta0 * xi));
                }
                __syncthreads();
#pragma unroll
                for (int g = 0; g < G; ++g)
#pragma unroll
                    for (int c = 0; c < 4; ++c) psi[g].c[c] = make_float2(0.f, 0.f);
                for (int j = 0; j < meff; ++j) {
                    const float2 cj = s_coef[j];
                    const float4 *bA = reinterpret_cast<const float4 *>(bas + (size_t)j * 2 * D);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        const C4 vj = load_group_g(bA, bA + D4, g * T + t);
#pragma unroll
                        for (int c = 0; c < 4; ++c) psi[g].c[c] = cfma(cj, vj.c[c], psi[g].c[c]);
                    }
                }
            }
            if (a.relax && relax_flip(a, col, s, 1u, item))
#pragma unroll
                for (int g = 0; g < G; ++g) st_flip(psi[g]);
        }
        if (t == 0)
            for (int c = 0; c < a.nobs; ++c) {
                a.yacc[((size_t)u * a.nobs + c) * 2] = ys[c];
                a.yacc[((size_t)u * a.nobs + c) * 2 + 1] = yc[c];
            }
        if (a.state) {
            float4 *dA = reinterpret_cast<float4 *>(a.state + (size_t)u * 2 * D);
#pragma unroll
            for (int g = 0; g < G; ++g) store_group_g(dA, dA + D4, g * T + t, psi[g]);
        }
        bool bad = false;
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int c = 0; c < 4; ++c) bad |= !isfinite(psi[g].c[c].x) || !isfinite(psi[g].c[c].y);
        if (bad) atomicOr(a.status, 1);
    }
}

// Matrix-free Lanczos for Hilbert dimensions whose Krylov vector exceeds LDS
// (D >= 16384): one 1024-thread workgroup per column; the state, the residual
// and the Krylov basis stream through global memory (per-slot scratch), and
// the operator gathers neighbour groups from the basis vector in L2/MALL.
// Per iteration: pass A computes w = H v_j - beta v_{j-1} with alpha and |w|^2
// in one reduction; pass B writes v_{j+1} = (w - alpha v_j) / beta.
// LV: v_j is also mirrored in LDS (8 D bytes, D <= 16384) so the matvec's
// neighbour reads stay on chip; the Krylov basis, residual and state stay in
// the per-slot global scratch either way.
// TILED (D >= 32768, LV false): the matvec stages v_j through LDS in tiles of
// kTileG groups, so only the nuclei on index bits >= kTileBits gather from
// global memory.
// (Unrolling the group and nuclear loops at compile time was measured slower
// here: these variants are bound by memory latency, not instruction issue.)
// One column of the streaming Lanczos kernels (1024 threads; unit u = item * ncols +
// column): the time loop with the Krylov basis, residual W and state P in the
// workgroup's global scratch `bas` ((m + 2) vectors), `lvs` the dynamic LDS (the
// mirror of v_j with LV, the tile of TILED), `scr` 2 x 17 x 16 doubles.
template <bool LV, bool TILED>
__device__ __forceinline__ void lanczos_big_column(const StepArgs &a, int u, float *bas, float4 *lvs, double *scr,
                                                   const BodyShared &sh) {
    constexpr int T = 1024;  // launched with exactly 1024 threads
    const int t = threadIdx.x, lane = t & 63;
    const int D = a.dim, D4 = D >> 2, ng = D >> 2;
    float4 *WA = reinterpret_cast<float4 *>(bas + (size_t)a.m * 2 * D), *WB = WA + D4;
    float4 *PA = reinterpret_cast<float4 *>(bas + (size_t)(a.m + 1) * 2 * D), *PB = PA + D4;
    auto BA = [&](int j) { return reinterpret_cast<float4 *>(bas + (size_t)j * 2 * D); };
    const int meff0 = a.m < D ? a.m : D;
    int par = 0;
    auto SCR = [&]() { par ^= 1; return scr + par * 17 * 16; };
    const int item = u / a.ncols, col = u - item * a.ncols;
    const float *co = a.coef + (size_t)item * a.coef_stride;
    fill_block_table(a, co, t, sh.blk);  // read after the workgroup's next barrier
    for (int nu = t; nu < ng; nu += T) {
        C4 x;
        if (a.init_mode == 0) {
            const int which = col / a.zdim, ncol = col - which * a.zdim;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                x.c[c] = (nu == ncol) ? make_float2(a.evec[which * 8 + 2 * c], a.evec[which * 8 + 2 * c + 1])
                                      : make_float2(0.f, 0.f);
        } else {
            const float *src = a.init_mode == 1 ? a.given + (size_t)col * 2 * D : a.state + (size_t)u * 2 * D;
            const float4 *sA = reinterpret_cast<const float4 *>(src);
            x = load_group_g(sA, sA + D4, nu);
        }
        store_group_g(PA, PB, nu, x);
    }
    if (t == 0)
        for (int c = 0; c < a.nobs; ++c) {
            sh.ys[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2] : 0.0;
            sh.yc[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2 + 1] : 0.0;
        }
    const double n0 = a.norm0 ? a.norm0[u] : 0.0;
    for (int s = a.s_begin;; ++s) {
        const bool end = (s == a.s_end);
        const bool flip0 = !end && a.relax &&
                           relax_flip(a, col, s, 0u, item);
        // observables at sample s, then the first relaxation half-step and |psi|^2
        {
            float Rf[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) Rf[k] = 0.f;
            double nn = 0.0;
            for (int nu = t; nu < ng; nu += T) {
                C4 x = load_group_g(PA, PB, nu);
                accum_rho(x, Rf);
                if (flip0) {
                    st_flip(x);
                    store_group_g(PA, PB, nu, x);
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) nn += (double)x.c[c].x * x.c[c].x + (double)x.c[c].y * x.c[c].y;
            }
            double R[17];
#pragma unroll
            for (int k = 0; k < 16; ++k) R[k] = Rf[k];
            R[16] = nn;
            block_sum<17>(R, SCR());
            if (t == 0 && (!end || s == a.nsamples - 1)) record_sample(a, u, s, R, sh.ys, sh.yc);
            if (end) break;
            const double dt = a.dts[s];
            const float *hd = a.dyn_head ? a.dyn_head + ((size_t)item * (a.nsamples - 1) + s) * 8 : co;
            const double beta0 = sqrt(R[16]);
            if (beta0 > 0.0) {
                const float inv = (float)(1.0 / beta0);
                for (int nu = t; nu < ng; nu += T) {
                    C4 x = load_group_g(PA, PB, nu);
#pragma unroll
                    for (int c = 0; c < 4; ++c) x.c[c] = make_float2(x.c[c].x * inv, x.c[c].y * inv);
                    store_group_g(BA(0), BA(0) + D4, nu, x);
                    if (LV) store_group(lvs, lvs + D4, nu, x);
                }
                double my_al = 0.0, my_be = 0.0, bprev = 0.0;
                int meff = meff0;
                for (int j = 0; j < meff0; ++j) {
                    __syncthreads();  // v_j written by the whole workgroup
                    const float4 *vA = LV ? lvs : BA(j), *vB = vA + D4;
                    const float4 *pA = BA(j > 0 ? j - 1 : 0), *pB = pA + D4;
                    const float fb = (float)bprev;
                    double red[2] = {0.0, 0.0};
                    // LV (D = 16384, four groups per thread): the residual stays in
                    // registers; otherwise it goes through the per-slot scratch W
                    C4 wr[LV ? 4 : 1];
                    auto pass_a = [&](int nu, const C4 &vj, const C4 &w0, C4 *keep) {
                        C4 w = w0;
                        if (j > 0) {
                            const C4 vp = load_group_g(pA, pB, nu);
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                w.c[c].x = fmaf(-fb, vp.c[c].x, w.c[c].x);
                                w.c[c].y = fmaf(-fb, vp.c[c].y, w.c[c].y);
                            }
                        }
                        // the group's 8 products in FP32, accumulated over groups in FP64
                        float ra = 0.f, rb = 0.f;
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            ra = fmaf(vj.c[c].x, w.c[c].x, fmaf(vj.c[c].y, w.c[c].y, ra));
                            rb = fmaf(w.c[c].x, w.c[c].x, fmaf(w.c[c].y, w.c[c].y, rb));
                        }
                        red[0] += (double)ra;
                        red[1] += (double)rb;
                        if (keep) *keep = w;
                        else store_group_g(WA, WB, nu, w);
                    };
                    if (TILED) {
                        float4 *tA = lvs, *tB = lvs + kTileG;
                        for (int tb = 0; tb < ng; tb += kTileG) {
                            if (tb > 0) __syncthreads();  // the previous tile is no longer read
                            for (int k = t; k < kTileG; k += T) store_group(tA, tB, k, load_group_g(vA, vB, tb + k));
                            __syncthreads();
                            for (int k = t; k < kTileG; k += T) {
                                const int nu = tb + k;
                                const C4 vj = load_group(tA, tB, k);
                                pass_a(nu, vj, apply_h<-1, true, true>(a, co, hd, nu, vj, tA, tB, vA, vB, sh.blk), nullptr);
                            }
                        }
                    } else if (LV) {
#pragma unroll
                        for (int g = 0; g < (LV ? 4 : 1); ++g) {
                            const int nu = g * T + t;
                            const C4 vj = load_group(vA, vB, nu);
                            pass_a(nu, vj, apply_h<-1, false, true>(a, co, hd, nu, vj, vA, vB, nullptr, nullptr, sh.blk), &wr[g]);
                        }
                    } else {
                        for (int nu = t; nu < ng; nu += T) {
                            const C4 vj = load_group_g(vA, vB, nu);
                            pass_a(nu, vj, apply_h<-1, false, true>(a, co, hd, nu, vj, vA, vB, nullptr, nullptr, sh.blk), nullptr);
                        }
                    }
                    block_sum<2>(red, SCR());
                    const double alpha = red[0];
                    if (lane == j) my_al = alpha;
                    if (j == meff0 - 1) break;
                    const float fa = (float)alpha;
                    double beta2 = red[1] - alpha * alpha;
                    if (!(beta2 > 0.02 * red[1])) {
                        double bb[1] = {0.0};
                        auto acc_bb = [&](const C4 &vj, const C4 &w) {
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                const float wx = fmaf(-fa, vj.c[c].x, w.c[c].x), wy = fmaf(-fa, vj.c[c].y, w.c[c].y);
                                bb[0] += (double)wx * wx + (double)wy * wy;
                            }
                        };
                        if (LV) {
#pragma unroll
                            for (int g = 0; g < (LV ? 4 : 1); ++g) acc_bb(load_group(vA, vB, g * T + t), wr[g]);
                        } else {
                            for (int nu = t; nu < ng; nu += T) acc_bb(load_group_g(vA, vB, nu), load_group_g(WA, WB, nu));
                        }
                        block_sum<1>(bb, SCR());
                        beta2 = bb[0];
                    }
                    const double beta = sqrt(fmax(beta2, 0.0));
                    if (!(beta > 1e-30)) {
                        meff = j + 1;
                        break;
                    }
                    if (lane == j) my_be = beta;
                    const float ib = (float)(1.0 / beta);
                    float4 *nA = BA(j + 1), *nB = nA + D4;
                    auto pass_b = [&](int nu, C4 w) {
                        const C4 vj = LV ? load_group(vA, vB, nu) : load_group_g(vA, vB, nu);
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            w.c[c] = make_float2(fmaf(-fa, vj.c[c].x, w.c[c].x) * ib, fmaf(-fa, vj.c[c].y, w.c[c].y) * ib);
                        store_group_g(nA, nB, nu, w);
                        if (LV) store_group(lvs, lvs + D4, nu, w);  // own elements: read above by this thread only
                    };
                    if (LV) {
#pragma unroll
                        for (int g = 0; g < (LV ? 4 : 1); ++g) pass_b(g * T + t, wr[g]);
                    } else {
                        for (int nu = t; nu < ng; nu += T) pass_b(nu, load_group_g(WA, WB, nu));
                    }
                    bprev = beta;
                }
                if (t < 64) {  // wavefront 0 evaluates the small exponential
                    double cre, cim;
                    cheb_expm_lanes(meff, dt, my_al, my_be, cre, cim, a.status);
                    const double sc = a.renorm ? n0 : beta0;
                    if (t < meff) sh.coef[t] = make_float2((float)(sc * cre), (float)(sc * cim));
                }
                __syncthreads();
                for (int nu = t; nu < ng; nu += T) {
                    C4 x;
#pragma unroll
                    for (int c = 0; c < 4; ++c) x.c[c] = make_float2(0.f, 0.f);
                    for (int j = 0; j < meff; ++j) {
                        const float2 cj = sh.coef[j];
                        const C4 vj = load_group_g(BA(j), BA(j) + D4, nu);
#pragma unroll
                        for (int c = 0; c < 4; ++c) x.c[c] = cfma(cj, vj.c[c], x.c[c]);
                    }
                    if (a.relax && relax_flip(a, col, s, 1u, item))
                        st_flip(x);
                    store_group_g(PA, PB, nu, x);
                }
            } else if (a.relax && relax_flip(a, col, s, 1u, item)) {
                for (int nu = t; nu < ng; nu += T) {
                    C4 x = load_group_g(PA, PB, nu);
                    st_flip(x);
                    store_group_g(PA, PB, nu, x);
                }
            }
            __syncthreads();  // psi complete before the next sample
        }
    }
    if (t == 0)
        for (int c = 0; c < a.nobs; ++c) {
            a.yacc[((size_t)u * a.nobs + c) * 2] = sh.ys[c];
            a.yacc[((size_t)u * a.nobs + c) * 2 + 1] = sh.yc[c];
        }
    bool bad = false;
    for (int nu = t; nu < ng; nu += T) {
        const C4 x = load_group_g(PA, PB, nu);
        if (a.state) {
            float4 *dA = reinterpret_cast<float4 *>(a.state + (size_t)u * 2 * D);
            store_group_g(dA, dA + D4, nu, x);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) bad |= !isfinite(x.c[c].x) || !isfinite(x.c[c].y);
    }
    if (bad) atomicOr(a.status, 1);
    __syncthreads();
}

template <bool LV, bool TILED = false>
__global__ __launch_bounds__(1024) void k_lanczos_big(const StepArgs *__restrict__ pa) {
    const StepArgs &a = *pa;
    extern __shared__ float4 lvs[];
    __shared__ double scr[2 * 17 * 16];  // block_sum of up to 17 values, double-buffered
    __shared__ double ys[kMaxObs], yc[kMaxObs];
    __shared__ float2 s_coef[kMaxKrylov];
    __shared__ float4 s_blk[kBlkTable];  // coupling blocks, both orientations
    __shared__ int s_unit;
    const BodyShared sh{ys, yc, s_coef, s_blk, nullptr};
    // per-slot scratch: basis (m vectors), residual W, state P
    float *bas = a.basis + (size_t)blockIdx.x * (a.m + 2) * 2 * a.dim;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(a.queue, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        __syncthreads();
        if (u >= a.n_units) break;
        lanczos_big_column<LV, TILED>(a, u, bas, lvs, scr, sh);
    }
}

// Norms of the initial columns (renormalisation targets).
__global__ void k_init_norms(const StepArgs *__restrict__ pa) {
    const StepArgs &a = *pa;
    // one wave per unit; initial columns follow init_mode as in k_lanczos
    const int u = blockIdx.x;
    if (u >= a.n_units) return;
    const int item = u / a.ncols, col = u - item * a.ncols;
    (void)item;
    const int D = a.dim;
    double acc = 0.0;
    if (a.init_mode == 0) {
        const int which = col / a.zdim;
        if (threadIdx.x == 0)
            for (int k = 0; k < 8; ++k) acc += (double)a.evec[which * 8 + k] * a.evec[which * 8 + k];
    } else {
        const float *src = a.init_mode == 1 ? a.given + (size_t)col * 2 * D : a.state + (size_t)u * 2 * D;
        for (int i = threadIdx.x; i < 2 * D; i += blockDim.x) acc += (double)src[i] * src[i];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (threadIdx.x == 0) a.norm0[u] = sqrt(acc);
}

__global__ void k_zero_int(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// out[item][k] = sum over the item's columns of in[unit][k] (unit = item*ncols + col),
// columns summed in order.  `in_stride` doubles per unit, `k` picks every
// `pick`-th entry (Kahan pairs store sum at even offsets).
__global__ void k_reduce_columns(const double *in, double *out, int n_items, int ncols, int nk, int in_stride,
                                 int pick) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items * nk) return;
    const int item = i / nk, k = i - item * nk;
    double s = 0.0, comp = 0.0;
    for (int c = 0; c < ncols; ++c) {
        const double x = in[(size_t)(item * ncols + c) * in_stride + (size_t)k * pick] -
                         (pick == 2 ? in[(size_t)(item * ncols + c) * in_stride + (size_t)k * pick + 1] : 0.0);
        const double y = x - comp;
        const double ns = s + y;
        comp = (ns - s) - y;
        s = ns;
    }
    out[i] = s;
}

// ---------------------------------------------------------------------------
// Matrix-free Lanczos with each column split over S = D / 8192 workgroups of T
// threads (D = 4 2^NN >= 16384).  A workgroup owns a slab of 2048 groups (the low
// index bits), G = 2048 / T per thread (2 at T = 1024 with four waves per SIMD and
// 128 VGPRs, 4 at T = 512 with two waves per SIMD and 256 VGPRs), and keeps the slab
// of v_j and v_{j-1} in LDS (128 KiB: one workgroup per CU); the residual stays in
// registers.  The nuclei on the higher bits couple slabs: their neighbours come from
// the slot's global vectors, complete at the slot's barrier.  Inner
// products are reduced over the workgroup, then over the slot's workgroups in slab
// order (slot_sum), so every workgroup of a slot sees the same alpha, beta and
// observables and follows the same control flow.  Per state-step HBM carries the m
// basis stores, the partner slabs of each v_j (L2 / MALL) and the m reads of the
// final combination.  All workgroups of a launch are resident at once (cooperative
// launch of a multiple of 8 slots); a barrier wait that does not end sets status
// bit 2 and ends the launch.
// ---------------------------------------------------------------------------

// Barrier over the S workgroups of a slot: counter `bar` counts arrivals over the
// launch, `nbar` the barriers this workgroup has passed.  Every thread calls it;
// false when the wait gave up (status bit 2).  The workgroups of a slot run on one
// XCD (blockIdx.x mod 8 selects the XCD), whose L2 they share: every wave waits for
// its own stores to reach L2 (L1 writes through), and after the arrival count is
// complete one thread invalidates the CU's L1 for the whole workgroup.
// The barrier in two halves: slot_arrive counts this workgroup in once its stores are
// in L2; slot_wait waits for the whole slot.  Work between the halves that reads only
// the workgroup's own data overlaps the other workgroups' arrival.
__device__ __forceinline__ void slot_arrive(int *bar) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's global stores are in L2
    __syncthreads();
    typedef __attribute__((address_space(1))) int gint;
    if (threadIdx.x == 0) __hip_atomic_fetch_add((gint *)bar, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool slot_wait(int *bar, int S, int &nbar, int *status, int *s_flag) {
    if (threadIdx.x == 0) {
        const int target = S * (nbar + 1);
        int ok = 1;
        typedef __attribute__((address_space(1))) int gint;
        for (int spins = 0; __hip_atomic_load((gint *)bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; ++spins) {
            if (spins > (1 << 26)) {  // about a second: a workgroup of the slot is gone
                ok = 0;
                atomicOr(status, 4);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        asm volatile("buffer_inv sc0" ::: "memory");  // drop this CU's stale L1 lines
        *s_flag = ok;
    }
    ++nbar;
    __syncthreads();
    return *s_flag != 0;
}
__device__ __forceinline__ bool slot_barrier(int *bar, int S, int &nbar, int *status, int *s_flag) {
    slot_arrive(bar);
    return slot_wait(bar, S, nbar, status, s_flag);
}

// Sums K workgroup totals (every thread holds them) over the slot's S workgroups in
// slab order, without an arrival counter: each workgroup stores its K totals as 64-bit
// words whose low 8 bits carry the sum's tag (1 ... 255 in turn; each total keeps 44
// mantissa bits, a relative change below 6e-14), one single-copy-atomic store per word
// (thread 0 stores them all: selecting a thread's word by its index put the totals in
// scratch), and wavefront 0 of every workgroup polls the slot's words in L2 until every tag
// matches, then adds them in slab order, so every workgroup gets the same bits.  The
// sum is complete one L2 round trip after the last word lands (the counter barrier took
// a store, a wait for it, an atomic, the poll, an L1 invalidation and a second read).
// xr holds 2 x S x 17 words, zeroed before the launch and used alternately: a workgroup
// rewrites a half two sums later, after every workgroup has passed the sum in between
// and so has read it; the stale words in a half carry a tag two steps back (255 is odd).
// `nsum` counts the sums this workgroup has passed.  Every thread calls it; false when
// the wait gave up (status bit 2).
template <int K>
__device__ __forceinline__ bool slot_sum(double (&v)[K], uint64_t *xr, int S, int slab, int &nsum, int *status,
                                         int *s_flag, double *s_tot) {
    constexpr uint64_t kTagBits = 0xFF;
    const uint64_t tag = 1 + (uint64_t)(nsum % 255);
    typedef __attribute__((address_space(1))) uint64_t guint64;  // global, not flat, accesses
    guint64 *buf = (guint64 *)(xr + (size_t)(nsum & 1) * S * 17);
    ++nsum;
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k)
            __hip_atomic_store(buf + slab * 17 + k, ((uint64_t)__double_as_longlong(v[k]) & ~kTagBits) | tag,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x < 64) {
        const int l = threadIdx.x;
        const bool on = l < S;  // lane q polls the words of slab q
        // the few words of the per-iteration sums stay in registers from the poll; the
        // 17 of the per-step sum are read once more after it (fewer live registers)
        constexpr bool keep = K <= 2;
        uint64_t w[keep ? K : 1];
        int ok = 1;
        for (int spins = 0;; ++spins) {
            bool good = true;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint64_t x = on ? __hip_atomic_load(buf + l * 17 + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : tag;
                if constexpr (keep) w[k] = x;
                good = good && (x & kTagBits) == tag;
            }
            if (__all(good)) break;
            if (spins > (1 << 26)) {  // about a second: a workgroup of the slot is gone
                ok = 0;
                if (l == 0) atomicOr(status, 4);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            uint64_t x;
            if constexpr (keep) x = w[k];
            else x = on ? __hip_atomic_load(buf + l * 17 + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            const double d = on ? __longlong_as_double((long long)(x & ~kTagBits)) : 0.0;
            double sum = 0.0;
            for (int q = 0; q < S; ++q) sum += __shfl(d, q, 64);
            if (l == 0) s_tot[k] = sum;
        }
        if (l == 0) *s_flag = ok;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = s_tot[k];
    return *s_flag != 0;
}

// AX: the system's electron-nucleus couplings are all axial (OP_AXIAL); each instance
// inlines one operator form (both forms in one kernel pushed the residual's registers to
// scratch).
template <int NN, int T, bool AX>
__global__ __launch_bounds__(T, T >= 1024 ? 4 : 2) void k_lanczos_split(const StepArgs *__restrict__ pa) {
    const StepArgs &a = *pa;
    constexpr int SG = kSplitSlab, G = SG / T, SB = Log2<SG>::value, S = 1 << (NN - SB);
    extern __shared__ float4 smem[];  // slab of v_j and v_{j-1}: planes A and B of each buffer
    auto LA = [&](int i) -> float4 * { return smem + (i ? 2 * SG : 0); };
    auto LB = [&](int i) -> float4 * { return smem + SG + (i ? 2 * SG : 0); };
    __shared__ double scr[2 * 17 * 16];  // block_sum of up to 17 values, double-buffered
    __shared__ double s_tot[17];
    __shared__ double ys[kMaxObs], yc[kMaxObs];
    __shared__ float2 s_coef[kMaxKrylov];
    __shared__ int s_flag;
    const int t = threadIdx.x, lane = t & 63;
    // block b runs on XCD b mod 8; the S workgroups of a slot share one XCD (and L2)
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int slab = local % S, slot = (local / S) * 8 + xcd, nslots = gridDim.x / S;
    const int base = slab << SB;  // first group of the slab
    const int D = a.dim, D4 = D >> 2;
    float *bas = a.basis + (size_t)slot * a.m * 2 * D;
    uint64_t *xr = reinterpret_cast<uint64_t *>(a.xred) + (size_t)slot * 2 * S * 17;
    int *bar = a.bar + slot;
    const bool lead = slab == 0 && t == 0;  // records samples and yields
    const int meff0 = a.m < D ? a.m : D;
    int nbar = 0, nsum = 0, par = 0;
    auto SCR = [&]() { par ^= 1; return scr + par * 17 * 16; };
    auto norm2 = [&](const C4 (&x)[G]) {
        double nn = 0.0;
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int c = 0; c < 4; ++c) nn += (double)x[g].c[c].x * x[g].c[c].x + (double)x[g].c[c].y * x[g].c[c].y;
        return nn;
    };

    for (int u = slot; u < a.n_units; u += nslots) {
        const int item = u / a.ncols, col = u - item * a.ncols;
        const float *co = a.coef + (size_t)item * a.coef_stride;
        C4 psi[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int nu = base + g * T + t;
            if (a.init_mode == 0) {
                const int which = col / a.zdim, ncol = col - which * a.zdim;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    psi[g].c[c] = (nu == ncol) ? make_float2(a.evec[which * 8 + 2 * c], a.evec[which * 8 + 2 * c + 1])
                                               : make_float2(0.f, 0.f);
            } else {
                const float *src = a.init_mode == 1 ? a.given + (size_t)col * 2 * D : a.state + (size_t)u * 2 * D;
                const float4 *sA = reinterpret_cast<const float4 *>(src);
                psi[g] = load_group_g(sA, sA + D4, nu);
            }
        }
        if (lead)
            for (int c = 0; c < a.nobs; ++c) {
                ys[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2] : 0.0;
                yc[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2 + 1] : 0.0;
            }
        const double n0 = a.norm0 ? a.norm0[u] : 0.0;

        for (int s = a.s_begin;; ++s) {
            const bool end = (s == a.s_end);
            // observables at sample s and |psi|^2, summed over the slot
            double R[17];
            {
                float Rf[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) Rf[k] = 0.f;
#pragma unroll
                for (int g = 0; g < G; ++g) accum_rho(psi[g], Rf);
#pragma unroll
                for (int k = 0; k < 16; ++k) R[k] = Rf[k];
                R[16] = norm2(psi);
                block_sum<17>(R, SCR());
                if (!slot_sum<17>(R, xr, S, slab, nsum, a.status, &s_flag, s_tot)) return;
                if (lead && (!end || s == a.nsamples - 1)) record_sample(a, u, s, R, ys, yc);
            }
            if (end) break;
            const double dt = a.dts[s];
            const float *hd = a.dyn_head ? a.dyn_head + ((size_t)item * (a.nsamples - 1) + s) * 8 : co;
            double nn = R[16];
            // relaxation half-step before (HSPropagator.cpp:137); the same decision in
            // every workgroup of the slot
            if (a.relax && relax_flip(a, col, s, 0u, item)) {
#pragma unroll
                for (int g = 0; g < G; ++g) st_flip(psi[g]);
                double q[1] = {norm2(psi)};
                block_sum<1>(q, SCR());
                if (!slot_sum<1>(q, xr, S, slab, nsum, a.status, &s_flag, s_tot)) return;
                nn = q[0];
            }
            const double beta0 = sqrt(nn);
            if (beta0 > 0.0) {
                {
                    const float inv = (float)(1.0 / beta0);
                    float4 *bA = reinterpret_cast<float4 *>(bas);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        C4 v0;
#pragma unroll
                        for (int c = 0; c < 4; ++c) v0.c[c] = make_float2(psi[g].c[c].x * inv, psi[g].c[c].y * inv);
                        store_group(LA(0), LB(0), g * T + t, v0);
                        store_group_g(bA, bA + D4, base + g * T + t, v0);
                    }
                }
                slot_arrive(bar);  // v_0 stored; the slot waits before the first cross-slab read
                double my_al = 0.0, my_be = 0.0;  // lane j keeps alpha_j and beta_j
                double bprev = 0.0;
                int meff = meff0;
                int cur = 0;
                for (int j = 0; j < meff0; ++j) {
                    const float4 *cA = LA(cur), *cB = LB(cur);
                    float4 *pA = LA(cur ^ 1), *pB = LB(cur ^ 1);
                    const float4 *gA = reinterpret_cast<const float4 *>(bas + (size_t)j * 2 * D);
                    const float fb = (float)bprev;
                    // w = H v_j - beta_{j-1} v_{j-1}; one reduction gives alpha and |w|^2
                    C4 w[G], vjs[G];
                    double red[2] = {0.0, 0.0};
#pragma unroll
                    for (int g = 0; g < G; ++g) vjs[g] = load_group(cA, cB, g * T + t);
                    // the slab's part of H v_j while the slot's other workgroups store
                    // theirs, then the nuclei across slabs once v_j is complete in the slot
                    // (lane-selected blocks: the LDS block table measured 8 % slower here);
                    // coefficients reloaded per iteration (5 % faster at D = 65536: 115 ms
                    // against 122 ms per launch)
                    const float *coj = opaque_uniform(co), *hdj = opaque_uniform(hd);
                    apply_h_groups_k<NN, G, Log2<T>::value, false, 0, 1, AX>(a, coj, hdj, t, vjs, cA, cB, w, base,
                                                                             nullptr, nullptr, nullptr);
                    // - beta_{j-1} v_{j-1} from the own slab, also before the wait
                    if (j > 0)
#pragma unroll
                        for (int g = 0; g < G; ++g) {
                            const C4 vp = load_group(pA, pB, g * T + t);
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                w[g].c[c].x = fmaf(-fb, vp.c[c].x, w[g].c[c].x);
                                w[g].c[c].y = fmaf(-fb, vp.c[c].y, w[g].c[c].y);
                            }
                        }
                    if (!slot_wait(bar, S, nbar, a.status, &s_flag)) return;
                    apply_h_groups_k<NN, G, Log2<T>::value, false, 0, 2, AX>(a, coj, hdj, t, vjs, cA, cB, w, base,
                                                                             gA, gA + D4, nullptr);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        // the group's 8 products in FP32, accumulated over groups in FP64
                        float ra = 0.f, rb = 0.f;
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            ra = fmaf(vjs[g].c[c].x, w[g].c[c].x, fmaf(vjs[g].c[c].y, w[g].c[c].y, ra));
                            rb = fmaf(w[g].c[c].x, w[g].c[c].x, fmaf(w[g].c[c].y, w[g].c[c].y, rb));
                        }
                        red[0] += (double)ra;
                        red[1] += (double)rb;
                    }
                    block_sum<2>(red, SCR());
                    if (!slot_sum<2>(red, xr, S, slab, nsum, a.status, &s_flag, s_tot)) return;
                    const double alpha = red[0];
                    if (lane == j) my_al = alpha;
                    const float fa = (float)alpha;
#pragma unroll
                    for (int g = 0; g < G; ++g)
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            w[g].c[c].x = fmaf(-fa, vjs[g].c[c].x, w[g].c[c].x);
                            w[g].c[c].y = fmaf(-fa, vjs[g].c[c].y, w[g].c[c].y);
                        }
                    if (j == meff0 - 1) break;
                    // |w - alpha v|^2 = |w|^2 - alpha^2 for orthonormal v; recomputed when it cancels
                    double beta2 = red[1] - alpha * alpha;
                    if (!(beta2 > 0.02 * red[1])) {
                        double bb[1] = {norm2(w)};
                        block_sum<1>(bb, SCR());
                        if (!slot_sum<1>(bb, xr, S, slab, nsum, a.status, &s_flag, s_tot)) return;
                        beta2 = bb[0];
                    }
                    const double beta = sqrt(fmax(beta2, 0.0));
                    if (!(beta > 1e-30)) {  // invariant Krylov subspace (happy breakdown)
                        meff = j + 1;
                        break;
                    }
                    if (lane == j) my_be = beta;
                    const float ib = (float)(1.0 / beta);
                    float4 *bA = reinterpret_cast<float4 *>(bas + (size_t)(j + 1) * 2 * D);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) w[g].c[c] = make_float2(w[g].c[c].x * ib, w[g].c[c].y * ib);
                        store_group(pA, pB, g * T + t, w[g]);  // v_{j+1} replaces v_{j-1}
                        store_group_g(bA, bA + D4, base + g * T + t, w[g]);
                    }
                    bprev = beta;
                    cur ^= 1;
                    slot_arrive(bar);  // v_{j+1} stored
                }
                // wavefront 0 evaluates the small exponential (the same in every
                // workgroup of the slot); psi = beta V c over the own slab
                if (t < 64) {
                    double cre, cim;
                    cheb_expm_lanes(meff, dt, my_al, my_be, cre, cim, a.status);
                    const double sc = a.renorm ? n0 : beta0;
                    if (t < meff) s_coef[t] = make_float2((float)(sc * cre), (float)(sc * cim));
                }
                __syncthreads();
#pragma unroll
                for (int g = 0; g < G; ++g)
#pragma unroll
                    for (int c = 0; c < 4; ++c) psi[g].c[c] = make_float2(0.f, 0.f);
                for (int j = 0; j < meff; ++j) {
                    const float2 cj = s_coef[j];
                    const float4 *bA = reinterpret_cast<const float4 *>(bas + (size_t)j * 2 * D);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        const C4 vj = load_group_g(bA, bA + D4, base + g * T + t);
#pragma unroll
                        for (int c = 0; c < 4; ++c) psi[g].c[c] = cfma(cj, vj.c[c], psi[g].c[c]);
                    }
                }
            }
            // relaxation half-step after (HSPropagator.cpp:139)
            if (a.relax && relax_flip(a, col, s, 1u, item))
#pragma unroll
                for (int g = 0; g < G; ++g) st_flip(psi[g]);
        }
        if (lead)
            for (int c = 0; c < a.nobs; ++c) {
                a.yacc[((size_t)u * a.nobs + c) * 2] = ys[c];
                a.yacc[((size_t)u * a.nobs + c) * 2 + 1] = yc[c];
            }
        if (a.state) {
            float4 *dA = reinterpret_cast<float4 *>(a.state + (size_t)u * 2 * D);
#pragma unroll
            for (int g = 0; g < G; ++g) store_group_g(dA, dA + D4, base + g * T + t, psi[g]);
        }
        bool bad = false;
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int c = 0; c < 4; ++c) bad |= !isfinite(psi[g].c[c].x) || !isfinite(psi[g].c[c].y);
        if (bad) atomicOr(a.status, 1);
    }
}

// ---------------------------------------------------------------------------
// Dense path: B <- U B on the FP32 matrix cores.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));

// Element (row, col) of the initial state of a dense-path unit.
__device__ __forceinline__ float2 dense_init(const StepArgs &a, int u, int col, int row) {
    const int D = a.dim;
    if (row >= D || col >= a.ncols) return make_float2(0.f, 0.f);
    if (a.init_mode == 0) {
        const int which = col / a.zdim, ncol = col - which * a.zdim;
        const int nu = row >> 2, c = row & 3;
        return nu == ncol ? make_float2(a.evec[which * 8 + 2 * c], a.evec[which * 8 + 2 * c + 1]) : make_float2(0.f, 0.f);
    }
    if (a.init_mode == 1) {
        const float *src = a.given + (size_t)col * 2 * D;
        const int nu = row >> 2, c = row & 3;
        const float *p = c < 2 ? src + 4 * nu + 2 * c : src + D + 4 * nu + 2 * (c - 2);
        return make_float2(p[0], p[1]);
    }
    const float *st = a.state + (size_t)u * 2 * a.dpad * 16;
    const int tcol = col & 15;
    return make_float2(st[row * 16 + tcol], st[a.dpad * 16 + row * 16 + tcol]);
}

// Waves per workgroup of the dense kernel: 8 from 12 row tiles up (at most four row tiles
// per wave keep the three products' accumulators in registers), else 4.
constexpr int dense_waves(int NT) { return NT >= 12 ? 8 : 4; }

// Dense-path propagation of P 16-column tiles of one item, one per subgroup of NW
// waves (unit u = item * tiles_per_item + tile; units u0 .. u0 + nact - 1 live,
// subgroups from nact up repeat u0 and discard their results).  NT: row tiles of 16
// (padded dimension Dp = 16 NT, a compile-time constant so the product loop and the
// per-wave tile loops unroll); wave w of a subgroup owns row tiles w + NW k.  `dsm`
// holds per subgroup the B tile planes (re, im, re + im), `scr` 2 x 256 doubles.
template <int NT, int NW, int P>
__device__ __forceinline__ void dense_tiles(const StepArgs &a, int u0, int nact, float *dsm, double *scr,
                                            const BodyShared &sh) {
    constexpr int Dp = 16 * NT;
    constexpr int KT = (NT + NW - 1) / NW;  // row tiles per wave
    static_assert(NW * P <= 16, "at most 16 waves per workgroup");
    const int sg = P == 1 ? 0 : (int)threadIdx.x / (64 * NW);
    const int t = (int)threadIdx.x - sg * 64 * NW, lane = t & 63, wave = t >> 6, gwave = threadIdx.x >> 6;
    const int w0 = sg * NW;
    const bool act = sg < nact;
    const int u = u0 + (act ? sg : 0);
    float *Bre = dsm + (size_t)sg * kDenseLdsPlanes * Dp * 16;
    float *Bim = Bre + Dp * 16, *Bsm = Bre + 2 * Dp * 16;  // B: re, im, re + im
    double *ys = sh.ys + sg * kMaxObs, *yc = sh.yc + sg * kMaxObs;
    const int q = lane >> 4, cl = lane & 15;
    constexpr int ntile = NT;
    const int tiles_per_item = (a.ncols + 15) >> 4;
    int par = 0, cpar = 0;
    auto SCR = [&]() { par ^= 1; return scr + par * 256; };
    // squared norm of this lane's column: lanes cl + 16 q of every wave of the subgroup hold its rows
    auto colnorm2 = [&](const floatx4 *xr_, const floatx4 *xi_) -> double {
        float p = 0.f;
#pragma unroll
        for (int k = 0; k < KT; ++k)
#pragma unroll
            for (int i = 0; i < 4; ++i) p = fmaf(xr_[k][i], xr_[k][i], fmaf(xi_[k][i], xi_[k][i], p));
        double d = p;
        d += __shfl_xor(d, 16, 64);
        d += __shfl_xor(d, 32, 64);
        cpar ^= 1;
        if (q == 0) sh.colsum[cpar * 256 + gwave * 16 + cl] = d;
        __syncthreads();
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) sum += sh.colsum[cpar * 256 + (w0 + w) * 16 + cl];
        return sum;
    };
    const int item = u / tiles_per_item, tile = u - item * tiles_per_item;
    const int col = tile * 16 + cl;  // the column this lane's accumulators hold
    // register copy of the state in MFMA output layout
    floatx4 xr[KT], xi[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const int rt = wave + NW * k;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float2 v = rt < ntile ? dense_init(a, u, col, rt * 16 + 4 * q + i) : make_float2(0.f, 0.f);
            xr[k][i] = v.x;
            xi[k][i] = v.y;
        }
    }
    if (t == 0)
        for (int c = 0; c < a.nobs; ++c) {
            ys[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2] : 0.0;
            yc[c] = a.s_begin > 0 ? a.yacc[((size_t)u * a.nobs + c) * 2 + 1] : 0.0;
        }
    const bool colok = col < a.ncols;
    // target norm of the column for the renormalised unitary part (symmetric recombination)
    const double cn0 = a.renorm ? sqrt(colnorm2(xr, xi)) : 0.0;
    for (int s = a.s_begin;; ++s) {
        const bool end = s == a.s_end;
        if (!end || s == a.nsamples - 1) {
            float Rf[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) Rf[k] = 0.f;
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                C4 g;
#pragma unroll
                for (int i = 0; i < 4; ++i) g.c[i] = make_float2(xr[k][i], xi[k][i]);
                accum_rho(g, Rf);
            }
            double R[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) R[k] = Rf[k];
            group_sum<16>(R, SCR(), NW, w0);
            if (t == 0 && act) record_sample(a, u, s, R, ys, yc);
        }
        if (end) break;
        if (a.relax && colok && relax_flip(a, col, s, 0u, item))
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                C4 g;
#pragma unroll
                for (int i = 0; i < 4; ++i) g.c[i] = make_float2(xr[k][i], xi[k][i]);
                st_flip(g);
#pragma unroll
                for (int i = 0; i < 4; ++i) { xr[k][i] = g.c[i].x; xi[k][i] = g.c[i].y; }
            }
        // publish B to LDS ([row][16 columns] planes)
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const int rt = wave + NW * k;
            if (rt < ntile)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    Bre[(rt * 16 + 4 * q + i) * 16 + cl] = xr[k][i];
                    Bim[(rt * 16 + 4 * q + i) * 16 + cl] = xi[k][i];
                    Bsm[(rt * 16 + 4 * q + i) * 16 + cl] = xr[k][i] + xi[k][i];
                }
        }
        __syncthreads();
        // B' = U B as three real products on the matrix cores (Gauss):
        // P1 = Ur Br, P2 = Ui Bi, P3 = (Ur + Ui)(Br + Bi); re = P1 - P2, im = P3 - P1 - P2
        const int uw = a.ustep ? s - a.s_begin : ((a.ulast && s == a.nsamples - 2) ? 1 : 0);
        const float *Ur = a.U + ((size_t)item * a.ustride + uw) * kDensePlanes * Dp * Dp;
        const float *Ui = Ur + (size_t)Dp * Dp;
        floatx4 x2[KT];  // P2 (xr holds P1, xi P3)
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            xr[k] = floatx4{0.f, 0.f, 0.f, 0.f};
            xi[k] = floatx4{0.f, 0.f, 0.f, 0.f};
            x2[k] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
        // k-groups of 16: one float4 per lane and plane carries four k-steps of A;
        // the A operands of group g + 1 are fetched while group g multiplies
        float4 nr4[KT], ni4[KT];
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const int rt = wave + NW * k;
            const size_t o = ((size_t)(rt * ntile) * 64 + lane) * 4;
            nr4[k] = rt < ntile ? load_f4_g(Ur + o) : float4{0.f, 0.f, 0.f, 0.f};
            ni4[k] = rt < ntile ? load_f4_g(Ui + o) : float4{0.f, 0.f, 0.f, 0.f};
        }
        for (int g = 0; g < ntile; ++g) {
            float br[4], bi[4], bs[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int kr = 16 * g + 4 * v + q;
                br[v] = Bre[kr * 16 + cl];
                bi[v] = Bim[kr * 16 + cl];
                bs[v] = Bsm[kr * 16 + cl];
            }
            float4 cr4[KT], ci4[KT];
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                cr4[k] = nr4[k];
                ci4[k] = ni4[k];
                const int rt = wave + NW * k;
                if (rt < ntile && g + 1 < ntile) {
                    const size_t o = ((size_t)(rt * ntile + g + 1) * 64 + lane) * 4;
                    nr4[k] = load_f4_g(Ur + o);  // global loads (no measurable change against flat here)
                    ni4[k] = load_f4_g(Ui + o);
                }
            }
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                const int rt = wave + NW * k;
                if (rt < ntile) {
                    const float4 ar4 = cr4[k], ai4 = ci4[k];
                    const float ar[4] = {ar4.x, ar4.y, ar4.z, ar4.w}, ai[4] = {ai4.x, ai4.y, ai4.z, ai4.w};
                    const float as[4] = {ar4.x + ai4.x, ar4.y + ai4.y, ar4.z + ai4.z, ar4.w + ai4.w};
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        xr[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[v], br[v], xr[k], 0, 0, 0);
                        x2[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[v], bi[v], x2[k], 0, 0, 0);
                        xi[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(as[v], bs[v], xi[k], 0, 0, 0);
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            xi[k] = xi[k] - xr[k] - x2[k];
            xr[k] = xr[k] - x2[k];
        }
        if (a.renorm) {
            const double n2 = colnorm2(xr, xi);
            const float f = n2 > 0.0 ? (float)(cn0 / sqrt(n2)) : 0.f;
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                xr[k] *= f;
                xi[k] *= f;
            }
        }
        if (a.relax && colok && relax_flip(a, col, s, 1u, item))
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                C4 g;
#pragma unroll
                for (int i = 0; i < 4; ++i) g.c[i] = make_float2(xr[k][i], xi[k][i]);
                st_flip(g);
#pragma unroll
                for (int i = 0; i < 4; ++i) { xr[k][i] = g.c[i].x; xi[k][i] = g.c[i].y; }
            }
    }
    if (t == 0 && act)
        for (int c = 0; c < a.nobs; ++c) {
            a.yacc[((size_t)u * a.nobs + c) * 2] = ys[c];
            a.yacc[((size_t)u * a.nobs + c) * 2 + 1] = yc[c];
        }
    bool bad = false;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const int rt = wave + NW * k;
        if (rt < ntile)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                bad |= !isfinite(xr[k][i]) || !isfinite(xi[k][i]);
                if (a.state && act) {
                    float *st = a.state + (size_t)u * 2 * Dp * 16;
                    st[(rt * 16 + 4 * q + i) * 16 + cl] = xr[k][i];
                    st[Dp * 16 + (rt * 16 + 4 * q + i) * 16 + cl] = xi[k][i];
                }
            }
    }
    if (bad && act) atomicOr(a.status, 1);
}

// One tile per workgroup of dense_waves(NT) waves, persistent over an atomic queue
// of (item, tile) units.
template <int NT>
__global__ __launch_bounds__(64 * dense_waves(NT)) void k_dense(const StepArgs *__restrict__ pa) {
    const StepArgs &a = *pa;
    extern __shared__ float dsm[];
    __shared__ double ys[kMaxObs], yc[kMaxObs];  // Kahan yield sums (thread 0)
    __shared__ double s_col[2 * 256];             // per-wave column norms (double-buffered)
    __shared__ int s_unit;
    double *scr = reinterpret_cast<double *>(dsm + kDenseLdsPlanes * 16 * NT * 16);
    const BodyShared sh{ys, yc, nullptr, nullptr, s_col};
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(a.queue, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        __syncthreads();
        if (u >= a.n_units) break;
        dense_tiles<NT, dense_waves(NT), 1>(a, u, 1, dsm, scr, sh);
    }
}

// ---------------------------------------------------------------------------
// Device FP64 matrix exponential for the dense path: exp(X) of a batch of n x n
// complex matrices by a degree-18 Taylor polynomial (Paterson-Stockmeyer, the
// products on rocBLAS) of X / 2^s followed by s squarings.  Row-major double2.
// ---------------------------------------------------------------------------

// out_b = c0 I + c1 X_b + c2 X2_b + c3 X3_b (a block of the Paterson-Stockmeyer
// evaluation of the Taylor polynomial); batch entries of n x n, contiguous.
__global__ void k_zlincomb(const double2 *__restrict__ X, const double2 *__restrict__ X2,
                           const double2 *__restrict__ X3, double2 *__restrict__ out, int n, size_t total, double c0,
                           double c1, double c2, double c3) {
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t e = idx % ((size_t)n * n);
        const bool diag = (e / n) == (e % n);
        const double2 x = X[idx], x2 = X2[idx], x3 = X3[idx];
        out[idx] = make_double2((diag ? c0 : 0.0) + c1 * x.x + c2 * x2.x + c3 * x3.x, c1 * x.y + c2 * x2.y + c3 * x3.y);
    }
}

// ---------------------------------------------------------------------------
// Infinite-time symmetric-decay yields (TaskStaticHSSymmetricDecay.cpp:114-158)
// from the eigenvectors of H: with M = V^dagger (row-major), Pt = M P M^dagger and
// Rt = M rho0 M^dagger,  Phi = sum_gh Re(Pt_gh conj(Rt_gh)) k^2 / (k^2 + (l_g - l_h)^2).
// ---------------------------------------------------------------------------
struct Block4 {
    double2 m[16];  // 4x4 electron-pair matrix, row-major
};

// Y = s * M (B (x) 1) in the device basis (index = nu * 4 + c): each row of M is
// multiplied on the right by the block-diagonal matrix.
__global__ void k_blockdiag_right(const double2 *__restrict__ M, Block4 B, double scale, double2 *__restrict__ Y,
                                  int n) {
    const size_t total = (size_t)n * n;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t row = idx / n;
        const int col = (int)(idx - row * n), c = col & 3, base = col - c;
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const double2 x = M[row * n + base + a], b = B.m[a * 4 + c];
            acc.x = fma(x.x, b.x, fma(-x.y, b.y, acc.x));
            acc.y = fma(x.x, b.y, fma(x.y, b.x, acc.y));
        }
        Y[idx] = make_double2(scale * acc.x, scale * acc.y);
    }
}

// rows[g] = sum_h (a Re z_gh + disp w_gh Im z_gh) sc / (a^2 + w_gh^2) with
// z_gh = conj(Pt_gh) Rt_gh and w_gh = lam_g - lam_h; one workgroup per row.
//  * symmetric decay: a = sc = k, disp = 0, i.e. Re z k^2 / (k^2 + w^2);
//  * time-infinity integral of Tr[P rho(t)] under the decay e^{-a t}: a = 2 kappa,
//    sc = 1, disp = 1, i.e. Re(z / (a + i w)).
__global__ __launch_bounds__(256) void k_lorentz_rows(const double2 *__restrict__ Pt, const double2 *__restrict__ Rt,
                                                      const double *__restrict__ lam, double a, double disp, double sc,
                                                      int n, double *__restrict__ rows) {
    __shared__ double red[256];
    const int g = blockIdx.x, t = threadIdx.x;
    const double lg = lam[g], a2 = a * a;
    double acc = 0.0;
    for (int h = t; h < n; h += blockDim.x) {
        const double2 p = Pt[(size_t)g * n + h], r = Rt[(size_t)g * n + h];
        const double w = lg - lam[h];
        const double re = p.x * r.x + p.y * r.y, im = p.x * r.y - p.y * r.x;
        acc += (a * re + disp * w * im) * (sc / (a2 + w * w));
    }
    red[t] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) rows[g] = red[0];
}

// out[0] = sum of rows[0..n) in a fixed order (one workgroup).
__global__ __launch_bounds__(256) void k_sum_rows(const double *__restrict__ rows, int n, double *__restrict__ out) {
    __shared__ double red[256];
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int i = t; i < n; i += blockDim.x) acc += rows[i];
    red[t] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) out[0] = red[0];
}

// ---------------------------------------------------------------------------
// Time-dependent generators of the dense path (dynamics=dynamic):
//   out_b = dt_b (G[item_b] + sum_k coef_b[k] Z_k)
// with G the static generator of each item and Z_k the time-dependent operators
// (-i S_a of the spins with time-dependent fields); n x n row-major, contiguous.
// ---------------------------------------------------------------------------
__global__ void k_dyn_generator(const double2 *__restrict__ G, const double2 *__restrict__ Z,
                                const double *__restrict__ coef, const int *__restrict__ item,
                                const double *__restrict__ dt, int nz, int n, size_t total,
                                double2 *__restrict__ out) {
    const size_t nn = (size_t)n * n;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t b = idx / nn, e = idx - b * nn;
        double2 acc = G[(size_t)item[b] * nn + e];
        for (int k = 0; k < nz; ++k) {
            const double c = coef[b * nz + k];
            const double2 z = Z[(size_t)k * nn + e];
            acc.x = fma(c, z.x, acc.x);
            acc.y = fma(c, z.y, acc.y);
        }
        const double h = dt[b];
        out[idx] = make_double2(h * acc.x, h * acc.y);
    }
}

// ---------------------------------------------------------------------------
// Time-infinity integral X = int_0^inf e^{Gt} rho0 e^{G^dagger t} dt by doubling
// (batches of n x n row-major matrices, contiguous):
//   X_tau = tau S_0 with S_n = rho0 + tau/(n+2) L(S_{n+1}), L(S) = G S + (G S)^dagger,
//   then X <- X + U X U^dagger and U <- U^2 (U = e^{G tau}) until U has decayed.
// ---------------------------------------------------------------------------

// S = s (R + c (Y + Y^dagger)) for every batch entry; Y = G S_prev (Hermitian S).
// One 32 x 32 output tile per workgroup (grid (tiles, tiles, batch), 32 x 8 threads);
// the mirrored tile of Y is staged through LDS so that both reads are coalesced.
__global__ __launch_bounds__(256) void k_herm_update(const double2 *__restrict__ Y, const double2 *__restrict__ R,
                                                     double2 *__restrict__ S, int n, double c, double s) {
    __shared__ double2 tile[32][33];
    const size_t off = (size_t)blockIdx.z * n * n;
    const int bi = blockIdx.y * 32, bj = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {  // tile[r][x] = Y[bj + r][bi + x]
        const int row = bj + r, col = bi + tx;
        if (row < n && col < n) tile[r][tx] = Y[off + (size_t)row * n + col];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = bi + r, j = bj + tx;
        if (i < n && j < n) {
            const size_t idx = off + (size_t)i * n + j;
            const double2 y = Y[idx], yt = tile[tx][r], q = R[idx];  // yt = Y[j][i]
            S[idx] = make_double2(s * (q.x + c * (y.x + yt.x)), s * (q.y + c * (y.y - yt.y)));
        }
    }
}

// One doubling accumulation W_b = X_b + Delta_b, with the partial Frobenius sums of
// Delta_b and X_b over slab p of each matrix: part[2 (b P + p)] = sum |Delta|^2,
// part[2 (b P + p) + 1] = sum |X|^2, P = gridDim.x.  Grid (P, batch); fixed
// reduction order, so the sums do not depend on scheduling.
__global__ __launch_bounds__(256) void k_accumulate_fro(const double2 *__restrict__ X, const double2 *__restrict__ Dl,
                                                        double2 *__restrict__ W, size_t nn, double *__restrict__ part) {
    __shared__ double rd[256], rx[256];
    const size_t P = gridDim.x, p = blockIdx.x, b = blockIdx.y;
    const size_t lo = nn * p / P, hi = nn * (p + 1) / P, off = b * nn;
    const int t = threadIdx.x;
    double sd = 0.0, sx = 0.0;
    for (size_t i = lo + t; i < hi; i += blockDim.x) {
        const double2 x = X[off + i], d = Dl[off + i];
        sd += d.x * d.x + d.y * d.y;
        sx += x.x * x.x + x.y * x.y;
        W[off + i] = make_double2(x.x + d.x, x.y + d.y);
    }
    rd[t] = sd;
    rx[t] = sx;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) {
            rd[t] += rd[t + w];
            rx[t] += rx[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        part[2 * (b * P + p)] = rd[0];
        part[2 * (b * P + p) + 1] = rx[0];
    }
}

// Electron-pair block of the partial trace over the nuclei (device index 4 nu + c):
// out[b][a][c] = sum_nu X_b[4 nu + a][4 nu + c]; grid (batch, 16).
__global__ __launch_bounds__(256) void k_pair_block_trace(const double2 *__restrict__ X, int n,
                                                          double2 *__restrict__ out) {
    __shared__ double rx[256], ry[256];
    const int b = blockIdx.x, ac = blockIdx.y, a = ac >> 2, c = ac & 3, t = threadIdx.x;
    const double2 *x = X + (size_t)b * n * n;
    double sx = 0.0, sy = 0.0;
    for (int nu = t; nu < n / 4; nu += blockDim.x) {
        const double2 v = x[(size_t)(4 * nu + a) * n + 4 * nu + c];
        sx += v.x;
        sy += v.y;
    }
    rx[t] = sx;
    ry[t] = sy;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) {
            rx[t] += rx[t + w];
            ry[t] += ry[t + w];
        }
        __syncthreads();
    }
    if (t == 0) out[(size_t)b * 16 + ac] = make_double2(rx[0], ry[0]);
}

// ---------------------------------------------------------------------------
// Uncoupled radicals (H = H_A (x) 1 + 1 (x) H_B, no electron-electron coupling):
// per radical, the electron-spin correlation functions
//   C_{mu,alpha}(t) = Tr[(sigma_mu (x) 1) U(t) (sigma_alpha (x) 1) U(t)^dagger]
// from the eigenpairs of the radical's Hamiltonian (M = V^dagger, row-major):
//   S_mu = M (sigma_mu (x) 1) M^dagger,  C(t) = sum_gh S_mu[g,h] conj(S_alpha[g,h]) e^{i(l_g - l_h) t}.
// ---------------------------------------------------------------------------

// Y = M (sigma_mu (x) 1_Z) for every batch entry; basis index e * Z + nu (electron first).
__global__ void k_pauli_right(const double2 *__restrict__ M, double2 *__restrict__ Y, int n, int mu, size_t total) {
    const int Z = n >> 1;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t rowbase = idx - idx % n;
        const int col = (int)(idx % n), e = col / Z, nu = col - e * Z;
        const double2 m0 = M[rowbase + nu], m1 = M[rowbase + Z + nu];  // electron up / down columns
        double2 v;
        if (mu == 0) v = e == 0 ? m1 : m0;                                               // sigma_x
        else if (mu == 1) v = e == 0 ? make_double2(-m1.y, m1.x) : make_double2(m0.y, -m0.x);  // sigma_y
        else v = e == 0 ? m0 : make_double2(-m1.x, -m1.y);                               // sigma_z
        Y[idx] = v;
    }
}

constexpr int kCorrSamples = 8;   // samples per lane in k_corr
constexpr int kCorrThreads = 64;  // lanes per k_corr workgroup

// Partial correlations C_p[item][s][3 mu + alpha] of pair partition p for the samples
// times[0..ns): the pairs g <= h, in row-major order, are split into gridDim.y equal
// ranges; every lane owns kCorrSamples consecutive samples and runs over its range's
// pairs in order (the pair (h, g) adds the conjugate of (g, h), so off-diagonal pairs
// count twice); the pair data are the same in every lane.  F = S_mu conj(S_alpha) is
// Hermitian in (mu, alpha), so its 9 terms take 15 FMAs per sample: 3 diagonal (real
// F) and 3 off-diagonal pairs Re(F e^{iwt}), Re(conj(F) e^{iwt}).  Phases: e^{i w t}
// of the lane's first sample by sincos, then the step factor e^{i w dt} (recomputed
// when a step differs from the uniform one).  Grid: (sample blocks, partitions,
// items); partition p writes at C + p pstride; no cross-lane traffic; the partitions
// are summed in order by k_unc_combine.
__global__ __launch_bounds__(kCorrThreads) void k_corr(const double2 *__restrict__ S, size_t plane,
                                                       const double *__restrict__ lam, const double *__restrict__ times,
                                                       int n, int ns, size_t pstride, double *__restrict__ C) {
    const int item = blockIdx.z, part = blockIdx.y, nparts = gridDim.y;
    const int s_first = (blockIdx.x * blockDim.x + threadIdx.x) * kCorrSamples;
    const size_t nn = (size_t)n * n;
    const double2 *Sx = S + (size_t)item * nn, *Sy = Sx + plane, *Sz = Sy + plane;
    const double *l = lam + (size_t)item * n;
    // this partition's range [k0, k1) of the n (n + 1) / 2 pairs, and its first (g, h)
    const long long npairs = (long long)n * (n + 1) / 2;
    const long long k0 = npairs * part / nparts, k1 = npairs * (part + 1) / nparts;
    int g = 0;
    long long row0 = 0;  // pair index of (g, g)
    while (g < n && row0 + (n - g) <= k0) { row0 += n - g; ++g; }
    int h = g + (int)(k0 - row0);
    double t_s[kCorrSamples];
#pragma unroll
    for (int j = 0; j < kCorrSamples; ++j) t_s[j] = times[min(s_first + j, ns - 1)];
    const double dt0 = ns > 1 ? times[1] - times[0] : 0.0;
    double acc[kCorrSamples][9];
#pragma unroll
    for (int j = 0; j < kCorrSamples; ++j)
#pragma unroll
        for (int q = 0; q < 9; ++q) acc[j][q] = 0.0;
    for (long long k = k0; k < k1; ++k) {
        const size_t e = (size_t)g * n + h;
        const double wt2 = h > g ? 2.0 : 1.0;
        const double2 sv[3] = {Sx[e], Sy[e], Sz[e]};
        double fd[3], fr[3], fi[3];  // diagonal |S_m|^2; off-diagonal S_m conj(S_q), (m, q) = (0,1), (0,2), (1,2)
#pragma unroll
        for (int m = 0; m < 3; ++m) fd[m] = wt2 * (sv[m].x * sv[m].x + sv[m].y * sv[m].y);
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            const int m = o < 2 ? 0 : 1, q = o < 1 ? 1 : 2;
            fr[o] = wt2 * (sv[m].x * sv[q].x + sv[m].y * sv[q].y);
            fi[o] = wt2 * (sv[m].y * sv[q].x - sv[m].x * sv[q].y);
        }
        const double w = l[g] - l[h];
        double ar, ai, zr, zi;
        sincos(w * t_s[0], &ai, &ar);
        sincos(w * dt0, &zi, &zr);
#pragma unroll
        for (int j = 0; j < kCorrSamples; ++j) {
#pragma unroll
            for (int m = 0; m < 3; ++m) acc[j][4 * m] = fma(fd[m], ar, acc[j][4 * m]);
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                const int m = o < 2 ? 0 : 1, q = o < 1 ? 1 : 2;
                acc[j][3 * m + q] = fma(fr[o], ar, fma(-fi[o], ai, acc[j][3 * m + q]));
                acc[j][3 * q + m] = fma(fr[o], ar, fma(fi[o], ai, acc[j][3 * q + m]));
            }
            if (j + 1 < kCorrSamples) {
                const double dtj = t_s[j + 1] - t_s[j];
                if (dtj == dt0) {
                    const double nr = ar * zr - ai * zi, ni = ar * zi + ai * zr;
                    ar = nr;
                    ai = ni;
                } else {
                    sincos(w * t_s[j + 1], &ai, &ar);
                }
            }
        }
        if (++h == n) { ++g; h = g; }
    }
    double *Cp = C + part * pstride;
#pragma unroll
    for (int j = 0; j < kCorrSamples; ++j)
        if (s_first + j < ns)
#pragma unroll
            for (int q = 0; q < 9; ++q) Cp[((size_t)item * ns + s_first + j) * 9 + q] = acc[j][q];
}

// Observables of the uncoupled pair from the radicals' correlations:
//   T_c(s) = sum_{mu nu alpha beta} Q_c[mu][alpha][nu][beta] CA_{mu alpha}(s) CB_{nu beta}(s)
// with the identity rows/columns of C filled in (C_00 = dimension, C_0a = C_m0 = 0);
// Q carries p^c_{mu nu} r_{alpha beta} / (16 Z_A Z_B) and the rate of observable c.
// CA, CB hold nparts partial correlations (k_corr's pair partitions, `pstride` apart),
// summed in partition order.  One workgroup of kCombineThreads per item: lane t takes
// the samples t, t + kCombineThreads, ... (Kahan-compensated yield sums in sample
// order), then the lanes' sums are added in a fixed tree order.
constexpr int kCombineThreads = 256;
__global__ __launch_bounds__(kCombineThreads) void k_unc_combine(
    const double *__restrict__ CA, const double *__restrict__ CB, int nparts, size_t pstride,
    const double *__restrict__ Q, const double *__restrict__ wy, const double *__restrict__ wt, int ns, int nobs,
    double dimA, double dimB, double *__restrict__ yields, double *__restrict__ trace) {
    const int item = blockIdx.x, t = threadIdx.x;
    __shared__ double red[kCombineThreads][kMaxObs];
    double ys[kMaxObs], yc[kMaxObs];
    for (int c = 0; c < nobs; ++c) ys[c] = yc[c] = 0.0;
    for (int s = t; s < ns; s += kCombineThreads) {
        double a[16], b[16];
        for (int i = 0; i < 16; ++i) a[i] = b[i] = 0.0;
        a[0] = dimA;
        b[0] = dimB;
        for (int p = 0; p < nparts; ++p) {
            const double *ca = CA + p * pstride + ((size_t)item * ns + s) * 9;
            const double *cb = CB + p * pstride + ((size_t)item * ns + s) * 9;
            for (int m = 0; m < 3; ++m)
                for (int q = 0; q < 3; ++q) {
                    a[(m + 1) * 4 + q + 1] += ca[3 * m + q];
                    b[(m + 1) * 4 + q + 1] += cb[3 * m + q];
                }
        }
        for (int c = 0; c < nobs; ++c) {
            const double *Qc = Q + (size_t)c * 256;
            double v = 0.0;
            for (int ma = 0; ma < 16; ++ma) {
                if (a[ma] == 0.0) continue;
                double inner = 0.0;
                for (int nb2 = 0; nb2 < 16; ++nb2) inner += Qc[ma * 16 + nb2] * b[nb2];
                v += a[ma] * inner;
            }
            const double term = wy[(size_t)s * nobs + c] * v - yc[c];
            const double nsum = ys[c] + term;
            yc[c] = (nsum - ys[c]) - term;
            ys[c] = nsum;
            if (trace) trace[((size_t)item * ns + s) * nobs + c] = wt[(size_t)s * nobs + c] * v;
        }
    }
    for (int c = 0; c < nobs; ++c) red[t][c] = ys[c];
    __syncthreads();
    for (int half = kCombineThreads / 2; half > 0; half >>= 1) {
        if (t < half)
            for (int c = 0; c < nobs; ++c) red[t][c] += red[t + half][c];
        __syncthreads();
    }
    if (t == 0)
        for (int c = 0; c < nobs; ++c) yields[(size_t)item * nobs + c] = red[0][c];
}

// Planes of the dense-path operand: U[b] row-major n x n (double2) -> two planes of
// Dp*Dp floats at dst + b*stride: re, then im.
// The planes are stored in the MFMA operand order of k_dense: for row tile rt and
// k-group g (16 columns), lane l = 16 q + i holds rows rt*16 + i and columns
// 16 g + 4 u + q for u = 0..3 as one float4, so a lane fetches four k-steps of
// its A operand with one 16-byte load (index ((rt * Dp/16 + g) * 64 + l) * 4 + u).
// Padding rows and columns are written as zero.
__global__ void k_to_planes(const double2 *__restrict__ U, float *__restrict__ dst, int n, int Dp, size_t stride) {
    const size_t b = blockIdx.y;
    const int ng = Dp >> 4;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < Dp * Dp; idx += gridDim.x * blockDim.x) {
        const int u = idx & 3, l = (idx >> 2) & 63, tg = idx >> 8;
        const int rt = tg / ng, g = tg - rt * ng;
        const int row = rt * 16 + (l & 15), col = 16 * g + 4 * u + (l >> 4);
        const double2 v = (row < n && col < n) ? U[b * n * n + (size_t)row * n + col] : make_double2(0.0, 0.0);
        dst[b * stride + idx] = (float)v.x;
        dst[b * stride + (size_t)Dp * Dp + idx] = (float)v.y;
    }
}

// ---------------------------------------------------------------------------
// Launchers (called from the host part of the library)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Uneven batch: radical pairs of different Hilbert dimension in one persistent
// launch (the loop over systems of RunSection.cpp:96-140 / main.cpp:634-686 and the
// per-system time loop of TaskHSGeneral.cpp:470-576, for every entry at once).
// Every workgroup has 1024 threads and the same LDS, and pops packs from one
// queue; the pack's class selects the body and its packing:
//   matrix-free, D = 4 ng:   ng  64  128  256  512 1024 2048 4096
//     groups per thread G:        1    1    1    1    2    2    4 (v_j mirrored in LDS)
//     threads per column T:      64  128  256  512  512 1024 1024
//     columns per workgroup P:   16    8    4    2    2    1    1
//   dense MFMA (D <= 512, NT row tiles of 16): 16 waves on one tile from NT = 12 up,
//     8 waves x 2 tiles at NT = 6, 8 and 4 waves x 4 tiles below.
// A pack's result does not depend on the other packs of the launch.  The packs
// are ordered by the host (largest first) so that the small ones fill the tail.
// ---------------------------------------------------------------------------
// The uneven kernel calls each body through a function of its own, so that every
// body gets the register allocation it has in its dedicated kernel (inlined into one
// switch, the bodies shared one allocation and spilled).  LDS pointers cross the
// call in the LDS address space, so the body's accesses stay LDS instructions.
#define MSG_LDS __attribute__((address_space(3)))
// Function arguments arrive in vector registers, and loads through them are vector
// loads; the entry's arguments are the same for the whole workgroup, so the body reads
// them through a wave-uniform pointer in the constant address space: scalar loads where
// a field is used, as in a dedicated kernel.  (A copy of all their words made uniform
// and held in scalar registers spilled: the launch took 1418 ms against 1377 ms.)
__device__ __forceinline__ const StepArgs &uniform_args(const StepArgs *a) {
    const uint64_t v = reinterpret_cast<uint64_t>(a);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const ConstPtr<StepArgs> c = (ConstPtr<StepArgs>)(((uint64_t)hi << 32) | lo);
    return *(const StepArgs *)c;
}
template <int G, int T, int P>
__device__ __noinline__ void uneven_lz(const StepArgs *__restrict__ a, int u0, int nact, float *bas, MSG_LDS float4 *lds,
                                       MSG_LDS double *scr, MSG_LDS double *ys, MSG_LDS double *yc,
                                       MSG_LDS float2 *coef, MSG_LDS float4 *blk) {
    const BodyShared sh{(double *)ys, (double *)yc, (float2 *)coef, (float4 *)blk, nullptr};
    const StepArgs &ua = uniform_args(a);
    lanczos_columns<G, T, P, true, false>(ua, __builtin_amdgcn_readfirstlane(u0),
                                          __builtin_amdgcn_readfirstlane(nact), bas, (float4 *)lds, (double *)scr, sh);
}
__device__ __noinline__ void uneven_big(const StepArgs *__restrict__ a, int u, float *bas, MSG_LDS float4 *lds, MSG_LDS double *scr,
                                        MSG_LDS double *ys, MSG_LDS double *yc, MSG_LDS float2 *coef,
                                        MSG_LDS float4 *blk) {
    const BodyShared sh{(double *)ys, (double *)yc, (float2 *)coef, (float4 *)blk, nullptr};
    const StepArgs &ua = uniform_args(a);
    lanczos_big_column<true, false>(ua, __builtin_amdgcn_readfirstlane(u), bas, (float4 *)lds,
                                    (double *)scr, sh);
}
template <int NT, int NW, int P>
__device__ __noinline__ void uneven_dense(const StepArgs *__restrict__ a, int u0, int nact, MSG_LDS float *lds, MSG_LDS double *scr,
                                          MSG_LDS double *ys, MSG_LDS double *yc, MSG_LDS double *colsum) {
    const BodyShared sh{(double *)ys, (double *)yc, nullptr, nullptr, (double *)colsum};
    const StepArgs &ua = uniform_args(a);
    dense_tiles<NT, NW, P>(ua, __builtin_amdgcn_readfirstlane(u0), __builtin_amdgcn_readfirstlane(nact),
                           (float *)lds, (double *)scr, sh);
}

__global__ __launch_bounds__(kUnevenThreads) void k_uneven(const UnevenArgs *__restrict__ pu) {
    extern __shared__ float4 usm[];  // kUnevenPlanesBytes of planes, then the reduction scratch
    __shared__ double ys[16 * kMaxObs], yc[16 * kMaxObs];
    __shared__ float2 s_coef[16 * kMaxKrylov];
    __shared__ float4 s_blk[kBlkTable];
    __shared__ double s_col[2 * 256];
    __shared__ int s_pack;
    double *scr = reinterpret_cast<double *>(reinterpret_cast<char *>(usm) + kUnevenPlanesBytes);
    float *bas = pu->basis + (size_t)blockIdx.x * pu->basis_stride;
    for (;;) {
        if (threadIdx.x == 0) s_pack = atomicAdd(pu->queue, 1);
        __syncthreads();
        const int k = __builtin_amdgcn_readfirstlane(s_pack);
        __syncthreads();
        if (k >= pu->n_packs) break;
        const UnevenPack pk = pu->packs[k];
        const StepArgs *a = pu->ents + pk.entry;
        if (pk.nact < 1 || pk.nact > uneven_pack_width(pk.cls)) {  // a pack the class's body cannot hold
            if (threadIdx.x == 0) atomicOr(a->status, 8);
            continue;
        }
        MSG_LDS float4 *L = (MSG_LDS float4 *)usm;
        MSG_LDS double *R = (MSG_LDS double *)scr;
        MSG_LDS double *Y = (MSG_LDS double *)ys, *Yc = (MSG_LDS double *)yc;
        MSG_LDS float2 *Cf = (MSG_LDS float2 *)s_coef;
        MSG_LDS float4 *Bk = (MSG_LDS float4 *)s_blk;
        MSG_LDS float *Lf = (MSG_LDS float *)usm;
        MSG_LDS double *Cs = (MSG_LDS double *)s_col;
        switch (pk.cls) {
            case UC_LZ64: uneven_lz<1, 64, 16>(a, pk.u0, pk.nact, bas, L, R, Y, Yc, Cf, Bk); break;
            case UC_LZ128: uneven_lz<1, 128, 8>(a, pk.u0, pk.nact, bas, L, R, Y, Yc, Cf, Bk); break;
            case UC_LZ256: uneven_lz<1, 256, 4>(a, pk.u0, pk.nact, bas, L, R, Y, Yc, Cf, Bk); break;
            case UC_LZ512: uneven_lz<1, 512, 2>(a, pk.u0, pk.nact, bas, L, R, Y, Yc, Cf, Bk); break;
            case UC_LZ1024: uneven_lz<2, 512, 2>(a, pk.u0, pk.nact, bas, L, R, Y, Yc, Cf, Bk); break;
            case UC_LZ2048: uneven_lz<2, 1024, 1>(a, pk.u0, pk.nact, bas, L, R, Y, Yc, Cf, Bk); break;
            case UC_LZ4096: uneven_big(a, pk.u0, bas, L, R, Y, Yc, Cf, Bk); break;
            case UC_DENSE + 0: uneven_dense<1, 4, 4>(a, pk.u0, pk.nact, Lf, R, Y, Yc, Cs); break;
            case UC_DENSE + 1: uneven_dense<2, 4, 4>(a, pk.u0, pk.nact, Lf, R, Y, Yc, Cs); break;
            case UC_DENSE + 2: uneven_dense<3, 4, 4>(a, pk.u0, pk.nact, Lf, R, Y, Yc, Cs); break;
            case UC_DENSE + 3: uneven_dense<4, 4, 4>(a, pk.u0, pk.nact, Lf, R, Y, Yc, Cs); break;
            case UC_DENSE + 4: uneven_dense<6, 8, 2>(a, pk.u0, pk.nact, Lf, R, Y, Yc, Cs); break;
            case UC_DENSE + 5: uneven_dense<8, 8, 2>(a, pk.u0, pk.nact, Lf, R, Y, Yc, Cs); break;
            case UC_DENSE + 6: uneven_dense<12, 16, 1>(a, pk.u0, pk.nact, Lf, R, Y, Yc, Cs); break;
            case UC_DENSE + 7: uneven_dense<16, 16, 1>(a, pk.u0, pk.nact, Lf, R, Y, Yc, Cs); break;
            case UC_DENSE + 8: uneven_dense<24, 16, 1>(a, pk.u0, pk.nact, Lf, R, Y, Yc, Cs); break;
            case UC_DENSE + 9: uneven_dense<32, 16, 1>(a, pk.u0, pk.nact, Lf, R, Y, Yc, Cs); break;
            default:
                if (threadIdx.x == 0) atomicOr(a->status, 8);
                break;
        }
    }
}

// LDS planes each class uses (bytes), for the host's checks.
static_assert(uneven_pack_width(UC_LZ64) == 16 && uneven_pack_width(UC_LZ128) == 8 &&
                  uneven_pack_width(UC_LZ256) == 4 && uneven_pack_width(UC_LZ512) == 2 &&
                  uneven_pack_width(UC_LZ1024) == 2 && uneven_pack_width(UC_LZ2048) == 1 &&
                  uneven_pack_width(UC_LZ4096) == 1 && uneven_pack_width(UC_DENSE + 3) == 4 &&
                  uneven_pack_width(UC_DENSE + 5) == 2 && uneven_pack_width(UC_DENSE + 6) == 1,
              "pack widths match the bodies of k_uneven");
static_assert(16 * 4 * 64 * 16 <= kUnevenPlanesBytes && 2 * 4 * 1024 * 16 <= kUnevenPlanesBytes &&
                  4 * 2048 * 16 <= kUnevenPlanesBytes && 2 * 4096 * 16 <= kUnevenPlanesBytes &&
                  3 * 512 * 16 * 4 <= kUnevenPlanesBytes && 2 * 3 * 128 * 16 * 4 <= kUnevenPlanesBytes &&
                  4 * 3 * 64 * 16 * 4 <= kUnevenPlanesBytes,
              "every class fits the uneven launch's LDS planes");

size_t uneven_lds_bytes() { return (size_t)kUnevenPlanesBytes + kUnevenScratchDoubles * sizeof(double); }

hipError_t launch_uneven(const UnevenArgs *du, int workgroups, hipStream_t st) {
    hipLaunchKernelGGL(k_uneven, dim3(workgroups), dim3(kUnevenThreads), uneven_lds_bytes(), st, du);
    return hipGetLastError();
}

int uneven_occupancy() {
    int nb = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_uneven, kUnevenThreads, uneven_lds_bytes()) == hipSuccess
               ? nb
               : 1;
}

#define MSG_DENSE_VARIANTS(X) X(1) X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(24) X(32)
#define MSG_LANCZOS_VARIANTS(X) \
    X(1, 64, true) X(1, 128, true) X(1, 256, true) X(1, 512, true) X(1, 1024, true) X(2, 512, true) X(2, 1024, true)
// shapes of the Lanczos kernel only (more groups per thread for tuning; no Arnoldi form)
#define MSG_LANCZOS_EXTRA(X) X(2, 128, true) X(2, 256, true) X(4, 256, true)
// and of the Arnoldi kernel (two groups per thread at D = 1024 and 2048, four at
// D = 16384: the Krylov vector takes 128 KiB of LDS)
#define MSG_ARNOLDI_EXTRA(X) X(2, 128, true) X(2, 256, true) X(4, 1024, true)
// shapes of the kernel with one nucleus of spin above 1/2 (M groups per thread, T threads;
// both Krylov vectors in LDS up to 96 KiB)
#define MSG_MIX_VARIANTS(X) \
    X(3, 64, true) X(3, 128, true) X(3, 256, true) X(3, 512, true) X(3, 1024, false) \
    X(4, 64, true) X(4, 128, true) X(4, 256, true) X(4, 512, false) X(4, 1024, false)
// shapes of the RK4 kernel (D = 1024 ... 16384)
#define MSG_RK4_VARIANTS(X) X(1, 256) X(1, 512) X(2, 512) X(2, 1024) X(4, 1024)

// Split-column kernels for D = 4 2^NN, NN = 12..16, with 1024 or 512 threads over a
// 2048-group slab.  At D = 65536 the 512-thread form (four groups per thread, 256
// VGPRs, 27 spilled against 71 at 1024 threads) takes 115 ms per trajectory launch
// against 148 ms.  (512-thread workgroups over 1024-group slabs, two per CU, measured
// slower at D = 16384 and 65536: more partner slabs and wider barriers.)
#define MSG_SPLIT_VARIANTS(X) X(12, 1024) X(13, 1024) X(14, 1024) X(15, 1024) X(16, 1024) \
    X(12, 512) X(13, 512) X(14, 512) X(15, 512) X(16, 512)

hipError_t launch_lanczos(const StepArgs &a, const StepArgs *da, int group_per_thread, int threads, int slots,
                          hipStream_t st) {
    if (group_per_thread == -3) {  // split columns: `slots` column slots of S workgroups each
        const int nn = 31 - __builtin_clz((unsigned)(a.dim >> 2));  // D / 4 = 2^nn
        const int S = (a.dim >> 2) / kSplitSlab;
        const size_t lds = (size_t)4 * kSplitSlab * sizeof(float4);
        void *args[] = {(void *)&da};
#define X(NN_, T_)                                                                                                \
        if (nn == NN_ && threads == T_)                                                                           \
            return hipLaunchCooperativeKernel((a.op_flags & OP_AXIAL) ? (const void *)k_lanczos_split<NN_, T_, true>   \
                                                                      : (const void *)k_lanczos_split<NN_, T_, false>, \
                                              dim3(slots * S), dim3(T_), args, (unsigned)lds, st);
        MSG_SPLIT_VARIANTS(X)
#undef X
        return hipErrorInvalidValue;
    }
    if (group_per_thread <= 0) {  // streaming variants for large D (v_j mirrored in LDS when G = -1)
        if (threads != 1024) return hipErrorInvalidValue;
        if (group_per_thread == -1)
            hipLaunchKernelGGL(k_lanczos_big<true>, dim3(slots), dim3(1024), (size_t)8 * a.dim, st, da);
        else if (group_per_thread == -2)
            hipLaunchKernelGGL((k_lanczos_big<false, true>), dim3(slots), dim3(1024), (size_t)32 * kTileG, st, da);
        else
            hipLaunchKernelGGL(k_lanczos_big<false>, dim3(slots), dim3(1024), 0, st, da);
        return hipGetLastError();
    }
    if (a.mixm) {
#define X(M_, T_, B_)                                                                                          \
        if (a.mixm == M_ && threads == T_) {                                                                   \
            const size_t lds_m = (size_t)(B_ ? 4 : 2) * a.dim * sizeof(float) + 2 * 256 * sizeof(double);     \
            hipLaunchKernelGGL((k_lanczos_mix<M_, T_, B_>), dim3(slots), dim3(threads), lds_m, st, da);        \
            return hipGetLastError();                                                                           \
        }
        MSG_MIX_VARIANTS(X)
#undef X
        return hipErrorInvalidValue;
    }
    if (a.rk4) {
        const size_t lds_r = (size_t)2 * a.dim * sizeof(float) + 2 * 256 * sizeof(double);
#define X(G_, T_)                                                                                 \
        if (group_per_thread == G_ && threads == T_) {                                            \
            hipLaunchKernelGGL((k_rk4<G_, T_>), dim3(slots), dim3(threads), lds_r, st, da);       \
            return hipGetLastError();                                                              \
        }
        MSG_RK4_VARIANTS(X)
#undef X
        return hipErrorInvalidValue;
    }
    if (a.arnoldi) {
        const size_t lds_a = (size_t)2 * a.dim * sizeof(float) + 2 * 256 * sizeof(double);
#define X(G_, T_, B_)                                                                                  \
        if (group_per_thread == G_ && threads == T_) {                                                 \
            hipLaunchKernelGGL((k_arnoldi<G_, T_>), dim3(slots), dim3(threads), lds_a, st, da);        \
            return hipGetLastError();                                                                   \
        }
        MSG_LANCZOS_VARIANTS(X)
        MSG_ARNOLDI_EXTRA(X)
#undef X
        return hipErrorInvalidValue;
    }
    const size_t lds = (size_t)4 * a.dim * sizeof(float) + 2 * 256 * sizeof(double);
#define X(G_, T_, B_)                                                                                   \
    if (group_per_thread == G_ && threads == T_) {                                                      \
        if (a.dyn_head)                                                                                  \
            hipLaunchKernelGGL((k_lanczos<G_, T_, B_, true>), dim3(slots), dim3(threads), lds, st, da);  \
        else                                                                                             \
            hipLaunchKernelGGL((k_lanczos<G_, T_, B_, false>), dim3(slots), dim3(threads), lds, st, da); \
        return hipGetLastError();                                                                        \
    }
    MSG_LANCZOS_VARIANTS(X)
    MSG_LANCZOS_EXTRA(X)
#undef X
    return hipErrorInvalidValue;
}

// Raises the dynamic-LDS limit of the large-LDS kernels to what the device
// allows per workgroup (minus their static LDS); best effort.
hipError_t set_lanczos_lds_limits() {
    int dev = 0, maxb = 65536;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&maxb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    const int lim = maxb - 2048;
#define X(G_, T_, B_)                                                                                               \
    (void)hipFuncSetAttribute((const void *)k_lanczos<G_, T_, B_, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
    (void)hipFuncSetAttribute((const void *)k_lanczos<G_, T_, B_, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
    (void)hipFuncSetAttribute((const void *)k_arnoldi<G_, T_>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    MSG_LANCZOS_VARIANTS(X)
#undef X
#define X(G_, T_, B_)                                                                                               \
    (void)hipFuncSetAttribute((const void *)k_lanczos<G_, T_, B_, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
    (void)hipFuncSetAttribute((const void *)k_lanczos<G_, T_, B_, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    MSG_LANCZOS_EXTRA(X)
#undef X
#define X(G_, T_, B_) \
    (void)hipFuncSetAttribute((const void *)k_arnoldi<G_, T_>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    MSG_ARNOLDI_EXTRA(X)
#undef X
#define X(NT_) (void)hipFuncSetAttribute((const void *)k_dense<NT_>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    MSG_DENSE_VARIANTS(X)
#undef X
#define X(NN_, T_) \
    (void)hipFuncSetAttribute((const void *)k_lanczos_split<NN_, T_, false>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                              maxb - 8192);                                                                         \
    (void)hipFuncSetAttribute((const void *)k_lanczos_split<NN_, T_, true>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                              maxb - 8192);
    MSG_SPLIT_VARIANTS(X)
#undef X
    (void)hipFuncSetAttribute((const void *)k_lanczos_big<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              maxb - 8192);
    (void)hipFuncSetAttribute((const void *)k_lanczos_big<false, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              maxb - 8192);
    (void)hipFuncSetAttribute((const void *)k_uneven, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)uneven_lds_bytes());
#define X(G_, T_) (void)hipFuncSetAttribute((const void *)k_rk4<G_, T_>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    MSG_RK4_VARIANTS(X)
#undef X
#define X(M_, T_, B_) \
    (void)hipFuncSetAttribute((const void *)k_lanczos_mix<M_, T_, B_>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    MSG_MIX_VARIANTS(X)
#undef X
    (void)hipGetLastError();
    return hipSuccess;
}

hipError_t launch_dense(const StepArgs &a, const StepArgs *da, int slots, hipStream_t st) {
    const size_t lds = (size_t)kDenseLdsPlanes * a.dpad * 16 * sizeof(float) + 2 * 256 * sizeof(double);
#define X(NT_)                                                                  \
    if (a.dpad == 16 * NT_) {                                                   \
        hipLaunchKernelGGL(k_dense<NT_>, dim3(slots), dim3(64 * dense_waves(NT_)), lds, st, da); \
        return hipGetLastError();                                               \
    }
    MSG_DENSE_VARIANTS(X)
#undef X
    return hipErrorInvalidValue;
}

int mix_occupancy(int mult, int threads, size_t lds) {
    int nb = 0;
#define X(M_, T_, B_)                                                                                           \
    if (mult == M_ && threads == T_)                                                                            \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_lanczos_mix<M_, T_, B_>, threads, lds) == hipSuccess ? nb : 1;
    MSG_MIX_VARIANTS(X)
#undef X
    return 1;
}
bool mix_dbuf(int mult, int threads) {
#define X(M_, T_, B_) \
    if (mult == M_ && threads == T_) return B_;
    MSG_MIX_VARIANTS(X)
#undef X
    return false;
}

int rk4_occupancy(int group_per_thread, int threads, size_t lds) {
    int nb = 0;
#define X(G_, T_)                                                                                           \
    if (group_per_thread == G_ && threads == T_)                                                            \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_rk4<G_, T_>, threads, lds) == hipSuccess ? nb : 1;
    MSG_RK4_VARIANTS(X)
#undef X
    return 1;
}

int lanczos_occupancy(int group_per_thread, int threads, size_t lds, bool arnoldi) {
    int nb = 0;
    if (arnoldi) {
#define X(G_, T_, B_)                                                                                         \
        if (group_per_thread == G_ && threads == T_)                                                          \
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_arnoldi<G_, T_>, threads, lds) == hipSuccess ? nb : 1;
        MSG_LANCZOS_VARIANTS(X)
        MSG_ARNOLDI_EXTRA(X)
#undef X
        return 1;
    }
    if (group_per_thread == -3)  // the split kernels of one width share one resource footprint
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
                   &nb, threads == 512 ? k_lanczos_split<14, 512, false> : k_lanczos_split<14, 1024, false>, threads,
                   lds) == hipSuccess
                   ? nb : 1;
    if (group_per_thread == 0)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_lanczos_big<false>, 1024, 0) == hipSuccess ? nb : 1;
    if (group_per_thread == -1)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_lanczos_big<true>, 1024, lds) == hipSuccess ? nb : 1;
    if (group_per_thread == -2)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_lanczos_big<false, true>, 1024, lds) == hipSuccess
                   ? nb : 1;
#define X(G_, T_, B_)                                                                                         \
    if (group_per_thread == G_ && threads == T_)                                                              \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_lanczos<G_, T_, B_, false>, threads, lds) == hipSuccess \
                   ? nb : 1;
    MSG_LANCZOS_VARIANTS(X)
    MSG_LANCZOS_EXTRA(X)
#undef X
    return 1;
}

int dense_threads(int dpad) { return 64 * dense_waves(dpad / 16); }

int dense_occupancy(int dpad, size_t lds) {
    int nb = 0;
#define X(NT_)                                                                                      \
    if (dpad == 16 * NT_)                                                                           \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_dense<NT_>, 64 * dense_waves(NT_), lds) == hipSuccess \
                   ? nb : 1;
    MSG_DENSE_VARIANTS(X)
#undef X
    return 1;
}


hipError_t launch_zlincomb(const void *X, const void *X2, const void *X3, void *out, int n, int batch,
                           const double *c, hipStream_t st) {
    const size_t total = (size_t)batch * n * n;
    const int blocks = (int)std::min<size_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_zlincomb, dim3(blocks), dim3(256), 0, st, (const double2 *)X, (const double2 *)X2,
                       (const double2 *)X3, (double2 *)out, n, total, c[0], c[1], c[2], c[3]);
    return hipGetLastError();
}

hipError_t launch_blockdiag_right(const void *M, const double *B4, double scale, void *Y, int n, hipStream_t st) {
    Block4 b;
    for (int i = 0; i < 16; ++i) b.m[i] = make_double2(B4[2 * i], B4[2 * i + 1]);
    const size_t total = (size_t)n * n;
    const int blocks = (int)std::min<size_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(k_blockdiag_right, dim3(blocks), dim3(256), 0, st, (const double2 *)M, b, scale, (double2 *)Y, n);
    return hipGetLastError();
}

hipError_t launch_pauli_right(const void *M, void *Y, int n, int mu, int batch, hipStream_t st) {
    const size_t total = (size_t)batch * n * n;
    const int blocks = (int)std::min<size_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_pauli_right, dim3(blocks), dim3(256), 0, st, (const double2 *)M, (double2 *)Y, n, mu, total);
    return hipGetLastError();
}

hipError_t launch_corr(const void *S, size_t plane, const double *lam, const double *times, int n, int ns, int batch,
                       int nparts, size_t pstride, double *C, hipStream_t st) {
    const int per_block = kCorrThreads * kCorrSamples;
    hipLaunchKernelGGL(k_corr, dim3((ns + per_block - 1) / per_block, nparts, batch), dim3(kCorrThreads), 0, st,
                       (const double2 *)S, plane, lam, times, n, ns, pstride, C);
    return hipGetLastError();
}

hipError_t launch_unc_combine(const double *CA, const double *CB, int nparts, size_t pstride, const double *Q,
                              const double *wy, const double *wt, int n_items, int ns, int nobs, double dimA,
                              double dimB, double *yields, double *trace, hipStream_t st) {
    hipLaunchKernelGGL(k_unc_combine, dim3(n_items), dim3(kCombineThreads), 0, st, CA, CB, nparts, pstride, Q, wy, wt,
                       ns, nobs, dimA, dimB, yields, trace);
    return hipGetLastError();
}

hipError_t launch_lorentz_sum(const void *Pt, const void *Rt, const double *lam, double k, int n, double *rows,
                              double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_lorentz_rows, dim3(n), dim3(256), 0, st, (const double2 *)Pt, (const double2 *)Rt, lam, k, 0.0,
                       k, n, rows);
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, st, (const double *)rows, n, out);
    return hipGetLastError();
}

hipError_t launch_timeinf_sum(const void *Pt, const void *Rt, const double *lam, double a, int n, double *rows,
                              double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_lorentz_rows, dim3(n), dim3(256), 0, st, (const double2 *)Pt, (const double2 *)Rt, lam, a, 1.0,
                       1.0, n, rows);
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, st, (const double *)rows, n, out);
    return hipGetLastError();
}

hipError_t launch_dyn_generator(const void *G, const void *Z, const double *coef, const int *item, const double *dt,
                               int nz, int n, int batch, void *out, hipStream_t st) {
    const size_t total = (size_t)batch * n * n;
    const int blocks = (int)std::min<size_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_dyn_generator, dim3(blocks), dim3(256), 0, st, (const double2 *)G, (const double2 *)Z, coef,
                       item, dt, nz, n, total, (double2 *)out);
    return hipGetLastError();
}

hipError_t launch_herm_update(const void *Y, const void *R, void *S, int n, int batch, double c, double s,
                              hipStream_t st) {
    const int tiles = (n + 31) / 32;
    hipLaunchKernelGGL(k_herm_update, dim3(tiles, tiles, batch), dim3(256), 0, st, (const double2 *)Y,
                       (const double2 *)R, (double2 *)S, n, c, s);
    return hipGetLastError();
}

hipError_t launch_accumulate_fro(const void *X, const void *Delta, void *W, int n, int batch, int parts, double *part,
                                 hipStream_t st) {
    hipLaunchKernelGGL(k_accumulate_fro, dim3(parts, batch), dim3(256), 0, st, (const double2 *)X,
                       (const double2 *)Delta, (double2 *)W, (size_t)n * n, part);
    return hipGetLastError();
}


hipError_t launch_pair_block_trace(const void *X, int n, int batch, void *out, hipStream_t st) {
    hipLaunchKernelGGL(k_pair_block_trace, dim3(batch, 16), dim3(256), 0, st, (const double2 *)X, n, (double2 *)out);
    return hipGetLastError();
}

hipError_t launch_to_planes(const void *U, float *dst, int n, int Dp, size_t stride, int batch, hipStream_t st) {
    const int blocks = std::min(1024, (Dp * Dp + 255) / 256);
    hipLaunchKernelGGL(k_to_planes, dim3(blocks, batch), dim3(256), 0, st, (const double2 *)U, dst, n, Dp, stride);
    return hipGetLastError();
}

hipError_t launch_init_norms(const StepArgs &a, const StepArgs *da, hipStream_t st) {
    hipLaunchKernelGGL(k_init_norms, dim3(a.n_units), dim3(64), 0, st, da);
    return hipGetLastError();
}

hipError_t launch_zero_int(int *p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_int, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_reduce_columns(const double *in, double *out, int n_items, int ncols, int nk, int in_stride, int pick,
                                 hipStream_t st) {
    const int n = n_items * nk;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_columns, dim3((n + 127) / 128), dim3(128), 0, st, in, out, n_items, ncols, nk, in_stride,
                       pick);
    return hipGetLastError();
}

}  // namespace msg
