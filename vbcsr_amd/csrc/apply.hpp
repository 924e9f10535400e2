// Internal interface and device code of the apply kernels (shared by
// apply.hip and kpm.hip).
//
// Two row kernels, both one wavefront per block row at a time:
//
// * row_loop (SpMM, and SpMV of blocks wider than 8): the 64 lanes are split
//   into Kp = min(64, pow2ceil(K)) vector columns x NG = 64/Kp lane groups; a
//   group holds rpp row slots x P block partitions (P > 1 when the block row
//   has fewer scalar rows than groups, merged at the end with xor shuffles),
//   and each lane keeps RB row accumulators (register blocking) so that one X
//   element feeds up to RB rows.  The row's block metadata comes from ONE
//   lane-parallel load of 64 slots (meta: x offset, width, panel column), so
//   the inner loop issues only the value and X loads -- no dependent index
//   chains.  Accumulation is in FP64 registers for every storage precision.
//
// * row_loop_spmv (K == 1, blocks up to 8 x 8): lanes are (row a, partition
//   p); the values of each run of 64/RPP consecutive blocks (contiguous in
//   memory) are first streamed into a wave-private LDS tile with 16-byte
//   coalesced loads, each lane loads one x entry of the run into the tile's x
//   strip, then every lane forms its block-row dot product from LDS.  The next
//   run's loads are issued before the current run is computed.
//
// The lane family and the matrix-core class (lanes.hpp) serve 2 <= K <= 64.
#pragma once

#include <initializer_list>
#include <type_traits>
#include <utility>

#include "common.hpp"

namespace vbh {

// T: vector (and accumulation input) type; V: matrix value type, T or its
// FP32 partner (precision 2: FP32 values against FP64 vectors).
template <typename T, typename V = T>
struct ApplyArgs {
    DevPattern p;
    const V* vals;
    int K;
    const T* X;
    int64_t ldx;
    T* Y;
    int64_t ldy;
    // fused epilogue: Y = alpha*AX + beta*X_own + gamma*Z; dots vs W (or Y) and X_own
    double alpha, beta, gamma;
    const T* Z;
    int64_t ldz;
    const T* W;
    int64_t ldw;
    double* partials;  // [G][2*max(dot_cols,1)] per-work-group FP64 partial dots, or null
    int dot_cols;      // K: one partial per column; 1 (or 0, no dots): summed over the columns
    int row_list;      // lane family: 1 = rows come from p.small_rows (the matrix-core class runs apart)
    int parts;         // lane family: partitions of each row's slots (1, 2, 4, 8)
    const int* split;  // G+1 row chunk boundaries
    int G;
};

template <typename T>
struct AdjointArgs {
    DevPattern p;
    const T* vals;
    int K;
    const T* X;
    int64_t ldx;
    T* Y;
    int64_t ldy;
    const int* t_ptr;
    const int64_t* t_voff;
    const int2* t_rows;
    const int* split;
    int G;
};

// Register accumulator of a storage type: FP64 for every storage precision,
// so an FP32 recursion rounds once per step (at the store), not per term.
template <typename T>
struct Wide {
    using type = double;
};
template <>
struct Wide<c64> {
    using type = c128;
};
template <>
struct Wide<c128> {
    using type = c128;
};
__device__ inline double widen(double x) { return x; }
__device__ inline double widen(float x) { return double(x); }
__device__ inline c128 widen(c128 x) { return x; }
__device__ inline c128 widen(c64 x) { return c128{double(x.re), double(x.im)}; }
template <typename T>
__device__ inline T narrow_to(typename Wide<T>::type x) {
    if constexpr (std::is_same_v<T, float>) return float(x);
    else if constexpr (std::is_same_v<T, c64>) return c64{float(x.re), float(x.im)};
    else return x;
}

__device__ inline int lane_read(int v, int lane_idx) { return __builtin_amdgcn_readlane(v, lane_idx); }
// A 64-bit value the wave holds in every lane, as a scalar.
__device__ inline int64_t uniform64(int64_t v) {
    const uint64_t u = uint64_t(v);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(u)), hi = __builtin_amdgcn_readfirstlane(uint32_t(u >> 32));
    return int64_t((uint64_t(hi) << 32) | lo);
}

// Series accumulation for Chebyshev sums sum_n c_n t_n built term by term:
// acc[row * ld + column] += (re + i im) * y for every stored (rounded) y of a
// fused step.  A kernel parameter of its own, not an ApplyArgs field: a larger
// argument struct raises the register use of every row kernel.
template <typename T>
struct SeriesAcc {
    T* acc;
    int64_t ld;
    double re, im;
};

// Fused epilogue for one output element (row yr, column v); ACC: also the
// series accumulation sa.
template <typename T, bool FUSED, bool ACC = false>
__device__ inline void emit(const ApplyArgs<T>& g, int64_t yr, int64_t xr, int v, typename Wide<T>::type acc,
                            double& d1, double& d2, const SeriesAcc<T>* sa = nullptr) {
    using A = typename Wide<T>::type;
    A y = acc;
    if constexpr (FUSED) {
        const A xo = widen(g.X[xr * g.ldx + v]);
        y = rscale(g.alpha, acc);
        if (g.beta != 0.0) y = y + rscale(g.beta, xo);
        if (g.gamma != 0.0) y = y + rscale(g.gamma, widen(g.Z[yr * g.ldz + v]));
        // dots of the stored (rounded) value: the state the recursion carries on
        const A ys = widen(narrow_to<T>(y));
        d1 += redot(g.W ? widen(g.W[yr * g.ldw + v]) : ys, ys);
        d2 += redot(xo, ys);
    }
    g.Y[yr * g.ldy + v] = narrow_to<T>(y);
    if constexpr (ACC) {
        T* ap = sa->acc + yr * sa->ld + v;
        const A c = Sc<A>::make(sa->re, sa->im);
        *ap = narrow_to<T>(Sc<A>::fma(c, widen(narrow_to<T>(y)), widen(*ap)));
    }
}

// SpMM row loop (see the file comment).  BS > 0: uniform BS x BS blocks.
template <typename T, int BS, int RB, bool FUSED>
__device__ inline void row_loop(const ApplyArgs<T>& g, int r0, int r1, int lane, int wave, double& d1, double& d2) {
    using A = typename Wide<T>::type;
    const DevPattern& p = g.p;
    const int K = g.K;
    const int Kp = K >= 64 ? 64 : pow2ceil(K);
    const int lgK = __builtin_ctz(Kp);
    const int NG = 64 >> lgK;
    const int v_l = lane & (Kp - 1);
    const int g_l = lane >> lgK;
    for (int i = r0 + wave; i < r1; i += kWaves) {
        const int r = BS > 0 ? BS : p.rbs[i];
        const int kb = p.row_ptr[i], ke = p.row_ptr[i + 1];
        const int rp = pow2ceil(r);
        const int rpp = rp < NG ? rp : NG;
        const int P = NG / rpp;
        const int a_l = g_l & (rpp - 1);
        const int p_l = g_l >> __builtin_ctz(rpp);
        const int64_t y0 = BS > 0 ? int64_t(i) * BS : p.roff[i];
        const int64_t x0 = BS > 0 ? int64_t(i) * BS : p.coff[i];
        const int64_t vrow = BS > 0 ? int64_t(kb) * (BS * BS) : p.voff[kb];
        for (int a0 = 0; a0 < r; a0 += rpp * RB) {
            for (int v0 = 0; v0 < K; v0 += Kp) {
                const int v = v0 + v_l;
                const bool vok = v < K;
                A acc[RB];
#pragma unroll
                for (int q = 0; q < RB; ++q) acc[q] = Sc<A>::zero();
                for (int t0 = kb; t0 < ke; t0 += 64) {
                    const int nb = ke - t0 < 64 ? ke - t0 : 64;
                    int mx = 0, mw = 0;
                    if (lane < nb) {
                        if constexpr (BS > 0) {
                            mx = p.col_ind[t0 + lane] * BS;
                            mw = ((t0 + lane - kb) * BS) << 8 | BS;
                        } else {
                            const int2 mm = p.meta[t0 + lane];
                            mx = mm.x;
                            mw = mm.y;
                        }
                    }
#pragma unroll 2
                    for (int ub = 0; ub < nb; ub += P) {
                        const int u = ub + p_l;
                        int xo, w;
                        if (P == 1) {
                            xo = lane_read(mx, ub);
                            w = lane_read(mw, ub);
                        } else {
                            xo = __shfl(mx, u, 64);
                            w = __shfl(mw, u, 64);
                        }
                        if (!vok || u >= nb) continue;
                        const int c = BS > 0 ? BS : (w & 255);
                        const T* ab = g.vals + vrow + int64_t(r) * (w >> 8);
                        const T* xb = g.X + int64_t(xo) * g.ldx + v;
                        if constexpr (BS > 0) {
                            T xv[BS];
#pragma unroll
                            for (int b = 0; b < BS; ++b) xv[b] = xb[int64_t(b) * g.ldx];
#pragma unroll
                            for (int q = 0; q < RB; ++q) {
                                const int a = a0 + a_l + q * rpp;
                                if (a < r) {
                                    T av[BS];
                                    const T* ar = ab + a * BS;
                                    if constexpr ((BS * sizeof(T)) % 16 == 0) {
                                        const T* al = static_cast<const T*>(__builtin_assume_aligned(ar, 16));
#pragma unroll
                                        for (int b = 0; b < BS; ++b) av[b] = al[b];
                                    } else {
#pragma unroll
                                        for (int b = 0; b < BS; ++b) av[b] = ar[b];
                                    }
#pragma unroll
                                    for (int b = 0; b < BS; ++b) acc[q] = Sc<A>::fma(widen(av[b]), widen(xv[b]), acc[q]);
                                }
                            }
                        } else {
                            for (int b = 0; b < c; ++b) {
                                const A xv = widen(xb[int64_t(b) * g.ldx]);
#pragma unroll
                                for (int q = 0; q < RB; ++q) {
                                    const int a = a0 + a_l + q * rpp;
                                    if (a < r) acc[q] = Sc<A>::fma(widen(ab[a * c + b]), xv, acc[q]);
                                }
                            }
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < RB; ++q)
                    for (int s = Kp * rpp; s < 64; s <<= 1) acc[q] = acc[q] + shfl_xor(acc[q], s);
                if (vok && p_l == 0) {
#pragma unroll
                    for (int q = 0; q < RB; ++q) {
                        const int a = a0 + a_l + q * rpp;
                        if (a < r) emit<T, FUSED>(g, y0 + a, x0 + a, v, acc[q], d1, d2);
                    }
                }
            }
        }
    }
}

// LDS tile of the SpMV per wave: a run's values (at most 64*rpp plus two
// partial 16-byte words), then the run's 64 x entries (kSpmvX onwards), then
// a zero element (kSpmvZero) that the columns past a block's width read.
constexpr int kSpmvTile(int rpp) { return 64 * rpp + 16 + 64 + 2; }
constexpr int kSpmvX(int rpp) { return 64 * rpp + 16; }
constexpr int kSpmvZero(int rpp) { return 64 * rpp + 16 + 64; }

// 16-byte vector of a storage type, for the coalesced LDS staging copies.
template <typename T>
struct Vec16 {
    static constexpr int n = 16 / sizeof(T);
    T v[16 / sizeof(T)];
};

// 16 raw bytes as a native vector type: staging words kept in registers
// across loop iterations stay registers (arrays of structs would not).
typedef int vbh_w4 __attribute__((ext_vector_type(4)));

// Words of one staged run of the LDS SpMV per lane.
template <typename T, int RPP>
constexpr int spmv_words() {
    return (kSpmvX(RPP) / Vec16<T>::n + 63) / 64;
}

// Issues the loads of one staged run into registers: its 16-byte value words
// and one x entry per lane -- entry b of the run's block j/RPP for lane j, so
// the run's x segments are read once, not once per row -- with clamped
// (unconditional, in-bounds) indices: a guarded load compiles to a branch with
// its own wait.
template <typename T, int RPP>
__device__ __forceinline__ void spmv_issue(const ApplyArgs<T>& g, int64_t vrow, int r, int nb, int ub, int mx, int mw,
                                           int lane, vbh_w4 (&stage)[spmv_words<T, RPP>()], T& xone, int& lead,
                                           int& nvec, int& pc_first, int& w, bool& act) {
    constexpr int CH = 64 / RPP;
    constexpr int EV = Vec16<T>::n;
    const int a_l = lane & (RPP - 1);
    const int p_l = lane / RPP;
    const int ulast = (ub + CH < nb ? ub + CH : nb) - 1;
    pc_first = lane_read(mw, ub) >> 8;
    const int wl = lane_read(mw, ulast);
    const int pc_end = (wl >> 8) + (wl & 255);
    const int64_t s0 = vrow + int64_t(r) * pc_first;  // first value of the run
    const int64_t sa = s0 & ~int64_t(EV - 1);         // 16-byte aligned start
    lead = int(s0 - sa);
    nvec = (lead + r * (pc_end - pc_first) + EV - 1) / EV;
    const vbh_w4* src = reinterpret_cast<const vbh_w4*>(g.vals + sa);
    const int u = ub + p_l;
    const int xo = __shfl(mx, u, 64);  // lanes past the row's slots hold x offset 0, width 0
    w = __shfl(mw, u, 64);
    act = u <= ulast && a_l < r;
#pragma unroll
    for (int q = 0; q < spmv_words<T, RPP>(); ++q) {
        // words past the run are skipped on a wave-uniform condition (a scalar
        // branch: no per-lane wait); inside the run the index is clamped
        if (q == 0 || q * 64 < nvec) {
            const int e = q * 64 + lane;
            // non-temporal: the value stream is read once, and evict-first
            // lines leave the L2 to the x entries, which are gathered again
            stage[q] = __builtin_nontemporal_load(src + (e < nvec ? e : nvec - 1));
        }
    }
    // lane = p*RPP + b loads x[xo_p + b] (clamped into the block's width;
    // complex types: zero past it)
    const int c = w & 255;
    const int cm = c > 0 ? c - 1 : 0;
    xone = g.X[int64_t(xo + (a_l < cm ? a_l : cm)) * g.ldx];
    if constexpr (Sc<T>::is_complex) xone = pick(a_l < c, xone, Sc<T>::zero());
}

// SpMV row loop for blocks up to RPP x RPP (RPP in {1, 2, 4, 8}); lds is this
// wave's 16-byte aligned tile of kSpmvTile(RPP) elements: a run of 64/RPP
// blocks holds at most 64*RPP values, plus up to two partial 16-byte words.
// Software-pipelined within each 64-slot chunk of a row: run n+1's loads are
// in flight while run n is computed from the tile.
template <typename T, int RPP, bool FUSED, bool ACC = false>
__device__ inline void row_loop_spmv(const ApplyArgs<T>& g, int r0, int r1, int lane, int wave, double& d1,
                                     double& d2, T* lds, const SeriesAcc<T>* sa = nullptr) {
    using A = typename Wide<T>::type;
    constexpr int CH = 64 / RPP;  // blocks per staged run = partitions
    const DevPattern& p = g.p;
    const int a_l = lane & (RPP - 1);
    if (lane == 0) lds[kSpmvZero(RPP)] = Sc<T>::zero();  // visible after the first run's wave barrier
    for (int i = r0 + wave; i < r1; i += kWaves) {
        const int r = p.rbs[i];
        const int kb = p.row_ptr[i], ke = p.row_ptr[i + 1];
        const int64_t vrow = p.voff[kb];
        A acc = Sc<A>::zero();
        for (int t0 = kb; t0 < ke; t0 += 64) {
            const int nb = ke - t0 < 64 ? ke - t0 : 64;
            int mx = 0, mw = 0;
            {
                const int2 mm = p.meta[t0 + (lane < nb ? lane : nb - 1)];
                mx = lane < nb ? mm.x : 0;
                mw = lane < nb ? mm.y : 0;
            }
            constexpr int NV = spmv_words<T, RPP>();
            vbh_w4 stage[NV];
            T xone;
            int lead = 0, nvec = 1, pc_first = 0, w = 0;
            bool act = false;
            spmv_issue<T, RPP>(g, vrow, r, nb, 0, mx, mw, lane, stage, xone, lead, nvec, pc_first, w, act);
            for (int ub = 0; ub < nb; ub += CH) {
                // this run's words to the tile (lanes past the run rewrite word
                // nvec-1 with the value it already holds), its x to xc
                {
                    vbh_w4* dstw = reinterpret_cast<vbh_w4*>(lds);
#pragma unroll
                    for (int q = 0; q < NV; ++q) {
                        if (q == 0 || q * 64 < nvec) {
                            const int e = q * 64 + lane;
                            dstw[e < nvec ? e : nvec - 1] = stage[q];
                        }
                    }
                }
                lds[kSpmvX(RPP) + lane] = xone;
                const int lead_c = lead, pc_first_c = pc_first, w_c = w;
                const bool act_c = act;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#ifndef VBH_SPMV_NOPIPE
                if (ub + CH < nb)  // overlaps the FMAs below
                    spmv_issue<T, RPP>(g, vrow, r, nb, ub + CH, mx, mw, lane, stage, xone, lead, nvec, pc_first, w,
                                       act);
#endif
                if (act_c) {
                    const int c = w_c & 255;
                    const T* ab = lds + lead_c + r * ((w_c >> 8) - pc_first_c) + a_l * c;
                    const T* xs = lds + kSpmvX(RPP) + (lane & ~(RPP - 1));  // this block's x entries
                    if constexpr (Sc<T>::is_complex) {
                        // columns b >= c of a narrower block multiply the zero
                        // element by a zero x entry: exact, without a select
                        // on the four words of the accumulator (propagation
                        // forward 68.9 -> 59.3 ms; for real types the select
                        // measured faster: 0.36 against 0.378 ms on the
                        // 1.54 GB SpMV)
                        const T* zp = lds + kSpmvZero(RPP);
#pragma unroll
                        for (int b = 0; b < RPP; ++b)
                            acc = Sc<A>::fma(widen(*(b < c ? ab + b : zp)), widen(xs[b]), acc);
                    } else {
                        const int cm = c > 0 ? c - 1 : 0;
#pragma unroll
                        for (int b = 0; b < RPP; ++b) {
                            const A t = Sc<A>::fma(widen(ab[b < cm ? b : cm]), widen(xs[b]), acc);
                            acc = pick(b < c, t, acc);
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#ifdef VBH_SPMV_NOPIPE
                if (ub + CH < nb)
                    spmv_issue<T, RPP>(g, vrow, r, nb, ub + CH, mx, mw, lane, stage, xone, lead, nvec, pc_first, w,
                                       act);
#endif
            }
        }
#pragma unroll
        for (int s = RPP; s < 64; s <<= 1) acc = acc + shfl_xor(acc, s);
        if (lane < r)
            emit<T, FUSED, ACC>(g, int64_t(p.roff[i]) + lane, int64_t(p.coff[i]) + lane, 0, acc, d1, d2, sa);
    }
}

// Per-column sums of the lane partials d1, d2 of a whole work-group, written
// to out[0..K) and out[K..2K) by wave 0 (K <= 64).  Every thread must call it.
__device__ inline void write_partials(int K, int lane, int wave, double d1, double d2, double* out) {
    __shared__ double red[kWaves][2][64];
    const int Kp = K >= 64 ? 64 : pow2ceil(K);
    for (int s = Kp; s < 64; s <<= 1) {
        d1 += __shfl_xor(d1, s, 64);
        d2 += __shfl_xor(d2, s, 64);
    }
    red[wave][0][lane] = d1;
    red[wave][1][lane] = d2;
    __syncthreads();
    if (wave == 0 && lane < K) {
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            s1 += red[w][0][lane];
            s2 += red[w][1][lane];
        }
        out[lane] = s1;
        out[K + lane] = s2;
    }
}

// Row-loop configuration chosen per launch: which kernel family and its
// compile-time parameters.
struct RowConfig {
    int spmv_rpp;  // > 0: LDS SpMV family with RPP = this
    int bs;        // SpMM family: uniform block size (4) or 0
    int rb;        // SpMM family: row accumulators per lane (1, 2, 4, 8)
};
RowConfig row_config(int K, int max_block, int uniform_bs);

// Lane-per-row family (lanes.hpp): S lanes per scalar row, KT columns each.
struct LaneShape {
    int S, KT;   // S == 0: family not applicable
    int P = 1;   // partitions of each row's slots (lane_parts)
    int rw = 0;  // > 0: the row-wave kernel instead, rw scalar rows per work-group
    int brw = 0;  // 1: the matrix-core class rows on the block-row wave kernel (vector ALUs)
};
// Columns per lane of the row-wave kernel (16 bytes of a row), or 0 when it
// does not serve (T, K): real types, K a multiple of it with K / it a power
// of two up to 64.
template <typename T>
inline int rowwave_cpl(int K) {
    if constexpr (Sc<T>::is_complex) {
        return 0;
    } else {
        const int cpl = int(16 / sizeof(T));
        if (K < cpl || K % cpl) return 0;
        const int lq = K / cpl;
        return (lq & (lq - 1)) == 0 && lq <= 64 ? cpl : 0;
    }
}
// Columns per lane of the block-row wave kernel (K = 16 * it: 1, 2 or 4), or
// 0 when it does not serve (T, K): real types only.
template <typename T>
inline int brw_cpl(int K) {
    if constexpr (Sc<T>::is_complex) {
        return 0;
    } else {
        return (K == 16 || K == 32 || K == 64) ? K / 16 : 0;
    }
}
// Instantiated shapes: real K <= 16 one lane per row (KT in 4, 8, 10, 12, 16),
// up to 64 with 2 or 4 lanes of 16 columns; complex KT = 8 with 1..8 lanes.
// K == 1 runs here only for mixed value/vector types (the SpMV families are
// single-type).
inline LaneShape lane_shape(bool complex, int K, bool mixed = false) {
    if (K == 1 && mixed) K = 2;
    if (K < 2 || K > 64) return {0, 0};
    if (complex) {
        if (K <= 4) return {1, 4};
        if (K <= 8) return {1, 8};
        return {K <= 16 ? 2 : K <= 32 ? 4 : 8, 8};
    }
    if (K <= 16) return {1, K <= 4 ? 4 : K <= 8 ? 8 : K <= 10 ? 10 : K <= 12 ? 12 : 16};
    return {K <= 32 ? 2 : 4, 16};
}
// Scalar rows handled by one work-group of the lane family.
inline int lane_rows_per_group(int S, int P = 1) { return kWaves * (64 / (S * P)); }
// Partitions of each row's slots: small launches split rows over more lanes,
// so every lane's dependent chain of block loads gets shorter and the chip
// fills.  Target: 4 waves on each of 1024 SIMDs for 4-byte vectors, 8 for
// wider ones (measured on the 200x200 graphene KPM step: FP32 is fastest at
// P = 1, FP64 at P = 2).  Quads (4x4 blocks) need 4 rows per lane group.
inline int lane_parts(int64_t rows, int S, int bs, int elem_bytes) {
    const int64_t target = int64_t(256) * 4 * (elem_bytes <= 4 ? 4 : 8) * 64;
    const int pmax = bs == 4 ? 16 / S : (64 / S < 8 ? 64 / S : 8);
    int P = 1;
#ifdef VBH_FORCE_PARTS
    while (P * 2 <= pmax && P < VBH_FORCE_PARTS) P *= 2;
    return P;
#endif
    while (P * 2 <= pmax && rows * S * P < target) P *= 2;
    return P;
}
// Work-groups of the lane family over n rows (a multiple of 8 for the XCD map).
inline int lane_groups(int64_t n_rows, int S, int P = 1) {
    const int64_t rpg = lane_rows_per_group(S, P);
    const int64_t g = (n_rows + rpg - 1) / rpg;
    return int(((g < 8 ? 8 : g) + 7) & ~int64_t(7));
}
// Widest vector (elements of T, at most 16 bytes, dividing KT and K, so no
// vector straddles column K) that every operand (base pointer, leading
// dimension) admits; null operands are skipped.
template <typename T>
int lane_vector(int KT, int K, std::initializer_list<std::pair<const void*, int64_t>> operands) {
    int vn = sizeof(T) >= 16 ? 1 : int(16 / sizeof(T));
    for (; vn > 1; vn >>= 1) {
        bool ok = KT % vn == 0 && K % vn == 0;
        for (const auto& o : operands)
            if (o.first && (reinterpret_cast<uintptr_t>(o.first) % (vn * sizeof(T)) != 0 || o.second % vn != 0))
                ok = false;
        if (ok) break;
    }
    return vn;
}

template <typename T, typename V>
constexpr bool is_mixed() {
    return !std::is_same_v<T, V>;
}
template <typename T, typename V>
inline LaneShape lane_shape_of(int K) {
    return lane_shape(Sc<T>::is_complex, K, is_mixed<T, V>());
}
// Whether launch_rows runs the lane family for these arguments (per-column
// dots stay with row_loop; mixed value/vector types exist only as lanes).
template <typename T, typename V>
inline bool uses_lanes(const ApplyArgs<T, V>& a) {
    return lane_shape_of<T, V>(a.K).S > 0 && a.dot_cols <= 1;
}
// Matrix-core class: block rows with kMfmaMinRows..16 rows of a real matrix
// run as 16x16 FP64 MFMA tiles, the other rows on the lane family.
constexpr int kMfmaRowsPerGroup = kWaves * kMfmaRowsPerWave;  // block rows per work-group
inline int mfma_groups(int n_big) {
    const int g = (n_big + kMfmaRowsPerGroup - 1) / kMfmaRowsPerGroup;
    return ((g < 8 ? 8 : g) + 7) & ~7;
}
template <typename T, typename V>
inline bool uses_mfma(const vbh_matrix* m, const ApplyArgs<T, V>& a) {
    return !Sc<T>::is_complex && a.p.n_big > 0 && m->max_block <= 16 && uses_lanes(a);
}
// Work-groups of the lane part (over the small rows when the matrix-core
// class runs apart) and of the matrix-core part.
template <typename T, typename V>
inline void split_groups(const vbh_matrix* m, const ApplyArgs<T, V>& a, int& g_lanes, int& g_mfma, int* parts = nullptr) {
    const bool mf = uses_mfma(m, a);
    const int64_t rows = mf ? a.p.n_small : m->n_rows;
    const int S = lane_shape_of<T, V>(a.K).S;
    const int P = lane_parts(rows, S, m->uniform_bs == 4 ? 4 : 0, int(sizeof(T)));
    if (parts) *parts = P;
    g_lanes = rows > 0 ? lane_groups(rows, S, P) : 0;
    g_mfma = mf ? mfma_groups(a.p.n_big) : 0;
}
// Work-groups (= partial slots) a launch_rows call with these arguments uses.
template <typename T, typename V>
inline int apply_groups(const vbh_matrix* m, const ApplyArgs<T, V>& a) {
    if (!uses_lanes(a)) return m->G;
    int gl, gm;
    split_groups(m, a, gl, gm);
    return gl + gm;
}

template <typename T, bool FUSED, typename V = T>
void launch_rows(const vbh_matrix* m, const ApplyArgs<T, V>& a, hipStream_t s);
template <typename T, typename V>
void launch_lanes(const ApplyArgs<T, V>& a, int64_t n_rows, LaneShape sh, int bs, hipStream_t s);
// Matrix-core class (a.G groups over p.big_brows); real types only.
template <typename T, typename V>
void launch_mfma(const ApplyArgs<T, V>& a, hipStream_t s);

// One matrix of a batched KPM launch.
template <typename T, typename V = T>
struct BatchItem {
    DevPattern p;
    const V* vals;
    T* buf[3];     // recursion vectors t_{n-1}, t_n, t_{n+1} rotate through these
    const T* v0;   // mode 1: the start vectors (W of the dot)
    const int* split;
    int64_t n_rows;
    int64_t lane_rows;  // rows of the lane launch: n_rows, or the small rows (row_list)
    int row_list;       // 1: the lane launch takes p.small_rows, the matrix-core launch p.big_brows
    int G;              // this matrix's lane (or row-loop) work-groups
    int g_base;         // its first work-group in the lane launch
    int g_base2;        // its first work-group in the matrix-core launch
    // block-row wave kernel (class rows on the vector ALUs): the column-major
    // FP32 values and the X row of every stored scalar column of each block
    // row (entries eoff[i] .. eoff[i+1])
    const float* vt;
    const int64_t* eoff;
    const int* ecol;
};
template <typename T, typename V>
void launch_batched_lanes(const BatchItem<T, V>* items, int nitems, int Gtot, int K, int ld, LaneShape sh, int step,
                          double alpha, double beta, double gamma0, double gamma, int mode, double* partials,
                          hipStream_t s);
template <typename T, typename V>
void launch_batched_rowwave(const BatchItem<T, V>* items, int nitems, int Gtot, int K, int ld, int rw, int step,
                            double alpha, double beta, double gamma0, double gamma, int mode, double* partials,
                            hipStream_t s);
template <typename T, typename V>
void launch_batched_brw(const BatchItem<T, V>* items, int nitems, int Gtot, int K, int ld, int step, double alpha,
                        double beta, double gamma0, double gamma, int mode, double* partials, hipStream_t s);
template <typename T, typename V>
void launch_batched_mfma(const BatchItem<T, V>* items, int nitems, int Gtot, int K, int ld, int step, double alpha,
                         double beta, double gamma0, double gamma, int mode, double* partials, hipStream_t s);
template <typename T, typename V = T>
ApplyArgs<T, V> make_args(const vbh_matrix* m, const V* vals, int K, const T* X, int64_t ldx, T* Y, int64_t ldy);
template <typename T>
const T* compute_values(vbh_matrix* m, hipStream_t s);
// One fused step through the K == 1 SpMV family with the series
// accumulation sa; returns false (nothing launched) when that family does not
// serve these arguments (lane family, blocks over 8, dots, mixed types): the
// caller then launches launch_rows and accumulates separately.
template <typename T, typename V>
bool launch_spmv_series(const vbh_matrix* m, const ApplyArgs<T, V>& a, const SeriesAcc<T>& sa, hipStream_t s);

template <typename T, typename V = T>
void apply_plain(vbh_matrix* m, const V* vals, int K, const void* X, int64_t ldx, void* Y, int64_t ldy, hipStream_t s);
template <typename T>
void apply_adjoint(vbh_matrix* m, const T* vals, int K, const void* X, int64_t ldx, void* Y, int64_t ldy,
                   hipStream_t s);
void reduce_partials(const double* partials, int G, int E, double* out, bool accumulate, hipStream_t s);

// dtype/precision dispatch: calls f(T*, V*) with typed null pointers of the
// vector type T and the matrix value type V.  precision 0: native (FP64 /
// complex128) throughout; 1: FP32 / complex64 throughout; 2: FP32 values
// against native vectors (the value stream halves, vectors keep full
// precision).
template <typename F>
void dispatch_compute(const vbh_matrix* m, int precision, F&& f) {
    VBH_REQUIRE(precision >= 0 && precision <= 2,
                "precision must be 0 (native), 1 (FP32) or 2 (FP32 values, native vectors)");
    if (m->dtype == VBH_F64) {
        if (precision == 0) f((double*)nullptr, (double*)nullptr);
        else if (precision == 1) f((float*)nullptr, (float*)nullptr);
        else f((double*)nullptr, (float*)nullptr);
    } else {
        if (precision == 0) f((c128*)nullptr, (c128*)nullptr);
        else if (precision == 1) f((c64*)nullptr, (c64*)nullptr);
        else f((c128*)nullptr, (c64*)nullptr);
    }
}

}  // namespace vbh
