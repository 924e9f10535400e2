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
//   coalesced loads, then every lane forms its block-row dot product from LDS
//   against x gathered from global memory.
#pragma once

#include "common.hpp"

namespace vbh {

template <typename T>
struct ApplyArgs {
    DevPattern p;
    const T* vals;
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
    double* partials;  // [G][2K] per-work-group FP64 partial dots, or null
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
    const int* t_row;
    const int* t_slot;
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

// Fused epilogue for one output element (row yr, column v).
template <typename T, bool FUSED>
__device__ inline void emit(const ApplyArgs<T>& g, int64_t yr, int64_t xr, int v, typename Wide<T>::type acc,
                            double& d1, double& d2) {
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

constexpr int kSpmvTile(int rpp) { return 64 * rpp + 16; }

// 16-byte vector of a storage type, for the coalesced LDS staging copies.
template <typename T>
struct Vec16 {
    static constexpr int n = 16 / sizeof(T);
    T v[16 / sizeof(T)];
};

// SpMV row loop for blocks up to RPP x RPP (RPP in {1, 2, 4, 8}); lds is this
// wave's 16-byte aligned tile of kSpmvTile(RPP) elements: a run of 64/RPP
// blocks holds at most 64*RPP values, plus up to two partial 16-byte words.
template <typename T, int RPP, bool FUSED>
__device__ inline void row_loop_spmv(const ApplyArgs<T>& g, int r0, int r1, int lane, int wave, double& d1,
                                     double& d2, T* lds) {
    using A = typename Wide<T>::type;
    constexpr int CH = 64 / RPP;  // blocks per staged run = partitions
    constexpr int EV = Vec16<T>::n;
    const DevPattern& p = g.p;
    const int a_l = lane & (RPP - 1);
    const int p_l = lane / RPP;
    for (int i = r0 + wave; i < r1; i += kWaves) {
        const int r = p.rbs[i];
        const int kb = p.row_ptr[i], ke = p.row_ptr[i + 1];
        const int64_t vrow = p.voff[kb];
        A acc = Sc<A>::zero();
        for (int t0 = kb; t0 < ke; t0 += 64) {
            const int nb = ke - t0 < 64 ? ke - t0 : 64;
            int mx = 0, mw = 0;
            if (lane < nb) {
                const int2 mm = p.meta[t0 + lane];
                mx = mm.x;
                mw = mm.y;
            }
            for (int ub = 0; ub < nb; ub += CH) {
                const int ulast = (ub + CH < nb ? ub + CH : nb) - 1;
                const int pc_first = lane_read(mw, ub) >> 8;
                const int wl = lane_read(mw, ulast);
                const int pc_end = (wl >> 8) + (wl & 255);
                const int64_t s0 = vrow + int64_t(r) * pc_first;      // first value of the run
                const int64_t sa = s0 & ~int64_t(EV - 1);             // 16-byte aligned start
                const int lead = int(s0 - sa);
                const int nvec = (lead + r * (pc_end - pc_first) + EV - 1) / EV;
                const Vec16<T>* src = reinterpret_cast<const Vec16<T>*>(g.vals + sa);
                Vec16<T>* dst = reinterpret_cast<Vec16<T>*>(lds);
                for (int e = lane; e < nvec; e += 64) dst[e] = src[e];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const int u = ub + p_l;
                const int xo = __shfl(mx, u, 64);
                const int w = __shfl(mw, u, 64);
                if (u <= ulast && a_l < r) {
                    const int c = w & 255;
                    const T* ab = lds + lead + r * ((w >> 8) - pc_first) + a_l * c;
                    const T* xb = g.X + int64_t(xo) * g.ldx;
                    for (int b = 0; b < c; ++b) acc = Sc<A>::fma(widen(ab[b]), widen(xb[int64_t(b) * g.ldx]), acc);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
#pragma unroll
        for (int s = RPP; s < 64; s <<= 1) acc = acc + shfl_xor(acc, s);
        if (lane < r) emit<T, FUSED>(g, int64_t(p.roff[i]) + lane, int64_t(p.coff[i]) + lane, 0, acc, d1, d2);
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

template <typename T, bool FUSED>
void launch_rows(const vbh_matrix* m, const ApplyArgs<T>& a, hipStream_t s);
template <typename T>
ApplyArgs<T> make_args(const vbh_matrix* m, const T* vals, int K, const T* X, int64_t ldx, T* Y, int64_t ldy);
template <typename T>
const T* compute_values(vbh_matrix* m, hipStream_t s);
template <typename T>
void apply_plain(vbh_matrix* m, const T* vals, int K, const void* X, int64_t ldx, void* Y, int64_t ldy, hipStream_t s);
template <typename T>
void apply_adjoint(vbh_matrix* m, const T* vals, int K, const void* X, int64_t ldx, void* Y, int64_t ldy,
                   hipStream_t s);
void reduce_partials(const double* partials, int G, int E, double* out, bool accumulate, hipStream_t s);

// dtype/precision dispatch: calls f with a typed null pointer of the compute type.
template <typename F>
void dispatch_compute(const vbh_matrix* m, int precision, F&& f) {
    VBH_REQUIRE(precision == 0 || precision == 1, "precision must be 0 (native) or 1 (FP32)");
    if (m->dtype == VBH_F64) {
        if (precision == 0) f((double*)nullptr);
        else f((float*)nullptr);
    } else {
        if (precision == 0) f((c128*)nullptr);
        else f((c64*)nullptr);
    }
}

}  // namespace vbh
