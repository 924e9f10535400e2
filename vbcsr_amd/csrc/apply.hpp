// Internal interface of the apply kernels (shared by apply.hip and kpm.hip).
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

template <typename T, int BS>
__device__ inline T block_row_dot(const T* __restrict__ blk, const T* __restrict__ xb, int64_t ldx, int c, T acc) {
    if constexpr (BS > 0) {
        T a[BS];
        if constexpr ((BS * sizeof(T)) % 16 == 0) {
            const T* q = static_cast<const T*>(__builtin_assume_aligned(blk, 16));
#pragma unroll
            for (int b = 0; b < BS; ++b) a[b] = q[b];
        } else {
#pragma unroll
            for (int b = 0; b < BS; ++b) a[b] = blk[b];
        }
#pragma unroll
        for (int b = 0; b < BS; ++b) acc = Sc<T>::fma(a[b], xb[int64_t(b) * ldx], acc);
        return acc;
    } else {
#pragma unroll 4
        for (int b = 0; b < c; ++b) acc = Sc<T>::fma(blk[b], xb[int64_t(b) * ldx], acc);
        return acc;
    }
}

template <typename T, int BS, bool FUSED>
__device__ inline void row_loop(const ApplyArgs<T>& g, int r0, int r1, int lane, int wave, double& d1, double& d2) {
    const DevPattern& p = g.p;
    const int K = g.K;
    const int Kp = K >= 64 ? 64 : pow2ceil(K);
    const int lgK = __builtin_ctz(Kp);
    const int v_l = lane & (Kp - 1);
    for (int i = r0 + wave; i < r1; i += kWaves) {
        const int r = BS > 0 ? BS : p.rbs[i];
        const int kb = p.row_ptr[i], ke = p.row_ptr[i + 1];
        const int Ar = 64 >> lgK;
        const int rp = pow2ceil(r);
        const int rpp = rp <= Ar ? rp : Ar;
        const int P = Ar / rpp;
        const int lgR = __builtin_ctz(rpp);
        const int a_l = (lane >> lgK) & (rpp - 1);
        const int p_l = lane >> (lgK + lgR);
        const int64_t yrow0 = BS > 0 ? int64_t(i) * BS : p.roff[i];
        const int64_t xown0 = BS > 0 ? int64_t(i) * BS : p.coff[i];
        for (int a0 = 0; a0 < r; a0 += rpp) {
            const int a = a0 + a_l;
            for (int v0 = 0; v0 < K; v0 += Kp) {
                const int v = v0 + v_l;
                const bool active = a < r && v < K;
                T acc = Sc<T>::zero();
                if (active) {
                    for (int k = kb + p_l; k < ke; k += P) {
                        const int j = p.col_ind[k];
                        const int c = BS > 0 ? BS : p.cbs[j];
                        const int64_t vo = BS > 0 ? int64_t(k) * (BS * BS) : p.voff[k];
                        const int64_t xo = BS > 0 ? int64_t(j) * BS : p.coff[j];
                        acc = block_row_dot<T, BS>(g.vals + vo + int64_t(a) * c, g.X + xo * g.ldx + v, g.ldx, c, acc);
                    }
                }
                for (int s = Kp * rpp; s < 64; s <<= 1) acc = acc + shfl_xor(acc, s);
                if (active && p_l == 0) {
                    const int64_t yr = yrow0 + a;
                    T y = acc;
                    if constexpr (FUSED) {
                        const T xo = g.X[(xown0 + a) * g.ldx + v];
                        y = rscale(g.alpha, acc);
                        if (g.beta != 0.0) y = y + rscale(g.beta, xo);
                        if (g.gamma != 0.0) y = y + rscale(g.gamma, g.Z[yr * g.ldz + v]);
                        d1 += redot(g.W ? g.W[yr * g.ldw + v] : y, y);
                        d2 += redot(xo, y);
                    }
                    g.Y[yr * g.ldy + v] = y;
                }
            }
        }
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
