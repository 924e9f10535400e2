// Adjoint of the Chebyshev propagation on gfx950: the stored Chebyshev basis,
// the Clenshaw recurrence with the adjoint operator, and the block-sampled
// outer-product sum (SDDMM) that turns both into the gradient with respect to
// the stored block values.
//
// Forward (examples/ex09_graphene_dynamics.py:102-133, vbcsr/utils.py:71-126):
//   t_0 = x, t_1 = H~ t_0, t_{n+1} = 2 H~ t_n - t_{n-1},  y = sum_n c_n t_n,
//   H~ = (H - center) / half_width.
// Reverse mode, for a real loss L and the gradient convention g_v = dL/dRe v +
// i dL/dIm v (torch's): with lambda = g_y,
//   b_n = conj(c_n) lambda + 2 H~^H b_{n+1} - b_{n+2}   (n >= 1, b_N = b_{N+1} = 0)
//   g_x = b_0 = conj(c_0) lambda + H~^H b_1 - b_2
//   g_{H~} = b_1 t_0^H + 2 sum_{n>=2} b_n t_{n-1}^H,   g_H = g_{H~} / half_width,
// restricted to the stored blocks (the pattern is fixed).  The basis t_n and
// the b_n are kept as column blocks of row-major n_rows x (nterms*K) arrays, so
// the SDDMM reads one contiguous run per row.
#include <algorithm>
#include <cstring>

#include "apply.hpp"

namespace vbh {

template <typename T>
__global__ void k_copy_rows(int64_t rows, int K, const T* __restrict__ in, int64_t ldi, T* __restrict__ out,
                            int64_t ldo) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= rows * K) return;
    const int64_t r = e / K;
    const int v = int(e - r * K);
    out[r * ldo + v] = in[r * ldi + v];
}

// G_blk(a,b) += alpha * sum_{q<L} U[row_a][q] conj(V[col_b][q]) for every
// stored block; lanes take consecutive entries of a block row's contiguous
// value run.  G is in the matrix's own type (FP64 accumulation).
template <typename T, typename G_t>
__global__ __launch_bounds__(kThreads) void k_sddmm(DevPattern p, int L, const T* __restrict__ U, int64_t ldu,
                                                    const T* __restrict__ V, int64_t ldv, double are, double aim,
                                                    G_t* __restrict__ Gv, const int* __restrict__ split, int Gn) {
    using A = typename Wide<T>::type;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int chunk = xcd_chunk(Gn);
    const A alpha = Sc<A>::make(are, aim);
    for (int i = split[chunk] + wave; i < split[chunk + 1]; i += kWaves) {
        const int r = p.rbs[i];
        const int kb = p.row_ptr[i], ke = p.row_ptr[i + 1];
        if (kb == ke) continue;
        const int64_t v0 = p.voff[kb];
        const int64_t len = p.voff[ke] - v0;
        const int64_t row0 = p.roff[i];
        for (int64_t e = lane; e < len; e += 64) {
            // slot of value e: the row's slots are in panel order
            int k = kb;
            while (k + 1 < ke && p.voff[k + 1] - v0 <= e) ++k;
            const int c = p.meta[k].y & 255;
            const int64_t in = e - (p.voff[k] - v0);
            const int a = int(in / c), b = int(in - int64_t(a) * c);
            const T* u = U + (row0 + a) * ldu;
            const T* v = V + (int64_t(p.meta[k].x) + b) * ldv;
            A acc = Sc<A>::zero();
#pragma unroll 4
            for (int q = 0; q < L; ++q) acc = Sc<A>::fma(widen(u[q]), widen(Sc<T>::conj(v[q])), acc);
            A g = acc * alpha;
            if constexpr (Sc<G_t>::is_complex) {
                Gv[v0 + e] = Gv[v0 + e] + G_t{g.re, g.im};
            } else {
                Gv[v0 + e] += Sc<A>::re(g);
            }
        }
    }
}

// b_n = conj(c) lam + s1 (h - center b1) - b2 over n_rows x K entries, where h
// = H^H b1 (absent for the first term), b1 = b_{n+1}, b2 = b_{n+2} (absent
// past the end).  All operands are row blocks of K columns with their own ld.
template <typename T>
__global__ void k_cheb_adj_combine(int64_t rows, int K, T* __restrict__ bn, int64_t ldb, const T* __restrict__ b1,
                                   const T* __restrict__ b2, const T* __restrict__ h, int64_t ldh,
                                   const T* __restrict__ lam, int64_t ldl, double cre, double cim, double s1,
                                   double center) {
    using A = typename Wide<T>::type;
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= rows * K) return;
    const int64_t r = e / K;
    const int v = int(e - r * K);
    // conj(c) * lambda
    A y = Sc<A>::make(cre, -cim) * widen(lam[r * ldl + v]);
    if constexpr (!Sc<T>::is_complex) y = cre * widen(lam[r * ldl + v]);
    if (h) y = y + rscale(s1, widen(h[r * ldh + v]) - rscale(center, widen(b1[r * ldb + v])));
    if (b2) y = y - widen(b2[r * ldb + v]);
    bn[r * ldb + v] = narrow_to<T>(y);
}

template <typename T, typename V>
static void cheb_basis(vbh_matrix* m, int K, const T* X, int64_t ldx, int nterms, T* Tb, int64_t ldt, double center,
                       double half, hipStream_t s) {
    const int64_t N = m->n_rows;
    if (N * K == 0 || nterms <= 0) return;
    const unsigned nb = unsigned((N * K + 255) / 256);
    hipLaunchKernelGGL((k_copy_rows<T>), dim3(nb), dim3(256), 0, s, N, K, X, ldx, Tb, ldt);
    VBH_CHECK_HIP(hipGetLastError());
    const V* vals = compute_values<V>(m, s);
    for (int n = 1; n < nterms; ++n) {
        const int st = n - 1;  // produces t_n from t_{n-1} (and t_{n-2})
        auto a = make_args<T, V>(m, vals, K, Tb + int64_t(st) * K, ldt, Tb + int64_t(n) * K, ldt);
        a.alpha = st == 0 ? 1.0 / half : 2.0 / half;
        a.beta = st == 0 ? -center / half : -2.0 * center / half;
        a.gamma = st == 0 ? 0.0 : -1.0;
        a.Z = st == 0 ? nullptr : Tb + int64_t(st - 1) * K;
        a.ldz = ldt;
        launch_rows<T, true, V>(m, a, s);
    }
}

template <typename T>
static void cheb_adjoint(vbh_matrix* m, int K, const T* lam, int64_t ldl, int nterms, const double* coef, T* B,
                         int64_t ldb, double center, double half, hipStream_t s) {
    const int64_t N = m->n_rows;
    if (N * K == 0 || nterms <= 0) return;
    const unsigned nb = unsigned((N * K + 255) / 256);
    T* h = (T*)scratch(sizeof(T) * size_t(N) * K, 4);
    const T* vals = compute_values<T>(m, s);
    for (int n = nterms - 1; n >= 0; --n) {
        T* bn = B + int64_t(n) * K;
        const T* b1 = n + 1 < nterms ? B + int64_t(n + 1) * K : nullptr;
        const T* b2 = n + 2 < nterms ? B + int64_t(n + 2) * K : nullptr;
        if (b1) apply_adjoint<T>(m, vals, K, b1, ldb, h, K, s);
        const double s1 = (n == 0 ? 1.0 : 2.0) / half;
        hipLaunchKernelGGL((k_cheb_adj_combine<T>), dim3(nb), dim3(256), 0, s, N, K, bn, ldb, b1, b2,
                           b1 ? (const T*)h : nullptr, int64_t(K), lam, ldl, coef[2 * n], coef[2 * n + 1], s1, center);
        VBH_CHECK_HIP(hipGetLastError());
    }
}

}  // namespace vbh

using namespace vbh;

extern "C" {

VBH_API int vbh_cheb_basis(vbh_matrix_t m, int K, const void* X, int ldx, int nterms, void* T, int ldt, double center,
                           double half_width, int precision, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m && X && T, "null argument");
        VBH_REQUIRE(m->nbc == m->nbr, "the Chebyshev basis needs a square local pattern without ghosts");
        VBH_REQUIRE(K >= 1 && nterms >= 1 && ldx >= K && ldt >= int64_t(nterms) * K, "ldt must be >= nterms*K");
        VBH_REQUIRE(half_width > 0.0, "half_width must be positive");
        VBH_REQUIRE(precision == 0 || precision == 1, "precision must be 0 or 1");
        dispatch_compute(m, precision, [&](auto* tag, auto* vtag) {
            using Tt = std::remove_pointer_t<decltype(tag)>;
            using V = std::remove_pointer_t<decltype(vtag)>;
            cheb_basis<Tt, V>(m, K, (const Tt*)X, ldx, nterms, (Tt*)T, ldt, center, half_width, (hipStream_t)stream);
        });
    });
}

VBH_API int vbh_cheb_adjoint(vbh_matrix_t m, int K, const void* L, int ldl, int nterms, const double* coef, void* B,
                             int ldb, double center, double half_width, int precision, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m && L && B && coef, "null argument");
        VBH_REQUIRE(m->nbc == m->nbr, "the Chebyshev adjoint needs a square local pattern without ghosts");
        VBH_REQUIRE(K >= 1 && nterms >= 1 && ldl >= K && ldb >= int64_t(nterms) * K, "ldb must be >= nterms*K");
        VBH_REQUIRE(half_width > 0.0, "half_width must be positive");
        VBH_REQUIRE(precision == 0 || precision == 1, "precision must be 0 or 1");
        dispatch_compute(m, precision, [&](auto* tag, auto*) {
            using Tt = std::remove_pointer_t<decltype(tag)>;
            cheb_adjoint<Tt>(m, K, (const Tt*)L, ldl, nterms, coef, (Tt*)B, ldb, center, half_width,
                             (hipStream_t)stream);
        });
    });
}

VBH_API int vbh_sddmm(vbh_matrix_t m, int L, const void* U, int ldu, const void* V, int ldv, const double* alpha,
                      void* G, int precision, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m && G && alpha, "null argument");
        VBH_REQUIRE(L >= 0 && ldu >= L && ldv >= L, "leading dimensions must be >= L");
        VBH_REQUIRE(precision == 0 || precision == 1, "precision must be 0 or 1");
        if (m->nbr == 0 || L == 0 || m->nnz == 0) return;
        VBH_REQUIRE(U && V, "null U or V");
        hipStream_t s = (hipStream_t)stream;
        dispatch_compute(m, precision, [&](auto* tag, auto*) {
            using Tt = std::remove_pointer_t<decltype(tag)>;
            // G has the matrix's own type: double for real, complex128 for complex
            if constexpr (!Sc<Tt>::is_complex) {
                hipLaunchKernelGGL((k_sddmm<Tt, double>), dim3(m->G), dim3(kThreads), 0, s, m->pattern(), L,
                                   (const Tt*)U, int64_t(ldu), (const Tt*)V, int64_t(ldv), alpha[0], alpha[1],
                                   (double*)G, (const int*)m->split, m->G);
            } else {
                hipLaunchKernelGGL((k_sddmm<Tt, c128>), dim3(m->G), dim3(kThreads), 0, s, m->pattern(), L,
                                   (const Tt*)U, int64_t(ldu), (const Tt*)V, int64_t(ldv), alpha[0], alpha[1],
                                   (c128*)G, (const int*)m->split, m->G);
            }
        });
        VBH_CHECK_HIP(hipGetLastError());
    });
}

}  // extern "C"
