// Block SpMV / SpMM and their adjoints on gfx950.
//
// One wavefront owns one block row at a time.  Its 64 lanes are split into
//   v  (fastest)  : Kp = min(64, pow2ceil(K)) vector columns,
//   a             : scalar rows of the block row handled in this pass,
//   p  (slowest)  : P partitions of the row's blocks (block k goes to
//                   partition (k - row_start) mod P),
// so a lane accumulates sum_k sum_b A_k[a][b] X[col_k + b][v] over its
// partition; the P partial sums are merged with cross-lane xor shuffles.  A
// block's row a is read as one contiguous run (vectorised when the block size
// is a compile-time constant), X rows are read coalesced across the v lanes,
// and every Y element is written exactly once -- no atomics, no Y zero pass.
// Work-groups take work-balanced contiguous chunks of block rows (weights
// sum r*c per row), remapped so that each XCD streams one contiguous run of
// rows and keeps the X rows it gathers in its own L2.
//
// The optional fused epilogue computes Y = alpha*AX + beta*X_own + gamma*Z and
// FP64 per-column partial dot products of Y with W (or Y) and with X_own, one
// partial per work-group, written (not accumulated) into partials[g][2K].
#include "apply.hpp"

namespace vbh {

template <typename T, int BS, int RB, bool FUSED>
__global__ __launch_bounds__(kThreads) void k_apply_rows(ApplyArgs<T> g) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int chunk = xcd_chunk(g.G);
    double d1 = 0.0, d2 = 0.0;
    row_loop<T, BS, RB, FUSED>(g, g.split[chunk], g.split[chunk + 1], lane, wave, d1, d2);
    if constexpr (FUSED) {
        if (g.partials != nullptr) {
            const int E = g.dot_cols > 1 ? g.K : 1;
            write_partials(E, lane, wave, d1, d2, g.partials + int64_t(chunk) * 2 * E);
        }
    }
}

template <typename T, int RPP, bool FUSED>
__global__ __launch_bounds__(kThreads) void k_apply_spmv(ApplyArgs<T> g) {
    __shared__ __attribute__((aligned(16))) T tile[kWaves][kSpmvTile(RPP)];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int chunk = xcd_chunk(g.G);
    double d1 = 0.0, d2 = 0.0;
    row_loop_spmv<T, RPP, FUSED>(g, g.split[chunk], g.split[chunk + 1], lane, wave, d1, d2, tile[wave]);
    if constexpr (FUSED) {
        if (g.partials != nullptr) write_partials(1, lane, wave, d1, d2, g.partials + int64_t(chunk) * 2);
    }
}

// The fused SpMV step with the series accumulation (no dots).
template <typename T, int RPP>
__global__ __launch_bounds__(kThreads) void k_apply_spmv_series(ApplyArgs<T> g, SeriesAcc<T> sa) {
    __shared__ __attribute__((aligned(16))) T tile[kWaves][kSpmvTile(RPP)];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int chunk = xcd_chunk(g.G);
    double d1 = 0.0, d2 = 0.0;
    row_loop_spmv<T, RPP, true, true>(g, g.split[chunk], g.split[chunk + 1], lane, wave, d1, d2, tile[wave], &sa);
}

RowConfig row_config(int K, int max_block, int uniform_bs) {
    RowConfig c{0, 0, 1};
    if (K == 1 && max_block <= 8) {
        c.spmv_rpp = pow2ceil(max_block);
        return c;
    }
    c.bs = uniform_bs == 4 ? 4 : 0;
    const int Kp = K >= 64 ? 64 : pow2ceil(K);
    const int NG = 64 / Kp;
    const int need = (max_block + NG - 1) / NG;
    c.rb = need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 8;
    return c;
}

// Adjoint: output block column j = sum over incoming (row i, slot k) of
// A_k^H X_i; the lane map is the forward one with the column's incoming list
// in place of the row's blocks and the block read down a column.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_apply_adjoint(AdjointArgs<T> g) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int chunk = xcd_chunk(g.G);
    const int c0 = g.split[chunk], c1 = g.split[chunk + 1];
    const DevPattern& p = g.p;
    const int K = g.K;
    const int Kp = K >= 64 ? 64 : pow2ceil(K);
    const int lgK = __builtin_ctz(Kp);
    const int v_l = lane & (Kp - 1);
    for (int j = c0 + wave; j < c1; j += kWaves) {
        const int c = p.cbs[j];
        const int qb = g.t_ptr[j], qe = g.t_ptr[j + 1];
        const int Ar = 64 >> lgK;
        const int cp = pow2ceil(c);
        const int cpp = cp <= Ar ? cp : Ar;
        const int P = Ar / cpp;
        const int lgC = __builtin_ctz(cpp);
        const int b_l = (lane >> lgK) & (cpp - 1);
        const int p_l = lane >> (lgK + lgC);
        for (int b0 = 0; b0 < c; b0 += cpp) {
            const int b = b0 + b_l;
            for (int v0 = 0; v0 < K; v0 += Kp) {
                const int v = v0 + v_l;
                const bool active = b < c && v < K;
                typename Wide<T>::type acc = Sc<typename Wide<T>::type>::zero();
                if (active) {
                    for (int q = qb + p_l; q < qe; q += P) {
                        const int2 rr = g.t_rows[q];  // (first scalar row, rows) of the block row
                        const int r = rr.y;
                        const T* blk = g.vals + g.t_voff[q] + b;
                        const T* xr = g.X + int64_t(rr.x) * g.ldx + v;
                        for (int a = 0; a < r; ++a)
                            acc = Sc<typename Wide<T>::type>::fma(widen(Sc<T>::conj(blk[int64_t(a) * c])), widen(xr[int64_t(a) * g.ldx]), acc);
                    }
                }
                for (int s = Kp * cpp; s < 64; s <<= 1) acc = acc + shfl_xor(acc, s);
                if (active && p_l == 0) g.Y[(int64_t(p.coff[j]) + b) * g.ldy + v] = narrow_to<T>(acc);
            }
        }
    }
}

// Sum per-work-group partials: out[e] (+)= sum_g partials[g][e], e < E.
__global__ void k_reduce_partials(const double* __restrict__ partials, int G, int E, double* __restrict__ out,
                                  int accumulate) {
    const int e = blockIdx.x;
    double s = 0.0;
    for (int gg = threadIdx.x; gg < G; gg += blockDim.x) s += partials[int64_t(gg) * E + e];
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int m = 128; m >= 1; m >>= 1) {
        if (threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[e] = accumulate ? out[e] + red[0] : red[0];
}

template <typename T, int BS, bool FUSED>
static void launch_spmm_family(int rb, const ApplyArgs<T>& a, hipStream_t s) {
    const dim3 grid(a.G), block(kThreads);
    switch (rb) {
        case 1: hipLaunchKernelGGL((k_apply_rows<T, BS, 1, FUSED>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_apply_rows<T, BS, 2, FUSED>), grid, block, 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_apply_rows<T, BS, 4, FUSED>), grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL((k_apply_rows<T, BS, 8, FUSED>), grid, block, 0, s, a); break;
    }
}

// Single-type row families: LDS SpMV, row_loop SpMM (defined below).
template <typename T, bool FUSED>
static void launch_single(const vbh_matrix* m, const ApplyArgs<T>& a, hipStream_t s);

template <typename T, bool FUSED, typename V>
void launch_rows(const vbh_matrix* m, const ApplyArgs<T, V>& a, hipStream_t s) {
    if (uses_lanes(a)) {
        LaneShape sh = lane_shape_of<T, V>(a.K);
        const bool mf = uses_mfma(m, a);
        int gl, gm;
        split_groups(m, a, gl, gm, &sh.P);
        if (gl > 0) {
            ApplyArgs<T, V> b = a;
            b.G = gl;
            b.row_list = mf ? 1 : 0;
            b.parts = sh.P;
            launch_lanes<T, V>(b, mf ? a.p.n_small : m->n_rows, sh, m->uniform_bs == 4 ? 4 : 0, s);
        }
        if (mf) {
            ApplyArgs<T, V> c = a;
            c.G = gm;
            c.partials = a.partials ? a.partials + 2 * int64_t(gl) : nullptr;
            launch_mfma<T, V>(c, s);
        }
        return;
    }
    if constexpr (is_mixed<T, V>()) {
        throw Error(VBH_ERR_NOTSUP,
                    "precision 2 (FP32 values, native vectors) supports 1 <= K <= 64 and column-summed dots only");
    } else {
        launch_single<T, FUSED>(m, a, s);
    }
}

template <typename T, bool FUSED>
static void launch_single(const vbh_matrix* m, const ApplyArgs<T>& a, hipStream_t s) {
    const RowConfig c = row_config(a.K, m->max_block, m->uniform_bs);
    const dim3 grid(a.G), block(kThreads);
    if (c.spmv_rpp > 0) {
        switch (c.spmv_rpp) {
            case 1: hipLaunchKernelGGL((k_apply_spmv<T, 1, FUSED>), grid, block, 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_apply_spmv<T, 2, FUSED>), grid, block, 0, s, a); break;
            case 4: hipLaunchKernelGGL((k_apply_spmv<T, 4, FUSED>), grid, block, 0, s, a); break;
            default: hipLaunchKernelGGL((k_apply_spmv<T, 8, FUSED>), grid, block, 0, s, a); break;
        }
    } else if (c.bs == 4) {
        launch_spmm_family<T, 4, FUSED>(c.rb, a, s);
    } else {
        launch_spmm_family<T, 0, FUSED>(c.rb, a, s);
    }
    VBH_CHECK_HIP(hipGetLastError());
}

template <typename T, typename V>
bool launch_spmv_series(const vbh_matrix* m, const ApplyArgs<T, V>& a, const SeriesAcc<T>& sa, hipStream_t s) {
    if constexpr (is_mixed<T, V>()) {
        return false;
    } else {
        if (uses_lanes(a) || a.partials != nullptr) return false;
        const RowConfig c = row_config(a.K, m->max_block, m->uniform_bs);
        if (c.spmv_rpp <= 0) return false;
        const dim3 grid(a.G), block(kThreads);
        switch (c.spmv_rpp) {
            case 1: hipLaunchKernelGGL((k_apply_spmv_series<T, 1>), grid, block, 0, s, a, sa); break;
            case 2: hipLaunchKernelGGL((k_apply_spmv_series<T, 2>), grid, block, 0, s, a, sa); break;
            case 4: hipLaunchKernelGGL((k_apply_spmv_series<T, 4>), grid, block, 0, s, a, sa); break;
            default: hipLaunchKernelGGL((k_apply_spmv_series<T, 8>), grid, block, 0, s, a, sa); break;
        }
        VBH_CHECK_HIP(hipGetLastError());
        return true;
    }
}

template <typename T, typename V>
ApplyArgs<T, V> make_args(const vbh_matrix* m, const V* vals, int K, const T* X, int64_t ldx, T* Y, int64_t ldy) {
    // row-class lists of the default matrix-core threshold, built on first use
    // (before any group count is derived from them; never inside a capture:
    // the captured batched launches take their lists beforehand)
    const auto& L = row_classes(const_cast<vbh_matrix*>(m), kMfmaMinRows, nullptr);
    ApplyArgs<T, V> a{};
    a.p = m->pattern(L);
    a.parts = 1;
    a.vals = vals;
    a.K = K;
    a.X = X;
    a.ldx = ldx;
    a.Y = Y;
    a.ldy = ldy;
    a.alpha = 1.0;
    a.split = m->split;
    a.G = m->G;
    return a;
}

template <typename T, typename V>
void apply_plain(vbh_matrix* m, const V* vals, int K, const void* X, int64_t ldx, void* Y, int64_t ldy, hipStream_t s) {
    if (m->nbr == 0 || K == 0) return;
    auto a = make_args<T, V>(m, vals, K, (const T*)X, ldx, (T*)Y, ldy);
    launch_rows<T, false, V>(m, a, s);
}

template <typename T>
void apply_adjoint(vbh_matrix* m, const T* vals, int K, const void* X, int64_t ldx, void* Y, int64_t ldy,
                   hipStream_t s) {
    if (m->nbc == 0 || K == 0) return;
    ensure_adjoint_plan(m, s);
    AdjointArgs<T> a{};
    a.p = m->pattern();
    a.vals = vals;
    a.K = K;
    a.X = (const T*)X;
    a.ldx = ldx;
    a.Y = (T*)Y;
    a.ldy = ldy;
    a.t_ptr = m->t_ptr;
    a.t_voff = m->t_voff;
    a.t_rows = m->t_rows;
    a.split = m->t_split;
    a.G = m->tG;
    hipLaunchKernelGGL(k_apply_adjoint<T>, dim3(a.G), dim3(kThreads), 0, s, a);
    VBH_CHECK_HIP(hipGetLastError());
}

void reduce_partials(const double* partials, int G, int E, double* out, bool accumulate, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(E), dim3(256), 0, s, partials, G, E, out, accumulate ? 1 : 0);
    VBH_CHECK_HIP(hipGetLastError());
}

template <typename T>
const T* compute_values(vbh_matrix* m, hipStream_t s) {
    if constexpr (std::is_same_v<T, float> || std::is_same_v<T, c64>) {
        ensure_vals32(m, s);
        return (const T*)m->vals32;
    } else {
        return (const T*)m->vals;
    }
}

// explicit instantiations used by kpm.hip
#define VBH_INST2(T, V)                                                                                          \
    template void launch_rows<T, true, V>(const vbh_matrix*, const ApplyArgs<T, V>&, hipStream_t);              \
    template void launch_rows<T, false, V>(const vbh_matrix*, const ApplyArgs<T, V>&, hipStream_t);             \
    template ApplyArgs<T, V> make_args<T, V>(const vbh_matrix*, const V*, int, const T*, int64_t, T*, int64_t); \
    template void apply_plain<T, V>(vbh_matrix*, const V*, int, const void*, int64_t, void*, int64_t, hipStream_t); \
    template bool launch_spmv_series<T, V>(const vbh_matrix*, const ApplyArgs<T, V>&, const SeriesAcc<T>&, hipStream_t);
#define VBH_INST(T)                                                          \
    VBH_INST2(T, T)                                                          \
    template const T* compute_values<T>(vbh_matrix*, hipStream_t);          \
    template void apply_adjoint<T>(vbh_matrix*, const T*, int, const void*, int64_t, void*, int64_t, hipStream_t);
VBH_INST(double)
VBH_INST(float)
VBH_INST(c128)
VBH_INST(c64)
VBH_INST2(double, float)
VBH_INST2(c128, c64)
#undef VBH_INST
#undef VBH_INST2

}  // namespace vbh

using namespace vbh;

extern "C" {

VBH_API int vbh_spmm(vbh_matrix_t m, int K, const void* X, int ldx, void* Y, int ldy, int precision, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m, "null matrix");
        VBH_REQUIRE(K >= 0 && ldx >= K && ldy >= K, "leading dimensions must be >= K");
        hipStream_t s = (hipStream_t)stream;
        dispatch_compute(m, precision, [&](auto* tag, auto* vtag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            using V = std::remove_pointer_t<decltype(vtag)>;
            apply_plain<T, V>(m, compute_values<V>(m, s), K, X, ldx, Y, ldy, s);
        });
    });
}

VBH_API int vbh_spmv(vbh_matrix_t m, const void* x, void* y, int precision, void* stream) {
    return vbh_spmm(m, 1, x, 1, y, 1, precision, stream);
}

VBH_API int vbh_spmm_adjoint(vbh_matrix_t m, int K, const void* X, int ldx, void* Y, int ldy, int precision,
                             void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m, "null matrix");
        VBH_REQUIRE(K >= 0 && ldx >= K && ldy >= K, "leading dimensions must be >= K");
        hipStream_t s = (hipStream_t)stream;
        VBH_REQUIRE(precision != 2, "the adjoint supports precision 0 and 1");
        dispatch_compute(m, precision, [&](auto* tag, auto*) {
            using T = std::remove_pointer_t<decltype(tag)>;
            apply_adjoint<T>(m, compute_values<T>(m, s), K, X, ldx, Y, ldy, s);
        });
    });
}

VBH_API int vbh_spmv_adjoint(vbh_matrix_t m, const void* x, void* y, int precision, void* stream) {
    return vbh_spmm_adjoint(m, 1, x, 1, y, 1, precision, stream);
}

VBH_API int vbh_spmm_fused(vbh_matrix_t m, int K, const void* X, int ldx, void* Y, int ldy, double alpha,
                           double beta, double gamma, const void* Z, int ldz, const void* W, int ldw, double* dots,
                           int precision, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m, "null matrix");
        VBH_REQUIRE(K >= 0 && ldx >= K && ldy >= K, "leading dimensions must be >= K");
        VBH_REQUIRE(gamma == 0.0 || (Z && ldz >= K), "gamma != 0 needs Z with ldz >= K");
        VBH_REQUIRE(!W || ldw >= K, "ldw must be >= K");
        VBH_REQUIRE(!dots || K <= 64, "fused dot products support K <= 64");
        // the epilogue reads X at the owned rows (column i of row i): needs nbc >= nbr
        VBH_REQUIRE(m->nbc >= m->nbr, "fused apply needs the owned columns of a square pattern");
        if (m->nbr == 0 || K == 0) return;
        hipStream_t s = (hipStream_t)stream;
        VBH_REQUIRE(precision != 2 || !dots, "per-column dots support precision 0 and 1");
        dispatch_compute(m, precision, [&](auto* tag, auto* vtag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            using V = std::remove_pointer_t<decltype(vtag)>;
            auto a = make_args<T, V>(m, compute_values<V>(m, s), K, (const T*)X, ldx, (T*)Y, ldy);
            a.alpha = alpha;
            a.beta = beta;
            a.gamma = gamma;
            a.Z = (const T*)Z;
            a.ldz = ldz;
            a.W = (const T*)W;
            a.ldw = ldw;
            a.dot_cols = dots ? K : 0;
            const int G = apply_groups(m, a);
            double* part = nullptr;
            if (dots) part = (double*)scratch(sizeof(double) * 2 * K * G, 2);
            a.partials = part;
            launch_rows<T, true, V>(m, a, s);
            if (dots) reduce_partials(part, G, 2 * K, dots, true, s);
        });
    });
}

}  // extern "C"
