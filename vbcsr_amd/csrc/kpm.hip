// KPM moments and Chebyshev series on gfx950.
//
// The reference composes the Chebyshev recursion in Python from mult_dense,
// scale/axpy and bdot (examples/ex08_graphene_dos.py:95-114,
// examples/inp_nanocrystal_dos/kpm.py:46-73, vbcsr/utils.py:71-126): per
// moment one SpMM plus ~5 full-vector passes.  Here one fused kernel per
// recursion step reads A, X = t_n, Z = t_{n-1} once, writes
// t_{n+1} = 2 H~ t_n - t_{n-1}, and emits per-work-group FP64 partial dots;
// the partials of all steps are reduced once at the end.  Modes:
//   0  plain: separate SpMM, axpby and dot kernels, one moment per matvec
//      (the unfused baseline the fused path is measured against);
//   1  fused, one moment per matvec: mu_n = <v0, t_n>;
//   2  fused, two moments per matvec: mu_2n = 2<t_n,t_n> - mu_0,
//      mu_2n-1 = 2<t_n,t_n-1> - mu_1.
// precision 1 keeps the recursion vectors and the matrix values in FP32 (the
// dot products stay FP64).  The batched entry point runs many independent
// matrices in one launch per step and captures all steps in one HIP graph.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>

#include "apply.hpp"

namespace vbh {

template <typename S, typename T>
__global__ void k_copy_convert(int64_t rows, int K, const S* __restrict__ in, int64_t ldi, T* __restrict__ out,
                               int64_t ldo) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= rows * K) return;
    const int64_t r = e / K;
    const int v = int(e - r * K);
    const S x = in[r * ldi + v];
    if constexpr (Sc<S>::is_complex) out[r * ldo + v] = Sc<T>::make(x.re, x.im);
    else out[r * ldo + v] = Sc<T>::make(double(x), 0.0);
}

// y = alpha*y + beta*x + gamma*z (the plain path's separate vector pass)
template <typename T>
__global__ void k_axpbypcz(int64_t n, double alpha, T* __restrict__ y, double beta, const T* __restrict__ x,
                           double gamma, const T* __restrict__ z) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    T v = rscale(alpha, y[e]);
    if (beta != 0.0) v = v + rscale(beta, x[e]);
    if (gamma != 0.0) v = v + rscale(gamma, z[e]);
    y[e] = v;
}

// partial[b][v] = Re sum_rows conj(w) y over rows [b*rp, (b+1)*rp) (plain path bdot)
template <typename T>
__global__ __launch_bounds__(256) void k_redot_partials(int64_t rows, int K, const T* __restrict__ w,
                                                        const T* __restrict__ y, int64_t rp,
                                                        double* __restrict__ partial, int E, int col0) {
    __shared__ double red[256];
    const int t = threadIdx.x;
    const int Kp = pow2ceil(K);
    const int v = t & (Kp - 1);
    const int64_t rb = int64_t(blockIdx.x) * rp;
    int64_t re_ = rb + rp;
    if (re_ > rows) re_ = rows;
    double s = 0.0;
    if (v < K)
        for (int64_t r = rb + t / Kp; r < re_; r += 256 / Kp) s += redot(w[r * K + v], y[r * K + v]);
    red[t] = s;
    __syncthreads();
    for (int m = 128; m >= Kp; m >>= 1) {
        if (t < m) red[t] += red[t + m];
        __syncthreads();
    }
    if (t < K) partial[int64_t(blockIdx.x) * E + col0 + t] = red[t];
}

// Sum partials per (step, segment): out[(s*nseg + h)*E + e] =
// sum_{g in seg h} partials[s][g][e]; grid (E, nseg, nsteps).
// Segment h also owns [off2 + seg2[h], off2 + seg2[h+1]) when seg2 is given
// (the matrix-core groups of a batch, stored after the lane groups).
__global__ void k_reduce_segments(const double* __restrict__ partials, int Gtot, int E, const int* __restrict__ seg,
                                  int nseg, double* __restrict__ out, const int* __restrict__ seg2, int off2) {
    const int e = blockIdx.x, h = blockIdx.y, s = blockIdx.z;
    const double* base = partials + int64_t(s) * Gtot * E;
    double acc = 0.0;
    for (int g = seg[h] + threadIdx.x; g < seg[h + 1]; g += blockDim.x) acc += base[int64_t(g) * E + e];
    if (seg2)
        for (int g = off2 + seg2[h] + threadIdx.x; g < off2 + seg2[h + 1]; g += blockDim.x)
            acc += base[int64_t(g) * E + e];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int m = 128; m >= 1; m >>= 1) {
        if (threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(int64_t(s) * nseg + h) * E + e] = red[0];
}

// ------------------------------------------------------------ batched step
// One launch per recursion step over every matrix of the batch: work-group g
// (XCD-remapped) finds its matrix by binary search over g_base, then runs the
// shared row loop on its chunk.  step >= 0 produces t_{step+1}.
template <typename T, int RB>
__global__ __launch_bounds__(kThreads) void k_apply_batched(const BatchItem<T>* __restrict__ items, int nitems,
                                                            int Gtot, int K, int step, double alpha, double beta,
                                                            double gamma0, double gamma, int mode,
                                                            double* __restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int gg = xcd_chunk(Gtot);
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].g_base <= gg) lo = mid;
        else hi = mid - 1;
    }
    const BatchItem<T>& it = items[lo];
    const int chunk = gg - it.g_base;
    ApplyArgs<T> a{};
    a.p = it.p;
    a.vals = it.vals;
    a.K = K;
    a.X = it.buf[step % 3];
    a.ldx = K;
    a.Y = it.buf[(step + 1) % 3];
    a.ldy = K;
    a.alpha = alpha;
    a.beta = beta;
    a.gamma = step == 0 ? gamma0 : gamma;
    a.Z = it.buf[(step + 2) % 3];
    a.ldz = K;
    a.W = mode == 1 ? it.v0 : nullptr;
    a.ldw = K;
    double d1 = 0.0, d2 = 0.0;
    if (chunk < it.G) row_loop<T, 0, RB, true>(a, it.split[chunk], it.split[chunk + 1], lane, wave, d1, d2);
    write_partials(1, lane, wave, d1, d2, partials + int64_t(gg) * 2);
}

// HIP-event timing of the last recursion on this thread: the step kernels
// only (setup, the final reduction and the host copy excluded), recorded on
// the launch stream.  Read through vbh_kpm_last_step_time.
static thread_local double g_last_step_ms = 0.0;
static thread_local int64_t g_last_launches = 0;

struct StepTimer {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipStream_t s;
    explicit StepTimer(hipStream_t st) : s(st) {
        VBH_CHECK_HIP(hipEventCreate(&t0));
        VBH_CHECK_HIP(hipEventCreate(&t1));
    }
    void start() { VBH_CHECK_HIP(hipEventRecord(t0, s)); }
    void stop(int64_t launches) {
        VBH_CHECK_HIP(hipEventRecord(t1, s));
        VBH_CHECK_HIP(hipEventSynchronize(t1));
        float ms = 0.f;
        VBH_CHECK_HIP(hipEventElapsedTime(&ms, t0, t1));
        g_last_step_ms = ms;
        g_last_launches = launches;
    }
    ~StepTimer() {
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
    }
};

template <typename T>
static int64_t step_count(int M, int mode) {
    if (M <= 1) return 0;
    if (mode == 2) return M / 2;  // steps producing t_1 .. t_ceil((M-1)/2)
    return M - 1;
}

// Bytes of one recursion buffer, rounded to 256 so that every buffer of a
// set starts 256-byte aligned (vector loads in the step kernels).
template <typename T>
static size_t vec_bytes_of(int64_t N, int K) {
    return (sizeof(T) * size_t(std::max<int64_t>(N * K, 1)) + 255) & ~size_t(255);
}

// Moments from the reduced per-step dots of one matrix, on the host (the
// batched driver): the rule of k_assemble_moments below, d[s][0..2K) per step,
// K partial columns.
static void assemble_moments(int M, int mode, int K, double mu0, const double* d, double* mu) {
    for (int n = 0; n < M; ++n) mu[n] = 0.0;
    mu[0] = mu0;
    if (M == 1) return;
    auto sum = [&](int s, int which) {
        double t = 0.0;
        for (int v = 0; v < K; ++v) t += d[int64_t(s) * 2 * K + which * K + v];
        return t;
    };
    if (mode != 2) {
        for (int n = 1; n < M; ++n) mu[n] = sum(n - 1, 0);  // step n-1 produced t_n; d1 = <v0, t_n>
        return;
    }
    // step 0 produced t_1: d1 = <t1,t1>, d2 = <t0,t1>
    mu[1] = sum(0, 1);
    for (int s = 0; s < step_count<double>(M, 2); ++s) {
        const int n = s + 1;  // t_n produced at step s
        const int even = 2 * n, odd = 2 * n - 1;
        if (even < M) mu[even] = 2.0 * sum(s, 0) - mu[0];
        if (odd > 1 && odd < M) mu[odd] = 2.0 * sum(s, 1) - mu[1];
    }
}

__global__ void k_zero_f64(double* __restrict__ p, int64_t n) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// Sum the per-work-group partials of every step: out[s*E + e] =
// sum_{g < count} partials[(s*stride + g)*E + e]; grid (E, nsteps).
__global__ void k_reduce_steps(const double* __restrict__ partials, int64_t stride, int count, int E,
                               double* __restrict__ out) {
    const int e = blockIdx.x;
    const int64_t s = blockIdx.y;
    const double* base = partials + s * stride * E;
    double acc = 0.0;
    for (int g = threadIdx.x; g < count; g += blockDim.x) acc += base[int64_t(g) * E + e];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int m = 128; m >= 1; m >>= 1) {
        if (threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s * E + e] = red[0];
}

// Moments from the reduced dots (one thread): d[s][0..Kd) = d1 and
// [Kd..2Kd) = d2 of step s (Kd = 1 when the step kernel summed the columns);
// d0[0..K) = <t0, t0> per column.  Mode 0/1: mu_n = <v0, t_n> (step n-1
// produced t_n).  Mode 2 (two moments per matvec, kpm.py:63-66): step 0 gives
// mu_1 = <t0, t1>; step s >= 0 gives t_n (n = s + 1) and mu_2n = 2<t_n, t_n> -
// mu_0, mu_2n-1 = 2<t_n, t_n-1> - mu_1.
__global__ void k_assemble_moments(int M, int mode, int Kd, int K, int64_t nsteps, const double* __restrict__ d,
                                   const double* __restrict__ d0, double* __restrict__ mu) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double mu0 = 0.0;
    for (int v = 0; v < K; ++v) mu0 += d0[v];
    for (int n = 0; n < M; ++n) mu[n] = 0.0;
    mu[0] = mu0;
    if (M == 1 || nsteps == 0) return;
    auto sum = [&](int64_t st, int which) {
        double t = 0.0;
        for (int v = 0; v < Kd; ++v) t += d[st * 2 * Kd + which * Kd + v];
        return t;
    };
    if (mode != 2) {
        for (int n = 1; n < M; ++n) mu[n] = sum(n - 1, 0);
        return;
    }
    mu[1] = sum(0, 1);
    for (int64_t st = 0; st < nsteps; ++st) {
        const int64_t n = st + 1;
        const int64_t even = 2 * n, odd = 2 * n - 1;
        if (even < M) mu[even] = 2.0 * sum(st, 0) - mu0;
        if (odd > 1 && odd < M) mu[odd] = 2.0 * sum(st, 1) - mu[1];
    }
}

// Device buffers of one single-matrix KPM call, laid out from one base
// address (every region 256-byte aligned): the recursion vectors t_{n-1},
// t_n, t_{n+1} (and v0 in modes 0/1), the per-step per-work-group partial dots,
// the partials of mu_0 = <t0, t0>, the reduced dots and the moments.
struct KpmLayout {
    int64_t N = 0, nsteps = 0, rp = 0, rp0 = 0;
    size_t vec_bytes = 0;
    bool keep_v0 = false;
    int Kd = 1, E = 2, G = 0, ndot = 0, ndot0 = 0;
    size_t off_part = 0, off_part0 = 0, off_red = 0, off_red0 = 0, off_mu = 0, total = 0;
};

static size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

// Sizes and work-group counts of a call.  Reads the matrix's row-class lists
// (built on first use).
template <typename T, typename V>
static KpmLayout kpm_layout(vbh_matrix* m, const V* vals, int K, int M, int mode) {
    KpmLayout L;
    L.N = m->n_rows;
    L.nsteps = L.N > 0 ? step_count<T>(M, mode) : 0;
    L.vec_bytes = vec_bytes_of<T>(L.N, K);
    L.keep_v0 = mode != 2;  // modes 0/1 dot every t_n against v0
    // fused steps sum their dots over the columns (E = 2); the plain path's
    // dot kernel keeps one partial per column (E = 2K)
    L.Kd = mode == 0 ? K : 1;
    L.E = 2 * L.Kd;
    L.G = m->G;
    if (L.nsteps > 0 && mode != 0) {
        auto probe = make_args<T, V>(m, vals, K, (const T*)nullptr, K, (T*)nullptr, K);
        probe.dot_cols = 1;
        L.G = apply_groups(m, probe);
    }
    L.rp = std::max<int64_t>(256, (L.N + L.G - 1) / std::max(L.G, 1));
    L.ndot = int((L.N + L.rp - 1) / L.rp);
    L.rp0 = std::max<int64_t>(256, (L.N + 1023) / 1024);
    L.ndot0 = int((L.N + L.rp0 - 1) / L.rp0);
    L.off_part = (L.keep_v0 ? 4 : 3) * L.vec_bytes;
    L.off_part0 = L.off_part + round256(sizeof(double) * size_t(L.nsteps) * L.G * L.E);
    L.off_red = L.off_part0 + round256(sizeof(double) * size_t(std::max(L.ndot0, 1)) * K);
    L.off_red0 = L.off_red + round256(sizeof(double) * size_t(L.nsteps) * L.E);
    L.off_mu = L.off_red0 + round256(sizeof(double) * K);
    L.total = L.off_mu + round256(sizeof(double) * size_t(M));
    return L;
}

// One single-matrix KPM call on the buffers at base (L.total bytes of device
// memory): moments (host, M doubles) of H~ = (H - center)/half.
template <typename T, typename S, typename V>
static void kpm_single(vbh_matrix* m, const V* vals, int K, const S* V0, int ldv, int M, int mode, double center,
                       double half, double* moments, hipStream_t s, const KpmLayout& L, char* base, bool timed) {
    const int64_t N = L.N, nsteps = L.nsteps;
    const size_t vb = L.vec_bytes;
    T* t[3] = {(T*)base, (T*)(base + vb), (T*)(base + 2 * vb)};
    T* v0 = L.keep_v0 ? (T*)(base + 3 * vb) : t[0];
    double* partials = (double*)(base + L.off_part);
    double* part0 = (double*)(base + L.off_part0);
    double* red = (double*)(base + L.off_red);
    double* red0 = (double*)(base + L.off_red0);
    double* mu = (double*)(base + L.off_mu);
    const int G = L.G, E = L.E;
    const unsigned nb = unsigned((N * K + 255) / 256);
    if (N * K > 0) {
        hipLaunchKernelGGL((k_copy_convert<S, T>), dim3(nb), dim3(256), 0, s, N, K, V0, int64_t(ldv), t[0], int64_t(K));
        VBH_CHECK_HIP(hipGetLastError());
        if (L.keep_v0) VBH_CHECK_HIP(hipMemcpyAsync(v0, t[0], sizeof(T) * N * K, hipMemcpyDeviceToDevice, s));
        // mu_0 = sum_v <t0, t0>
        hipLaunchKernelGGL(k_redot_partials<T>, dim3(L.ndot0), dim3(256), 0, s, N, K, (const T*)t[0],
                           (const T*)t[0], L.rp0, part0, K, 0);
        VBH_CHECK_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_reduce_steps, dim3(K, 1), dim3(256), 0, s, part0, int64_t(0), L.ndot0, K, red0);
        VBH_CHECK_HIP(hipGetLastError());
    } else {
        hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, red0, int64_t(K));
        VBH_CHECK_HIP(hipGetLastError());
    }
    if (nsteps > 0) {
        const double a1 = 1.0 / half, b1 = -center / half;        // t1 = H~ t0
        const double a2 = 2.0 / half, b2 = -2.0 * center / half;  // t_{n+1} = 2 H~ t_n - t_{n-1}
        std::unique_ptr<StepTimer> timer;
        if (timed) {
            timer.reset(new StepTimer(s));
            timer->start();
        }
        if (mode == 0) {
            // plain: SpMM kernel, then a vector pass, then a dot kernel per step
            VBH_REQUIRE(L.ndot <= G, "internal: dot partial blocks exceed work-groups");
            for (int64_t st = 0; st < nsteps; ++st) {
                T* X = t[st % 3];
                T* Y = t[(st + 1) % 3];
                T* Z = t[(st + 2) % 3];
                apply_plain<T, V>(m, vals, K, X, K, Y, K, s);
                hipLaunchKernelGGL(k_axpbypcz<T>, dim3(nb), dim3(256), 0, s, N * K, st == 0 ? a1 : a2, Y,
                                   st == 0 ? b1 : b2, (const T*)X, st == 0 ? 0.0 : -1.0, (const T*)Z);
                VBH_CHECK_HIP(hipGetLastError());
                hipLaunchKernelGGL(k_redot_partials<T>, dim3(L.ndot), dim3(256), 0, s, N, K, (const T*)v0,
                                   (const T*)Y, L.rp, partials + st * G * E, E, 0);
                VBH_CHECK_HIP(hipGetLastError());
            }
        } else {
            for (int64_t st = 0; st < nsteps; ++st) {
                auto a = make_args<T, V>(m, vals, K, t[st % 3], K, t[(st + 1) % 3], K);
                a.alpha = st == 0 ? a1 : a2;
                a.beta = st == 0 ? b1 : b2;
                a.gamma = st == 0 ? 0.0 : -1.0;
                a.Z = t[(st + 2) % 3];
                a.ldz = K;
                a.W = mode == 1 ? v0 : nullptr;
                a.ldw = K;
                a.dot_cols = 1;
                a.partials = partials + st * G * E;
                launch_rows<T, true, V>(m, a, s);
            }
        }
        if (timer) timer->stop(mode == 0 ? 3 * nsteps : nsteps);
        // one reduction over all steps (the plain path's dot kernel fills
        // the first ndot partial rows of a step, the fused kernels all G)
        hipLaunchKernelGGL(k_reduce_steps, dim3(E, unsigned(nsteps)), dim3(256), 0, s, partials, int64_t(G),
                           mode == 0 ? L.ndot : G, E, red);
        VBH_CHECK_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_assemble_moments, dim3(1), dim3(64), 0, s, M, mode, L.Kd, K, nsteps, (const double*)red,
                       (const double*)red0, mu);
    VBH_CHECK_HIP(hipGetLastError());
    VBH_CHECK_HIP(hipMemcpyAsync(moments, mu, sizeof(double) * M, hipMemcpyDeviceToHost, s));
    VBH_CHECK_HIP(hipStreamSynchronize(s));
}

// f(T*, S*, V*): recursion vector type T, start-vector (input) type S, matrix
// value type V, for precision 0 (native), 1 (FP32) or 2 (FP32 values).
template <typename F>
static void dispatch_kpm(const vbh_matrix* m, int precision, F&& f) {
    VBH_REQUIRE(precision >= 0 && precision <= 2, "precision must be 0, 1 or 2");
    if (m->dtype == VBH_F64) {
        if (precision == 0) f((double*)nullptr, (double*)nullptr, (double*)nullptr);
        else if (precision == 1) f((float*)nullptr, (double*)nullptr, (float*)nullptr);
        else f((double*)nullptr, (double*)nullptr, (float*)nullptr);
    } else {
        if (precision == 0) f((c128*)nullptr, (c128*)nullptr, (c128*)nullptr);
        else if (precision == 1) f((c64*)nullptr, (c128*)nullptr, (c64*)nullptr);
        else f((c128*)nullptr, (c128*)nullptr, (c64*)nullptr);
    }
}

// ------------------------------------------------------------ batched driver
// Work layout of a batch: lane groups per matrix (over the small rows when the
// matrix has matrix-core rows, real types), matrix-core groups per matrix,
// each launch's total rounded up to a multiple of 8 (trailing groups find no
// rows); K == 1: row loop over each matrix's work-balanced split.
struct BatchLayout {
    std::vector<int> seg, seg2;
    int Gtot1 = 0, Gtot2 = 0;
};
static BatchLayout batch_layout(int count, vbh_matrix* const* ms, const vbh_matrix::RowLists* const* lists,
                                const LaneShape& sh, bool mfma_ok) {
    BatchLayout L;
    L.seg.assign(count + 1, 0);
    L.seg2.assign(count + 1, 0);
    for (int h = 0; h < count; ++h) {
        const int64_t rpg = sh.rw ? sh.rw : sh.S ? lane_rows_per_group(sh.S, sh.P) : 1;
        const bool split = mfma_ok && lists[h]->n_big > 0;
        const int64_t lrows = split ? lists[h]->n_small : ms[h]->n_rows;
        L.seg[h + 1] = L.seg[h] + (sh.S ? int((lrows + rpg - 1) / rpg) : ms[h]->G);
        L.seg2[h + 1] = L.seg2[h] + (split ? (lists[h]->n_big + kMfmaRowsPerGroup - 1) / kMfmaRowsPerGroup : 0);
    }
    L.Gtot1 = (L.seg[count] + 7) & ~7;
    L.Gtot2 = (L.seg2[count] + 7) & ~7;
    return L;
}

// Tile plan of a batch: the class threshold (block rows of t..16 rows in the
// class, t = 17: none), the kernel of the class rows (the FP64 matrix cores,
// or with brw the block-row wave kernel on the vector ALUs) and the kernel of
// the other rows: the lane family with P slot partitions, or the row-wave
// kernel with rw rows per work-group.  With autotuning on, the first call on a batch times one
// recursion step of every candidate and keeps the fastest for later calls on
// the same batch.
struct BatchPlan {
    int t = kMfmaMinRows;
    int P = 1;
    bool tuned = false;
    int rw = 0;
    bool brw = false;
};
// Plans per batch (process-wide, guarded by g_plan_mu), the switches, and the
// plan of this thread's last batched call.
static std::mutex g_plan_mu;
static std::map<std::vector<uintptr_t>, BatchPlan> g_batch_plans;
static int g_autotune = 1;
static BatchPlan g_forced_plan{0, 1, false, 0, false};  // t > 0: used instead of the cached or tuned plan
static thread_local BatchPlan g_last_plan;

template <typename T, typename S, typename V>
static void kpm_batched(int count, vbh_matrix* const* ms, int K, const void* const* V0s, const int* ldv, int M, int mode,
                        double center, double half, int use_graph, double* moments, hipStream_t s) {
    VBH_REQUIRE(mode == 1 || mode == 2, "batched KPM supports the fused modes 1 and 2");
    VBH_REQUIRE(K >= 1 && K <= 64, "batched KPM supports 1 <= K <= 64");
    using Item = BatchItem<T, V>;
    const int64_t nsteps = step_count<T>(M, mode);
    std::vector<Item> items(count);
    std::vector<size_t> off(count + 1, 0);
    for (int h = 0; h < count; ++h) {
        VBH_REQUIRE(ms[h]->nbc == ms[h]->nbr, "batched KPM needs square local patterns without ghosts");
        off[h + 1] = off[h] + vec_bytes_of<T>(ms[h]->n_rows, K) * (mode == 1 ? 4 : 3);
    }
    // Dots are summed over the columns (E = 2) everywhere.
    LaneShape sh = lane_shape_of<T, V>(K);
    bool mfma_ok = !Sc<T>::is_complex && sh.S > 0;
    for (int h = 0; h < count; ++h) mfma_ok = mfma_ok && ms[h]->max_block <= 16;
    // the row-wave kernel: real types, K = CPL * 2^k, block widths below 256
    // and fewer than 2^24 columns (packed into its LDS entries)
    bool rowwave_ok = sh.S > 0 && rowwave_cpl<T>(K) > 0;
    for (int h = 0; h < count; ++h) rowwave_ok = rowwave_ok && ms[h]->n_cols < (int64_t(1) << 24);
    // the block-row wave kernel for the class rows: real vectors, FP32 values,
    // K = 16, 32 or 64, leading dimension K (the recursion buffers)
    const bool brw_ok = mfma_ok && std::is_same_v<V, float> && brw_cpl<T>(K) > 0;
    // row-class lists of every matrix for the plan's matrix-core threshold
    std::vector<const vbh_matrix::RowLists*> lists(count);
    auto select_all = [&](int t) {
        for (int h = 0; h < count; ++h) lists[h] = &row_classes(ms[h], t, s);
    };
    auto default_parts = [&]() {
        int64_t lrows = 0;
        for (int h = 0; h < count; ++h)
            lrows += (mfma_ok && lists[h]->n_big > 0) ? lists[h]->n_small : ms[h]->n_rows;
        return lane_parts(lrows, sh.S, 0, int(sizeof(T)));
    };
    std::vector<uintptr_t> pkey = {uintptr_t(count), uintptr_t(K), uintptr_t(sizeof(T)), uintptr_t(sizeof(V)),
                                   uintptr_t(mode)};
    for (int h = 0; h < count; ++h) {
        pkey.push_back(uintptr_t(ms[h]));
        pkey.push_back(uintptr_t(ms[h]->nnzb));
    }
    BatchPlan plan, forced_plan;
    bool have_known = false;
    int autotune = 1;
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        auto known = g_batch_plans.find(pkey);
        if (known != g_batch_plans.end()) {
            have_known = true;
            plan = known->second;
        }
        forced_plan = g_forced_plan;
        autotune = g_autotune;
    }
    const bool forced = forced_plan.t > 0 && sh.S > 0;
    const bool tune = !forced && sh.S > 0 && autotune && nsteps >= 2 && !have_known;
    if (forced) {
        plan = forced_plan;
        plan.P = std::min(plan.P, 64 / sh.S);
        if (!rowwave_ok) plan.rw = 0;
        if (!brw_ok) plan.brw = false;
    } else if (have_known) {
    } else if (sh.S) {
        select_all(kMfmaMinRows);
        plan.P = default_parts();
    }
    if (brw_ok && (tune || plan.brw))
        for (int h = 0; h < count; ++h) ensure_brw(ms[h], s);
    select_all(plan.t);
    sh.P = sh.S ? plan.P : 1;
    sh.rw = sh.S ? plan.rw : 0;
    sh.brw = sh.S && plan.brw ? 1 : 0;
    BatchLayout lay = batch_layout(count, ms, lists.data(), sh, mfma_ok);
    std::vector<int>& seg = lay.seg;
    std::vector<int>& seg2 = lay.seg2;
    int Gtot1 = lay.Gtot1, Gtot2 = lay.Gtot2;
    int Gtot = Gtot1 + Gtot2;
    const int E = 2;
    char* vecs = (char*)scratch(std::max<size_t>(off[count], 64), 4);
    // plan-dependent fields of the item table: the active row lists and the
    // groups of every matrix in the layout
    auto fill_items = [&](const BatchLayout& L) {
        for (int h = 0; h < count; ++h) {
            vbh_matrix* m = ms[h];
            Item& it = items[h];
            it.p = m->pattern(*lists[h]);
            it.p.bs = 0;
            it.split = m->split;
            it.n_rows = m->n_rows;
            it.row_list = (mfma_ok && lists[h]->n_big > 0) ? 1 : 0;
            it.lane_rows = it.row_list ? lists[h]->n_small : m->n_rows;
            it.G = L.seg[h + 1] - L.seg[h];
            it.g_base = L.seg[h];
            it.g_base2 = L.seg2[h];
            it.vt = m->vt;
            it.eoff = m->eoff;
            it.ecol = m->ecol;
        }
    };
    std::vector<double> mu0(count, 0.0);
    for (int h = 0; h < count; ++h) {
        vbh_matrix* m = ms[h];
        const int64_t N = m->n_rows;
        const size_t vb = vec_bytes_of<T>(N, K);
        Item& it = items[h];
        it.vals = compute_values<V>(m, s);
        for (int q = 0; q < 3; ++q) it.buf[q] = (T*)(vecs + off[h] + q * vb);
        it.v0 = mode == 1 ? (const T*)(vecs + off[h] + 3 * vb) : nullptr;
        if (N * K > 0) {
            hipLaunchKernelGGL((k_copy_convert<S, T>), dim3(unsigned((N * K + 255) / 256)), dim3(256), 0, s, N, K,
                               (const S*)V0s[h], int64_t(ldv[h]), it.buf[0], int64_t(K));
            VBH_CHECK_HIP(hipGetLastError());
            if (mode == 1)
                VBH_CHECK_HIP(hipMemcpyAsync((void*)it.v0, it.buf[0], vb, hipMemcpyDeviceToDevice, s));
            std::vector<double> dd(2 * K);
            const int dt = std::is_same_v<T, double> ? VBH_F64 : std::is_same_v<T, float> ? VBH_F32
                           : std::is_same_v<T, c128> ? VBH_C128 : VBH_C64;
            if (vbh_vec_bdot(dt, N, K, it.buf[0], K, it.buf[0], K, dd.data(), s) != VBH_OK)
                throw Error(VBH_ERR_HIP, vbh_last_error());
            for (int v = 0; v < K; ++v) mu0[h] += dd[2 * v];
        }
    }
    const double a1 = 1.0 / half, b1 = -center / half;
    const double a2 = 2.0 / half, b2 = -2.0 * center / half;
    StreamState& sst = current_stream_state();
    StreamFork& fork = sst.fork;
    if (sh.S && mfma_ok) fork.init();
    // One recursion step of the lane family (and the forked matrix-core
    // launch) for the items at ditems in layout L.
    auto lane_step = [&](const Item* ditems, const BatchLayout& L, const LaneShape& shp, int64_t step, double* pt,
                         hipStream_t st) {
        const double al = step == 0 ? a1 : a2, be = step == 0 ? b1 : b2;
        if (L.Gtot2 > 0) {
            // the matrix-core launch of a step is independent of its lane
            // launch: forked onto the side stream and joined back, so the two
            // run side by side
            VBH_CHECK_HIP(hipEventRecord(fork.fork, st));
            VBH_CHECK_HIP(hipStreamWaitEvent(fork.side, fork.fork, 0));
            if (shp.brw)
                launch_batched_brw<T, V>(ditems, count, L.Gtot2, K, K, int(step), al, be, 0.0, -1.0, mode,
                                         pt + int64_t(L.Gtot1) * E, fork.side);
            else
                launch_batched_mfma<T, V>(ditems, count, L.Gtot2, K, K, int(step), al, be, 0.0, -1.0, mode,
                                          pt + int64_t(L.Gtot1) * E, fork.side);
        }
        if (L.Gtot1 > 0) {
            if (shp.rw)
                launch_batched_rowwave<T, V>(ditems, count, L.Gtot1, K, K, shp.rw, int(step), al, be, 0.0, -1.0, mode,
                                             pt, st);
            else
                launch_batched_lanes<T, V>(ditems, count, L.Gtot1, K, K, shp, int(step), al, be, 0.0, -1.0, mode, pt,
                                           st);
        }
        if (L.Gtot2 > 0) {
            VBH_CHECK_HIP(hipEventRecord(fork.join, fork.side));
            VBH_CHECK_HIP(hipStreamWaitEvent(st, fork.join, 0));
        }
    };
    if (tune) {
        // Candidates: class thresholds (4: the 4..16-row classes, 8: 8..16
        // only, 17: none) with the class rows on the matrix cores, or (encoded
        // as -1 (every row), -4, -8) on the block-row wave kernel, times slot
        // partitions around the default.  Each is timed on recursion step 1 (reads buffers 0, 1,
        // writes buffer 2; buffer 1 is cleared first), which the real run
        // overwrites; its partials go to a scratch of their own.
        for (int h = 0; h < count; ++h) {
            const size_t vb = vec_bytes_of<T>(ms[h]->n_rows, K);
            VBH_CHECK_HIP(hipMemsetAsync(items[h].buf[1], 0, vb, s));
        }
        std::vector<int> ts = mfma_ok ? std::vector<int>{4, 8, 17} : std::vector<int>{plan.t};
        if (brw_ok)
            for (int t : {-1, -4, -8}) ts.push_back(t);
        const int pmax = 64 / sh.S < 16 ? 64 / sh.S : 16;
        float best = 0.f;
        hipEvent_t e0, e1;
        VBH_CHECK_HIP(hipEventCreate(&e0));
        VBH_CHECK_HIP(hipEventCreate(&e1));
        for (int ts_ : ts) {
            const int t = ts_ < 0 ? -ts_ : ts_;
            select_all(t);
            const int p0 = default_parts();
            // lane family: P around the row-count rule (rw = 0); row-wave
            // kernel: rw rows per work-group (encoded as P = -rw)
            std::vector<int> cands = {p0 / 2, p0, 2 * p0, 4 * p0};
            if (rowwave_ok)
                for (int rw : {8, 16, 32}) cands.push_back(-rw);
            for (int P : cands) {
                if (P == 0 || P > pmax || P < -64) continue;
                LaneShape shc = sh;
                shc.P = P > 0 ? P : 1;
                shc.rw = P < 0 ? -P : 0;
                shc.brw = ts_ < 0 ? 1 : 0;
                BatchLayout Lc = batch_layout(count, ms, lists.data(), shc, mfma_ok);
                fill_items(Lc);
                Item* titems = (Item*)scratch(sizeof(Item) * count, 8);
                double* tpart = (double*)scratch(sizeof(double) * std::max(Lc.Gtot1 + Lc.Gtot2, 8) * E, 9);
                VBH_CHECK_HIP(hipMemcpyAsync(titems, items.data(), sizeof(Item) * count, hipMemcpyHostToDevice, s));
                lane_step(titems, Lc, shc, 1, tpart, s);  // warm-up
                VBH_CHECK_HIP(hipEventRecord(e0, s));
                for (int rep = 0; rep < 3; ++rep) lane_step(titems, Lc, shc, 1, tpart, s);
                VBH_CHECK_HIP(hipEventRecord(e1, s));
                VBH_CHECK_HIP(hipEventSynchronize(e1));
                float ms_ = 0.f;
                VBH_CHECK_HIP(hipEventElapsedTime(&ms_, e0, e1));
                if (!plan.tuned || ms_ < best) {
                    best = ms_;
                    plan.t = t;
                    plan.P = shc.P;
                    plan.rw = shc.rw;
                    plan.brw = shc.brw != 0;
                    plan.tuned = true;
                }
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        {
            std::lock_guard<std::mutex> lk(g_plan_mu);
            g_batch_plans[pkey] = plan;
        }
        select_all(plan.t);
        sh.P = plan.P;
        sh.rw = plan.rw;
        sh.brw = plan.brw ? 1 : 0;
        lay = batch_layout(count, ms, lists.data(), sh, mfma_ok);
        Gtot1 = lay.Gtot1;
        Gtot2 = lay.Gtot2;
        Gtot = Gtot1 + Gtot2;
    } else if (!forced && !have_known) {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        g_batch_plans[pkey] = plan;
    }
    g_last_plan = plan;
    fill_items(lay);
    double* partials = (double*)scratch(sizeof(double) * std::max<int64_t>(nsteps, 1) * std::max(Gtot, 8) * E, 5);
    if (nsteps > 0 && Gtot > 0) {
        Item* ditems = (Item*)scratch(sizeof(Item) * count + 2 * sizeof(int) * (count + 1), 6);
        int* dseg = (int*)(ditems + count);
        int* dseg2 = dseg + (count + 1);
        VBH_CHECK_HIP(hipMemcpyAsync(ditems, items.data(), sizeof(Item) * count, hipMemcpyHostToDevice, s));
        VBH_CHECK_HIP(hipMemcpyAsync(dseg, seg.data(), sizeof(int) * (count + 1), hipMemcpyHostToDevice, s));
        VBH_CHECK_HIP(hipMemcpyAsync(dseg2, seg2.data(), sizeof(int) * (count + 1), hipMemcpyHostToDevice, s));
        VBH_CHECK_HIP(hipStreamSynchronize(s));
        int max_block = 1;
        for (int h = 0; h < count; ++h) max_block = std::max(max_block, ms[h]->max_block);
        const int rb = row_config(K < 2 ? 2 : K, max_block, 0).rb;
        auto launch_all = [&](hipStream_t st) {
            for (int64_t step = 0; step < nsteps; ++step) {
                const Item* it = ditems;
                const double al = step == 0 ? a1 : a2, be = step == 0 ? b1 : b2;
                double* pt = partials + step * Gtot * E;
                if (sh.S) {
                    lane_step(ditems, lay, sh, step, pt, st);
                    continue;
                }
                if constexpr (!is_mixed<T, V>()) {  // K == 1, single type: row loop
                    switch (rb) {
                        case 1: hipLaunchKernelGGL((k_apply_batched<T, 1>), dim3(Gtot), dim3(kThreads), 0, st, it,
                                                   count, Gtot, K, int(step), al, be, 0.0, -1.0, mode, pt); break;
                        case 2: hipLaunchKernelGGL((k_apply_batched<T, 2>), dim3(Gtot), dim3(kThreads), 0, st, it,
                                                   count, Gtot, K, int(step), al, be, 0.0, -1.0, mode, pt); break;
                        case 4: hipLaunchKernelGGL((k_apply_batched<T, 4>), dim3(Gtot), dim3(kThreads), 0, st, it,
                                                   count, Gtot, K, int(step), al, be, 0.0, -1.0, mode, pt); break;
                        default: hipLaunchKernelGGL((k_apply_batched<T, 8>), dim3(Gtot), dim3(kThreads), 0, st, it,
                                                    count, Gtot, K, int(step), al, be, 0.0, -1.0, mode, pt); break;
                    }
                }
            }
        };
        if (use_graph) {
            // The captured graph is kept and replayed while the batch (matrices,
            // buffers, shape and scaling) is unchanged; a new batch re-captures.
            // one cached graph per stream state
            GraphCache& cache = sst.graph;
            std::vector<uintptr_t> key = {uintptr_t(ditems), uintptr_t(vecs), uintptr_t(partials), uintptr_t(count),
                                          uintptr_t(K), uintptr_t(nsteps), uintptr_t(mode), uintptr_t(sizeof(T)),
                                          uintptr_t(Gtot), uintptr_t(Gtot1), uintptr_t(sh.P), uintptr_t(sh.rw),
                                          uintptr_t(sh.brw)};
            double ab[4] = {a1, b1, a2, b2};
            for (double v : ab) {
                uint64_t bits;
                std::memcpy(&bits, &v, sizeof(bits));
                key.push_back(uintptr_t(bits));
            }
            for (int h = 0; h < count; ++h) {
                key.push_back(uintptr_t(items[h].vals));
                key.push_back(uintptr_t(ms[h]->split));
            }
            if (!cache.cs) VBH_CHECK_HIP(hipStreamCreateWithFlags(&cache.cs, hipStreamNonBlocking));
            if (!cache.exec || cache.key != key) {
                if (cache.exec) VBH_CHECK_HIP(hipGraphExecDestroy(cache.exec));
                if (cache.graph) VBH_CHECK_HIP(hipGraphDestroy(cache.graph));
                cache.exec = nullptr;
                cache.graph = nullptr;
                VBH_CHECK_HIP(hipStreamBeginCapture(cache.cs, hipStreamCaptureModeThreadLocal));
                launch_all(cache.cs);
                VBH_CHECK_HIP(hipStreamEndCapture(cache.cs, &cache.graph));
                VBH_CHECK_HIP(hipGraphInstantiate(&cache.exec, cache.graph, nullptr, nullptr, 0));
                cache.key = key;
            }
            StepTimer timer(cache.cs);
            timer.start();
            VBH_CHECK_HIP(hipGraphLaunch(cache.exec, cache.cs));
            timer.stop(nsteps);
        } else {
            StepTimer timer(s);
            timer.start();
            launch_all(s);
            VBH_CHECK_HIP(hipGetLastError());
            timer.stop(nsteps);
        }
        double* red = (double*)scratch(sizeof(double) * nsteps * count * E, 7);
        hipLaunchKernelGGL(k_reduce_segments, dim3(E, count, unsigned(nsteps)), dim3(256), 0, s, partials, Gtot, E,
                           dseg, count, red, Gtot2 > 0 ? (const int*)dseg2 : nullptr, Gtot1);
        VBH_CHECK_HIP(hipGetLastError());
        std::vector<double> d(nsteps * count * E);
        VBH_CHECK_HIP(hipMemcpyAsync(d.data(), red, sizeof(double) * d.size(), hipMemcpyDeviceToHost, s));
        VBH_CHECK_HIP(hipStreamSynchronize(s));
        std::vector<double> dh(nsteps * E);
        for (int h = 0; h < count; ++h) {
            for (int64_t st = 0; st < nsteps; ++st)
                std::memcpy(&dh[st * E], &d[(st * count + h) * E], sizeof(double) * E);
            assemble_moments(M, mode, 1, mu0[h], dh.data(), moments + int64_t(h) * M);
        }
    } else {
        for (int h = 0; h < count; ++h) {
            for (int n = 0; n < M; ++n) moments[int64_t(h) * M + n] = 0.0;
            if (M >= 1) moments[int64_t(h) * M] = mu0[h];
        }
    }
    // (the captured graph holds the plan's lists, which stay allocated until
    // the matrices are destroyed)
}

// ------------------------------------------------------------ Chebyshev apply
template <typename T>
__global__ void k_cheb_accum(int64_t rows, int K, const T* __restrict__ t, T* __restrict__ Y, int64_t ldy, double cre,
                             double cim, int first) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= rows * K) return;
    const int64_t r = e / K;
    const int v = int(e - r * K);
    const T c = Sc<T>::make(cre, cim);
    const T x = t[r * K + v];
    T& y = Y[r * ldy + v];
    y = first ? c * x : Sc<T>::fma(c, x, y);
}
__global__ void k_cheb_accum_d(int64_t rows, int K, const double* __restrict__ t, double* __restrict__ Y, int64_t ldy,
                               double c, int first) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= rows * K) return;
    const int64_t r = e / K;
    const int v = int(e - r * K);
    Y[r * ldy + v] = first ? c * t[r * K + v] : __builtin_fma(c, t[r * K + v], Y[r * ldy + v]);
}
__global__ void k_cheb_accum_f(int64_t rows, int K, const float* __restrict__ t, float* __restrict__ Y, int64_t ldy,
                               double c, int first) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= rows * K) return;
    const int64_t r = e / K;
    const int v = int(e - r * K);
    Y[r * ldy + v] = first ? float(c) * t[r * K + v] : __builtin_fmaf(float(c), t[r * K + v], Y[r * ldy + v]);
}

template <typename T>
static void cheb_accum(int64_t N, int K, const T* t, T* Y, int64_t ldy, const double* coef, int n, hipStream_t s) {
    const unsigned nb = unsigned((N * K + 255) / 256);
    if constexpr (std::is_same_v<T, double>)
        hipLaunchKernelGGL(k_cheb_accum_d, dim3(nb), dim3(256), 0, s, N, K, t, Y, ldy, coef[2 * n], n == 0);
    else if constexpr (std::is_same_v<T, float>)
        hipLaunchKernelGGL(k_cheb_accum_f, dim3(nb), dim3(256), 0, s, N, K, t, Y, ldy, coef[2 * n], n == 0);
    else
        hipLaunchKernelGGL(k_cheb_accum<T>, dim3(nb), dim3(256), 0, s, N, K, t, Y, ldy, coef[2 * n], coef[2 * n + 1],
                           n == 0);
    VBH_CHECK_HIP(hipGetLastError());
}

template <typename T, typename V>
static void cheb_apply(vbh_matrix* m, int K, const T* X, int ldx, T* Y, int ldy, int nterms, const double* coef,
                       double center, double half, hipStream_t s) {
    const int64_t N = m->n_rows;
    if (N * K == 0 || nterms <= 0) return;
    const size_t vb = sizeof(T) * N * K;
    T* buf = (T*)scratch(3 * vb, 4);
    T* t[3] = {buf, (T*)((char*)buf + vb), (T*)((char*)buf + 2 * vb)};
    const unsigned nb = unsigned((N * K + 255) / 256);
    hipLaunchKernelGGL((k_copy_convert<T, T>), dim3(nb), dim3(256), 0, s, N, K, X, int64_t(ldx), t[0], int64_t(K));
    VBH_CHECK_HIP(hipGetLastError());
    cheb_accum<T>(N, K, t[0], Y, ldy, coef, 0, s);
    const V* vals = compute_values<V>(m, s);
    for (int n = 1; n < nterms; ++n) {
        const int st = n - 1;
        auto a = make_args<T, V>(m, vals, K, t[st % 3], K, t[(st + 1) % 3], K);
        a.alpha = st == 0 ? 1.0 / half : 2.0 / half;
        a.beta = st == 0 ? -center / half : -2.0 * center / half;
        a.gamma = st == 0 ? 0.0 : -1.0;
        a.Z = t[(st + 2) % 3];
        a.ldz = K;
        // Y += c_n t_n in the same launch where the SpMV family runs the step
        const SeriesAcc<T> sa{Y, int64_t(ldy), coef[2 * n], coef[2 * n + 1]};
        if (!launch_spmv_series<T, V>(m, a, sa, s)) {
            launch_rows<T, true, V>(m, a, s);
            cheb_accum<T>(N, K, t[(st + 1) % 3], Y, ldy, coef, n, s);
        }
    }
}

}  // namespace vbh

using namespace vbh;

static void check_kpm_args(const vbh_matrix* m, int K, const void* V0, int ldv, int M, int mode, double half_width,
                           const double* moments) {
    VBH_REQUIRE(m && V0 && moments, "null argument");
    VBH_REQUIRE(m->nbc == m->nbr, "KPM needs a square local pattern without ghosts");
    VBH_REQUIRE(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
    VBH_REQUIRE(K >= 1 && K <= 64 && ldv >= K, "KPM supports 1 <= K <= 64 columns per call, ldv >= K");
    VBH_REQUIRE(half_width > 0.0, "half_width must be positive");
    (void)M;
}

// Scratch of vbh_kpm_moments_noalloc: a static device array, one copy per
// GPU, present from the library's load until the process ends.  Host code
// never passes the array's name to a kernel or a copy (on the host it is only
// a stand-in symbol); it uses the device address that hipGetSymbolAddress
// returns for the current device, looked up once per device.
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_noalloc_mu;

static void* noalloc_scratch_ptr() {
    static std::mutex mu;
    static std::map<int, void*> ptrs;
    int dev = 0;
    VBH_CHECK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = ptrs.find(dev);
    if (it != ptrs.end()) return it->second;
    void* p = nullptr;
    VBH_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    VBH_REQUIRE(p != nullptr, "internal: no device address for the static scratch");
    ptrs[dev] = p;
    return p;
}

extern "C" {

VBH_API int vbh_kpm_moments(vbh_matrix_t m, int K, const void* V0, int ldv, int M, int mode, double center,
                            double half_width, int precision, double* moments, void* stream) {
    return guard_on(stream, [&] {
        check_kpm_args(m, K, V0, ldv, M, mode, half_width, moments);
        if (M <= 0) return;
        dispatch_kpm(m, precision, [&](auto* t, auto* sty, auto* vty) {
            using T = std::remove_pointer_t<decltype(t)>;
            using S = std::remove_pointer_t<decltype(sty)>;
            using V = std::remove_pointer_t<decltype(vty)>;
            hipStream_t s = (hipStream_t)stream;
            const bool steps = m->n_rows > 0 && step_count<T>(M, mode) > 0;
            const V* vals = steps ? compute_values<V>(m, s) : nullptr;
            const KpmLayout L = kpm_layout<T, V>(m, vals, K, M, mode);
            char* base = (char*)vbh::scratch(L.total, 4);
            kpm_single<T, S, V>(m, vals, K, (const S*)V0, ldv, M, mode, center, half_width, moments, s, L, base, true);
        });
    });
}

VBH_API int vbh_kpm_moments_noalloc(vbh_matrix_t m, int K, const void* V0, int ldv, int M, int mode, double center,
                                    double half_width, int precision, double* moments, void* stream) {
    return guard([&] {
        check_kpm_args(m, K, V0, ldv, M, mode, half_width, moments);
        if (M <= 0) return;
        char* base = (char*)noalloc_scratch_ptr();
        std::lock_guard<std::mutex> lk(g_noalloc_mu);  // one call per process at a time owns the array
        dispatch_kpm(m, precision, [&](auto* t, auto* sty, auto* vty) {
            using T = std::remove_pointer_t<decltype(t)>;
            using S = std::remove_pointer_t<decltype(sty)>;
            using V = std::remove_pointer_t<decltype(vty)>;
            hipStream_t s = (hipStream_t)stream;
            const bool steps = m->n_rows > 0 && step_count<T>(M, mode) > 0;
            const V* vals = nullptr;
            if (steps) {
                // the matrix's lazily built structures must exist already:
                // building them allocates
                VBH_REQUIRE_STATE(has_row_classes(m, kMfmaMinRows),
                                  "matrix not prepared: its row-class lists are built by its first apply or KPM "
                                  "call (which allocates); run one before vbh_kpm_moments_noalloc");
                if constexpr (std::is_same_v<V, float> || std::is_same_v<V, c64>) {
                    VBH_REQUIRE_STATE(has_vals32(m),
                                      "matrix not prepared: the FP32 value copy is built by its first apply or KPM "
                                      "call at precision 1 or 2 (which allocates); run one before "
                                      "vbh_kpm_moments_noalloc");
                }
                vals = compute_values<V>(m, s);
            }
            const KpmLayout L = kpm_layout<T, V>(m, vals, K, M, mode);
            if (L.total > kNoAllocScratchBytes)
                throw Error(VBH_ERR_ARG, "vbh_kpm_moments_noalloc: this call needs " + std::to_string(L.total) +
                                             " bytes of scratch; the static scratch holds " +
                                             std::to_string(kNoAllocScratchBytes) +
                                             " (use vbh_kpm_moments, or fewer rows, columns or moments)");
            kpm_single<T, S, V>(m, vals, K, (const S*)V0, ldv, M, mode, center, half_width, moments, s, L, base,
                                false);
        });
    });
}

VBH_API int vbh_kpm_moments_batched(int count, const vbh_matrix_t* ms, int K, const void* const* V0s, const int* ldv,
                                    int M, int mode, double center, double half_width, int precision, int graph,
                                    double* moments, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(count >= 1 && ms && V0s && ldv && moments, "null argument");
        VBH_REQUIRE(half_width > 0.0, "half_width must be positive");
        for (int h = 0; h < count; ++h) {
            VBH_REQUIRE(ms[h] && V0s[h] && ldv[h] >= K, "bad batch entry");
            VBH_REQUIRE(ms[h]->dtype == ms[0]->dtype, "batched KPM needs one dtype");
        }
        if (M <= 0) return;
        dispatch_kpm(ms[0], precision, [&](auto* t, auto* sty, auto* vty) {
            using T = std::remove_pointer_t<decltype(t)>;
            using S = std::remove_pointer_t<decltype(sty)>;
            using V = std::remove_pointer_t<decltype(vty)>;
            kpm_batched<T, S, V>(count, ms, K, V0s, ldv, M, mode, center, half_width, graph, moments, (hipStream_t)stream);
        });
    });
}

VBH_API int vbh_set_autotune(int on) {
    return guard([&] {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        g_autotune = on ? 1 : 0;
    });
}

VBH_API int vbh_set_batch_plan(int threshold, int parts) {
    return guard([&] {
        const int mag = parts < 0 ? -parts : parts;
        const int tmag = threshold < 0 ? -threshold : threshold;
        VBH_REQUIRE(threshold == 0 || (tmag >= 1 && tmag <= 17 && threshold != -17 && mag >= 1 && mag <= 64 &&
                                       (mag & (mag - 1)) == 0),
                    "threshold must be 0 (release), 1..17 (class rows on the matrix cores) or -16..-1 (class rows "
                    "on the block-row wave kernel), with parts a power of two in 1..64 (lane family) or its "
                    "negative (row-wave kernel, rows per work-group)");
        std::lock_guard<std::mutex> lk(g_plan_mu);
        g_forced_plan.t = tmag;
        g_forced_plan.brw = threshold < 0;
        g_forced_plan.P = parts > 0 ? parts : 1;
        g_forced_plan.rw = parts < 0 ? -parts : 0;
        g_forced_plan.tuned = false;
    });
}

VBH_API int vbh_kpm_last_plan(int* threshold, int* parts, int* tuned) {
    return guard([&] {
        VBH_REQUIRE(threshold && parts && tuned, "null argument");
        *threshold = g_last_plan.brw ? -g_last_plan.t : g_last_plan.t;
        *parts = g_last_plan.rw ? -g_last_plan.rw : g_last_plan.P;
        *tuned = g_last_plan.tuned ? 1 : 0;
    });
}

VBH_API int vbh_kpm_last_step_time(double* ms, int64_t* launches) {
    return guard([&] {
        VBH_REQUIRE(ms && launches, "null argument");
        *ms = g_last_step_ms;
        *launches = g_last_launches;
    });
}

VBH_API int vbh_cheb_apply(vbh_matrix_t m, int K, const void* X, int ldx, void* Y, int ldy, int nterms,
                           const double* coef, double center, double half_width, int precision, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m && X && Y && coef, "null argument");
        VBH_REQUIRE(m->nbc == m->nbr, "Chebyshev apply needs a square local pattern without ghosts");
        VBH_REQUIRE(ldx >= K && ldy >= K && K >= 1, "bad leading dimensions");
        VBH_REQUIRE(half_width > 0.0, "half_width must be positive");
        dispatch_compute(m, precision, [&](auto* tag, auto* vtag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            using V = std::remove_pointer_t<decltype(vtag)>;
            cheb_apply<T, V>(m, K, (const T*)X, ldx, (T*)Y, ldy, nterms, coef, center, half_width, (hipStream_t)stream);
        });
    });
}

}  // extern "C"
