// Vector and multivector kernels behind DistVector / DistMultiVector
// (vbcsr/core/dist_vector.hpp:82-123, dist_multivector.hpp:179-206,283-372):
// axpby, scale, pointwise products, column-wise conjugate dot products with
// FP64 accumulation, and the counter-based random fill.
#include <cmath>

#include "common.hpp"

namespace vbh {

template <typename T>
__global__ void k_vaxpby(int64_t rows, int K, T alpha, const T* __restrict__ x, int64_t ldx, T beta,
                         T* __restrict__ y, int64_t ldy) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= rows * K) return;
    const int64_t r = e / K;
    const int v = int(e - r * K);
    T& yy = y[r * ldy + v];
    yy = Sc<T>::fma(alpha, x[r * ldx + v], beta * yy);
}

template <typename T>
__global__ void k_vfill(int64_t rows, int K, T val, T* __restrict__ y, int64_t ldy) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= rows * K) return;
    const int64_t r = e / K;
    y[r * ldy + (e - r * K)] = val;
}

template <typename T>
__global__ void k_vscale(int64_t rows, int K, T alpha, T* __restrict__ y, int64_t ldy) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= rows * K) return;
    const int64_t r = e / K;
    const int v = int(e - r * K);
    y[r * ldy + v] = alpha * y[r * ldy + v];
}

template <typename T>
__global__ void k_vpointwise(int64_t rows, int K, const T* __restrict__ x, int64_t ldx, int x_is_vector,
                             T* __restrict__ y, int64_t ldy) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= rows * K) return;
    const int64_t r = e / K;
    const int v = int(e - r * K);
    const T xv = x_is_vector ? x[r] : x[r * ldx + v];
    y[r * ldy + v] = y[r * ldy + v] * xv;
}

// Column dot partials: block b covers rows [b*rows_per, (b+1)*rows_per);
// partial[b][v] = (re, im) of sum conj(x) y, in double.
template <typename T>
__global__ __launch_bounds__(256) void k_bdot(int64_t rows, int K, int v0, int Kt, const T* __restrict__ x, int64_t ldx,
                                              const T* __restrict__ y, int64_t ldy, int64_t rows_per,
                                              double* __restrict__ partial) {
    __shared__ double sre[256], sim[256];
    const int t = threadIdx.x;
    const int Kp = pow2ceil(Kt);
    const int v = t & (Kp - 1);
    const int rstep = 256 / Kp;
    const int64_t rb = int64_t(blockIdx.x) * rows_per;
    int64_t re_ = rb + rows_per;
    if (re_ > rows) re_ = rows;
    double are = 0.0, aim = 0.0;
    if (v < Kt) {
        for (int64_t r = rb + t / Kp; r < re_; r += rstep) {
            const T a = x[r * ldx + v0 + v];
            const T b = y[r * ldy + v0 + v];
            are += redot(a, b);
            if constexpr (Sc<T>::is_complex) aim += double(a.re) * double(b.im) - double(a.im) * double(b.re);
        }
    }
    sre[t] = are;
    sim[t] = aim;
    __syncthreads();
    for (int m = 128; m >= Kp; m >>= 1) {
        if (t < m) {
            sre[t] += sre[t + m];
            sim[t] += sim[t + m];
        }
        __syncthreads();
    }
    if (t < Kt) {
        partial[(int64_t(blockIdx.x) * Kt + t) * 2] = sre[t];
        partial[(int64_t(blockIdx.x) * Kt + t) * 2 + 1] = sim[t];
    }
}

__global__ void k_bdot_final(const double* __restrict__ partial, int nblocks, int E, double* __restrict__ out) {
    const int e = blockIdx.x;  // e < E (= 2*Kt)
    __shared__ double red[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partial[int64_t(b) * E + e];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int m = 128; m >= 1; m >>= 1) {
        if (threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[e] = red[0];
}

// splitmix64 finaliser and the (seed, block, row, column) key mix of
// random_vectors.hpp:64-79 (Steele, Lea & Flood's published constants).
__device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
__device__ inline uint64_t entry_hash(uint64_t seed, uint64_t block, uint64_t row, uint64_t col) {
    uint64_t h = splitmix64(seed + 0x9e3779b97f4a7c15ULL);
    h = splitmix64(h ^ (block * 0xbf58476d1ce4e5b9ULL));
    h = splitmix64(h ^ (row * 0x94d049bb133111ebULL));
    return splitmix64(h ^ (col * 0xd6e8feb86659fd93ULL));
}
__device__ inline double uniform01(uint64_t h) { return double(h >> 11) * (1.0 / 9007199254740992.0); }

// One thread per (block, column): kind 0 Rademacher (+1 iff bit 63 set),
// kind 1 random phase exp(2 pi i u), kind 2 Gaussian (Box-Muller), as in
// random_vectors.hpp:99-143.
template <typename T>
__global__ void k_set_random(int nblocks, const int* __restrict__ bs, const int* __restrict__ gids,
                             const int64_t* __restrict__ offs, int K, uint64_t seed, int kind, double scale,
                             T* __restrict__ y, int64_t ldy) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(nblocks) * K) return;
    const int b = int(e / K);
    const int v = int(e - int64_t(b) * K);
    const uint64_t g = uint64_t(gids[b]);
    for (int a = 0; a < bs[b]; ++a) {
        const uint64_t h = entry_hash(seed, g, uint64_t(a), uint64_t(v));
        double re = 0.0, im = 0.0;
        if (kind == 0) {
            re = (h >> 63) ? 1.0 : -1.0;
        } else if (kind == 1) {
            const double th = 2.0 * M_PI * uniform01(h);
            re = cos(th);
            im = sin(th);
        } else {
            const uint64_t h2 = splitmix64(h);
            auto normal = [](uint64_t h1, uint64_t hh2) {
                double u1 = uniform01(h1);
                if (u1 < 1e-300) u1 = 1e-300;
                return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * uniform01(hh2));
            };
            if constexpr (Sc<T>::is_complex) {
                const double s = 1.0 / sqrt(2.0);
                re = s * normal(h, h2);
                im = s * normal(h2, splitmix64(h2));
            } else {
                re = normal(h, h2);
            }
        }
        y[(offs[b] + a) * ldy + v] = Sc<T>::make(re * scale, im * scale);
    }
}

static inline unsigned nblk(int64_t n) { return unsigned((n + 255) / 256); }

template <typename F>
static void dispatch_any(int dtype, F&& f) {
    switch (dtype) {
        case VBH_F64: f((double*)nullptr); break;
        case VBH_C128: f((c128*)nullptr); break;
        case VBH_F32: f((float*)nullptr); break;
        case VBH_C64: f((c64*)nullptr); break;
        default: throw Error(VBH_ERR_ARG, "unknown dtype");
    }
}

}  // namespace vbh

using namespace vbh;

extern "C" {

VBH_API int vbh_vec_axpby(int dtype, int64_t rows, int K, const double* alpha, const void* x, int ldx,
                          const double* beta, void* y, int ldy, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(ldx >= K && ldy >= K && alpha && beta, "bad axpby arguments");
        if (rows * K == 0) return;
        dispatch_any(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL(k_vaxpby<T>, dim3(nblk(rows * K)), dim3(256), 0, (hipStream_t)stream, rows, K,
                               Sc<T>::make(alpha[0], alpha[1]), (const T*)x, int64_t(ldx),
                               Sc<T>::make(beta[0], beta[1]), (T*)y, int64_t(ldy));
        });
        VBH_CHECK_HIP(hipGetLastError());
    });
}

VBH_API int vbh_vec_scale(int dtype, int64_t rows, int K, const double* alpha, void* y, int ldy, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(ldy >= K && alpha, "bad scale arguments");
        if (rows * K == 0) return;
        dispatch_any(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL(k_vscale<T>, dim3(nblk(rows * K)), dim3(256), 0, (hipStream_t)stream, rows, K,
                               Sc<T>::make(alpha[0], alpha[1]), (T*)y, int64_t(ldy));
        });
        VBH_CHECK_HIP(hipGetLastError());
    });
}

VBH_API int vbh_vec_fill(int dtype, int64_t rows, int K, const double* value, void* y, int ldy, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(ldy >= K && value, "bad fill arguments");
        if (rows * K == 0) return;
        dispatch_any(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL(k_vfill<T>, dim3(nblk(rows * K)), dim3(256), 0, (hipStream_t)stream, rows, K,
                               Sc<T>::make(value[0], value[1]), (T*)y, int64_t(ldy));
        });
        VBH_CHECK_HIP(hipGetLastError());
    });
}

VBH_API int vbh_vec_pointwise(int dtype, int64_t rows, int K, const void* x, int ldx, int x_is_vector, void* y,
                              int ldy, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(ldy >= K && (x_is_vector || ldx >= K), "bad pointwise arguments");
        if (rows * K == 0) return;
        dispatch_any(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL(k_vpointwise<T>, dim3(nblk(rows * K)), dim3(256), 0, (hipStream_t)stream, rows, K,
                               (const T*)x, int64_t(ldx), x_is_vector, (T*)y, int64_t(ldy));
        });
        VBH_CHECK_HIP(hipGetLastError());
    });
}

VBH_API int vbh_vec_bdot(int dtype, int64_t rows, int K, const void* x, int ldx, const void* y, int ldy, double* out,
                         void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(ldx >= K && ldy >= K && out, "bad bdot arguments");
        hipStream_t s = (hipStream_t)stream;
        for (int e = 0; e < 2 * K; ++e) out[e] = 0.0;
        if (rows == 0 || K == 0) return;
        const int nb = int(std::min<int64_t>(1024, std::max<int64_t>(1, rows / 256)));
        const int64_t rows_per = (rows + nb - 1) / nb;
        for (int v0 = 0; v0 < K; v0 += 256) {
            const int Kt = std::min(256, K - v0);
            double* part = (double*)scratch(sizeof(double) * 2 * Kt * (nb + 1), 3);
            double* fin = part + int64_t(2) * Kt * nb;
            dispatch_any(dtype, [&](auto* tag) {
                using T = std::remove_pointer_t<decltype(tag)>;
                hipLaunchKernelGGL(k_bdot<T>, dim3(nb), dim3(256), 0, s, rows, K, v0, Kt, (const T*)x, int64_t(ldx),
                                   (const T*)y, int64_t(ldy), rows_per, part);
            });
            VBH_CHECK_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_bdot_final, dim3(2 * Kt), dim3(256), 0, s, part, nb, 2 * Kt, fin);
            VBH_CHECK_HIP(hipGetLastError());
            VBH_CHECK_HIP(hipMemcpyAsync(out + 2 * v0, fin, sizeof(double) * 2 * Kt, hipMemcpyDeviceToHost, s));
            VBH_CHECK_HIP(hipStreamSynchronize(s));
        }
    });
}

VBH_API int vbh_vec_set_random(int dtype, int nblocks, const int* bs_dev, const int* gids_dev, const int64_t* offs_dev,
                               int K, uint64_t seed, int kind, double scale, void* y, int ldy, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(ldy >= K, "ldy must be >= K");
        VBH_REQUIRE(kind >= 0 && kind <= 2, "kind must be 0 (Rademacher), 1 (random phase) or 2 (Gaussian)");
        VBH_REQUIRE(!(kind == 1 && (dtype == VBH_F64 || dtype == VBH_F32)),
                    "RandomKind::RandomPhase needs a complex scalar type; use Rademacher for a real one.");
        if (int64_t(nblocks) * K == 0) return;
        dispatch_any(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL(k_set_random<T>, dim3(nblk(int64_t(nblocks) * K)), dim3(256), 0, (hipStream_t)stream,
                               nblocks, bs_dev, gids_dev, offs_dev, K, seed, kind, scale, (T*)y, int64_t(ldy));
        });
        VBH_CHECK_HIP(hipGetLastError());
    });
}

}  // extern "C"
