// Thresholded block SpGEMM C = A B on gfx950.
//
// Rules (the reference's, restated): block norms are Frobenius norms in FP64
// (block_csr.hpp:2486-2502); the symbolic estimate est_ij = sum_k
// |A_ik||B_kj| keeps column j iff est_ij > thr, strictly
// (ops/spmm/common.hpp:507-690, keep rule :664); the numeric phase skips a
// pair when |A_ik||B_kj| < thr / max(1, nnzb_row_A(i)) (ops/spmm/vbcsr.hpp:
// 371-387) and accumulates the rest into C's symbolic blocks; finally, when
// thr > 0, blocks with |C_ij| < thr are dropped (block_csr.hpp:1399-1528).
//
// Symbolic phase: one work-group per row of A with an open-addressing hash
// table (key = global column, value = FP64 estimate) sized to
// pow2ceil(2 min(bound_i, nbc) + 1), where bound_i = sum_k nnzb(B_k) is the
// number of multiplication paths of the row.  Tables up to 4096 slots live in
// LDS (three size classes, so small rows do not pay for big tables); wider
// rows use a per-row slice of device memory.  Kept columns are written
// unsorted to a per-row region; the host sorts each row (ascending global
// column, common.hpp:668) and builds C's pattern.
//
// Numeric phase: one work-group per row of C.  The A slots of the row are
// visited in order.  Blocks up to 8 x 8 (k_numeric_owned): each C block is
// owned by one wave, which scans the B slots of every row k lane-parallel and
// looks C_ij up in an LDS column map of the row; no barriers.  Larger blocks
// (k_numeric): the four waves take the B slots of row k round-robin, locate
// C_ij by binary search in C's sorted row, and one lane per element of the
// r x c output block accumulates sum_m A[a][m] B[m][b], with a work-group
// barrier between consecutive A slots.  Either way every C element is
// updated by one thread at a time in a fixed order -- deterministic, no
// atomics, the same sums in the same order.  numeric = 1 runs the block
// products in FP32 (FP32 accumulation, FP64 result); numeric = 2 rounds A and
// B to BF16 and multiplies on the CDNA4 matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation), 16-column strips of the
// row's output panel per wave, B blocks found through an LDS column map.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "common.hpp"

namespace vbh {

template <typename T>
__global__ __launch_bounds__(256) void k_norms(const T* __restrict__ vals, const int64_t* __restrict__ voff,
                                               int64_t nnzb, double* __restrict__ out) {
    const int64_t k = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= nnzb) return;
    double s = 0.0;
    for (int64_t i = voff[k] + lane; i < voff[k + 1]; i += 64) s += sqabs(vals[i]);
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if (lane == 0) out[k] = sqrt(s);
}

struct SymArgs {
    const int* a_row_ptr;
    const int* a_col_ind;
    const double* a_norm;
    const int* b_row_ptr;
    const int* b_col_ind;
    const double* b_norm;
    const int* rows;        // rows handled by this launch
    int nrows;
    const int* tsize;       // hash table size per row (power of two)
    const int64_t* out_off; // per row: start of its kept-column region
    int* out_cols;          // kept columns (unsorted)
    int* out_count;         // per row count
    int64_t* pairs;         // per row multiplication paths
    double thr;
    int upper;              // keep only block columns j >= the row i (upper triangle)
    int* g_keys;            // global tables (LDS == false)
    double* g_vals;
    const int64_t* g_off;   // per launch-row table offset
};

template <bool LDS>
__global__ __launch_bounds__(256) void k_symbolic(SymArgs a) {
    // double-typed so the dynamic area is 8-byte aligned: the FP64 values are
    // updated with 8-byte LDS atomics, which require natural alignment
    extern __shared__ double smem[];
    const int q = blockIdx.x;
    if (q >= a.nrows) return;
    const int i = a.rows[q];
    const int S = a.tsize[i];
    int* keys;
    double* vals;
    if constexpr (LDS) {
        vals = smem;
        keys = reinterpret_cast<int*>(smem + S);
    } else {
        keys = a.g_keys + a.g_off[q];
        vals = a.g_vals + a.g_off[q];
    }
    for (int t = threadIdx.x; t < S; t += blockDim.x) {
        keys[t] = -1;
        vals[t] = 0.0;
    }
    __syncthreads();
    const int mask = S - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t npairs = 0;
    for (int ka = a.a_row_ptr[i] + wave; ka < a.a_row_ptr[i + 1]; ka += 4) {
        const int k = a.a_col_ind[ka];
        const double na = a.a_norm[ka];
        const int b0 = a.b_row_ptr[k], b1 = a.b_row_ptr[k + 1];
        npairs += b1 - b0;
        for (int kb = b0 + lane; kb < b1; kb += 64) {
            const int j = a.b_col_ind[kb];
            if (a.upper && j < i) continue;  // common.hpp:625 (upper_only)
            const double v = na * a.b_norm[kb];
            int h = j & mask;
            while (true) {
                const int prev = atomicCAS(&keys[h], -1, j);
                if (prev == -1 || prev == j) {
                    atomicAdd(&vals[h], v);
                    break;
                }
                h = (h + 1) & mask;
            }
        }
    }
    __syncthreads();
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int* out = a.out_cols + a.out_off[i];
    for (int t = threadIdx.x; t < S; t += blockDim.x) {
        if (keys[t] >= 0 && vals[t] > a.thr) {
            const int pos = atomicAdd(&cnt, 1);
            out[pos] = keys[t];
        }
    }
    if (lane == 0) atomicAdd((unsigned long long*)&a.pairs[i], (unsigned long long)npairs);
    __syncthreads();
    if (threadIdx.x == 0) a.out_count[i] = cnt;
}

// Symbolic phase for rows whose hash table would not fit in LDS, when C has
// nbc <= kNumMapCols block columns and thr >= 0: the estimates live in a
// dense per-column FP64 array in LDS (no hashing, no global tables).  Column j
// is kept iff est_j > thr, which for thr >= 0 also drops every column no path
// reached (est 0), as the hash table's absent keys are.  The kept columns are
// written in ascending order (a block-wide scan over per-thread column
// ranges), so the host sort finds the row already sorted.
__global__ __launch_bounds__(256) void k_symbolic_dense(SymArgs a, int nbc) {
    extern __shared__ double est[];  // [nbc]
    __shared__ int scan[256];
    const int q = blockIdx.x;
    if (q >= a.nrows) return;
    const int i = a.rows[q];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int t = tid; t < nbc; t += 256) est[t] = 0.0;
    __syncthreads();
    int64_t npairs = 0;
    for (int ka = a.a_row_ptr[i] + wave; ka < a.a_row_ptr[i + 1]; ka += 4) {
        const int k = a.a_col_ind[ka];
        const double na = a.a_norm[ka];
        const int b0 = a.b_row_ptr[k], b1 = a.b_row_ptr[k + 1];
        npairs += b1 - b0;
        for (int kb = b0 + lane; kb < b1; kb += 64) {
            const int j = a.b_col_ind[kb];
            if (a.upper && j < i) continue;  // common.hpp:625 (upper_only)
            atomicAdd(&est[j], na * a.b_norm[kb]);
        }
    }
    __syncthreads();
    // thread t owns columns [t*C, (t+1)*C): count, exclusive scan, write in order
    const int C = (nbc + 255) / 256;
    const int j0 = tid * C, j1 = j0 + C < nbc ? j0 + C : nbc;
    int mine = 0;
    for (int j = j0; j < j1; ++j) mine += est[j] > a.thr;
    scan[tid] = mine;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int v = tid >= off ? scan[tid - off] : 0;
        __syncthreads();
        scan[tid] += v;
        __syncthreads();
    }
    int* out = a.out_cols + a.out_off[i];
    int pos = scan[tid] - mine;
    for (int j = j0; j < j1; ++j)
        if (est[j] > a.thr) out[pos++] = j;
    if (lane == 0) atomicAdd((unsigned long long*)&a.pairs[i], (unsigned long long)npairs);
    if (tid == 255) a.out_count[i] = scan[255];
}

// Dense symbolic phase for products whose candidate pattern is dense: the
// estimates est_ij = sum_k ||A_ik|| ||B_kj|| (common.hpp:620-664) of all rows
// at once as one FP64 product of the block-norm matrices (zeros where no block
// is stored; the padded extents are multiples of the tile sizes).  A
// work-group computes a 64 x 64 tile over k-steps of 16 staged in LDS; each
// wave a 32 x 32 part as 2 x 2 tiles of v_mfma_f64_16x16x4f64 (operand lane
// l: row or column l % 16, k l / 16; result register e of lane l: row
// 4e + l / 16, column l % 16).  Before, one thread per 4 x 4 micro-tile on
// the vector ALUs: 12.1 ms for a 5000^3 product.
constexpr int kEstTile = 64;
constexpr int kEstK = 16;
typedef double est_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_gemm_f64(const double* __restrict__ A, const double* __restrict__ B,
                                                  double* __restrict__ C, int64_t Kp, int64_t Np, int tiles_n,
                                                  int n_tiles) {
    __shared__ double as[kEstK][kEstTile + 1];
    __shared__ double bs[kEstK][kEstTile + 1];
    const int gid = xcd_chunk(int(gridDim.x));
    if (gid >= n_tiles) return;
    const int tm = gid / tiles_n, tn = gid - tm * tiles_n;
    const int64_t m0 = int64_t(tm) * kEstTile, n0 = int64_t(tn) * kEstTile;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1, li = lane & 15, lk = lane >> 4;
    est_d4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = est_d4{0.0, 0.0, 0.0, 0.0};
    for (int64_t k0 = 0; k0 < Kp; k0 += kEstK) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            as[e & 15][e >> 4] = A[(m0 + (e >> 4)) * Kp + k0 + (e & 15)];
            bs[e >> 6][e & 63] = B[(k0 + (e >> 6)) * Np + n0 + (e & 63)];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kEstK; kk += 4) {
            double a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = as[kk + lk][wm * 32 + i * 16 + li];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = bs[kk + lk][wn * 32 + j * 16 + li];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                C[(m0 + wm * 32 + i * 16 + 4 * e + lk) * Np + n0 + wn * 32 + j * 16 + li] = acc[i][j][e];
}

// Kept columns of row i from the dense estimates (est_ij > thr, j >= i when
// upper), ascending, at out_cols[out_off[i]..], and their count.
__global__ __launch_bounds__(256) void k_est_compact(const double* __restrict__ est, int64_t ld, int nbc, double thr,
                                                     int upper, const int64_t* __restrict__ out_off,
                                                     int* __restrict__ out_cols, int* __restrict__ out_count) {
    __shared__ int scan[256];
    const int i = blockIdx.x, tid = threadIdx.x;
    const double* row = est + int64_t(i) * ld;
    const int C = (nbc + 255) / 256;
    const int j0 = tid * C, j1 = j0 + C < nbc ? j0 + C : nbc;
    auto keep = [&](int j) { return row[j] > thr && (!upper || j >= i); };
    int mine = 0;
    for (int j = j0; j < j1; ++j) mine += keep(j);
    scan[tid] = mine;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int v = tid >= off ? scan[tid - off] : 0;
        __syncthreads();
        scan[tid] += v;
        __syncthreads();
    }
    int* out = out_cols + out_off[i];
    int pos = scan[tid] - mine;
    for (int j = j0; j < j1; ++j)
        if (keep(j)) out[pos++] = j;
    if (tid == 255) out_count[i] = scan[255];
}

// Row i's kept columns, out_cols[out_off[i] .. + (crp[i+1] - crp[i])), to
// cci[crp[i] ..]: one work-group per row.
__global__ __launch_bounds__(256) void k_compact_cols(const int* __restrict__ out_cols,
                                                      const int64_t* __restrict__ out_off,
                                                      const int* __restrict__ crp, int* __restrict__ cci) {
    const int i = blockIdx.x;
    const int n = crp[i + 1] - crp[i];
    const int* src = out_cols + out_off[i];
    int* dst = cci + crp[i];
    for (int t = threadIdx.x; t < n; t += 256) dst[t] = src[t];
}

template <typename T, typename CT>
struct NumArgs {
    const int* a_row_ptr;
    const int* a_col_ind;
    const int64_t* a_voff;
    const T* a_vals;
    const double* a_norm;
    const int* a_rbs;
    const int* a_cbs;  // = B's row block sizes
    const int* b_row_ptr;
    const int* b_col_ind;
    const int64_t* b_voff;
    const T* b_vals;
    const double* b_norm;
    const int* c_row_ptr;
    const int* c_col_ind;
    const int64_t* c_voff;
    const int* c_cbs;
    T* c_vals;
    unsigned long long* npairs;
    double thr;
    // CT != T (FP32 products) in k_numeric_owned: A and B from their CT value
    // copies, C accumulated in a CT array (widened into c_vals afterwards)
    const CT* a_c;
    const CT* b_c;
    CT* c_c;
    double* c_norm;  // [nnzb(C)] block Frobenius norms for the filter, or null
};

__device__ inline int find_col(const int* __restrict__ cols, int lo, int hi, int j) {
    // first position in [lo, hi) with cols[pos] >= j; -1 if absent
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cols[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <typename T>
__device__ inline T to_ct(double x) { return T(x); }

// CT: the product/accumulation type (T itself, or float / c64).
template <typename T, typename CT>
__device__ inline CT narrow(T x) {
    if constexpr (std::is_same_v<T, CT>) return x;
    else if constexpr (Sc<T>::is_complex) return CT{float(x.re), float(x.im)};
    else return CT(x);
}
template <typename T, typename CT>
__device__ inline T widen(CT x) {
    if constexpr (std::is_same_v<T, CT>) return x;
    else if constexpr (Sc<T>::is_complex) return T{double(x.re), double(x.im)};
    else return T(x);
}

// Frobenius norms of C row i's blocks from the values a numeric kernel has
// just written (the filter's norm pass, fused into the numeric phase): wave
// w of nw takes the row's slots c0 + w, c0 + w + nw, ...; lane l sums the
// squares of elements l, l + 64, ... of a block, then a butterfly over the
// wave -- the order of k_block_norms, so the filter keeps exactly the blocks
// a separate norm pass would keep.  With NS the block is read from the
// narrow accumulation array and widened into C's values on the way.
template <typename T, typename CT, bool NS>
__device__ inline void row_block_norms(const CT* __restrict__ src, T* __restrict__ vals,
                                       const int64_t* __restrict__ voff, int c0, int c1, int wave, int nw, int lane,
                                       double* __restrict__ norm) {
    for (int q = c0 + wave; q < c1; q += nw) {
        const int64_t b = voff[q], e = voff[q + 1];
        double sum = 0.0;
        for (int64_t x = b + lane; x < e; x += 64) {
            T v;
            if constexpr (NS) {
                v = widen<T, CT>(src[x]);
                vals[x] = v;
            } else {
                v = vals[x];
            }
            sum += sqabs(v);
        }
        for (int m = 32; m >= 1; m >>= 1) sum += __shfl_xor(sum, m, 64);
        if (lane == 0 && norm) norm[q] = sqrt(sum);
    }
}

template <typename T, typename CT>
__global__ __launch_bounds__(256) void k_numeric(NumArgs<T, CT> a) {
    const int i = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int a0 = a.a_row_ptr[i], a1 = a.a_row_ptr[i + 1];
    const int c0 = a.c_row_ptr[i], c1 = a.c_row_ptr[i + 1];
    if (c0 == c1) return;
    const int r = a.a_rbs[i];
    const int row_count = a1 - a0;
    const double row_eps = a.thr / double(row_count > 1 ? row_count : 1);
    unsigned long long used = 0;
    for (int ka = a0; ka < a1; ++ka) {
        const int k = a.a_col_ind[ka];
        const int m = a.a_cbs[k];
        const double na = a.a_norm[ka];
        const T* A = a.a_vals + a.a_voff[ka];
        for (int kb = a.b_row_ptr[k] + wave; kb < a.b_row_ptr[k + 1]; kb += 4) {
            if (na * a.b_norm[kb] < row_eps) continue;
            const int j = a.b_col_ind[kb];
            const int pos = find_col(a.c_col_ind, c0, c1, j);
            if (pos >= c1 || a.c_col_ind[pos] != j) continue;
            ++used;
            const int c = a.c_cbs[j];
            const T* B = a.b_vals + a.b_voff[kb];
            T* C = a.c_vals + a.c_voff[pos];
            for (int e = lane; e < r * c; e += 64) {
                const int ra = e / c, cb = e - ra * c;
                CT acc = narrow<T, CT>(C[e]);
                for (int q = 0; q < m; ++q)
                    acc = Sc<CT>::fma(narrow<T, CT>(A[ra * m + q]), narrow<T, CT>(B[int64_t(q) * c + cb]), acc);
                C[e] = widen<T, CT>(acc);
            }
        }
        __syncthreads();
    }
    if (lane == 0 && used) atomicAdd(&a.npairs[0], used);
    if (a.c_norm) row_block_norms<T, CT, false>(nullptr, a.c_vals, a.c_voff, c0, c1, wave, 4, lane, a.c_norm);
}

// Numeric phase for blocks up to 8 x 8 (the common case): C slot q of the row
// belongs to wave q % 4, so the waves never touch each other's blocks and
// run without barriers.  For every A slot each wave scans the B slots of row
// k 64 at a time, one per lane: the skip test, C's slot of column j (from an
// LDS map of the row when MAP, else a per-lane binary search), and the
// offsets, all lane-parallel; the pairs the wave owns are compacted by a
// ballot and processed one after another by lanes (a, b) = (lane / 8,
// lane % 8), each holding its row of A_ik in registers; for real types a B
// block is loaded whole (one coalesced load, element e in lane e) and each
// lane gathers its column b by cross-lane shuffles.  Every C element
// still receives its products in A-slot order and, within a product, in
// inner-index order, exactly as in k_numeric.
constexpr int kNumMapCols = 8192;  // LDS column map (nbc ints, dynamic) when nbc <= this
#ifndef VBH_NUM_WAVES
#define VBH_NUM_WAVES 4
#endif
constexpr int kNumWaves = VBH_NUM_WAVES;  // waves per work-group, each owning every kNumWaves-th C slot (4 beats 2 and 8)

// FP32 value rounded to the nearest BF16 (ties to even), kept in FP32: the
// product of two such values is exact in FP32, so FP32 FMAs of them are BF16
// products with FP32 accumulation, as on the BF16 matrix cores.
__device__ inline float bf16_round(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

// BF: BF16 products (numeric 2) from the FP32 value copies, rounded to BF16
// as they are loaded; accumulation in FP32 as for numeric 1.
template <typename T, typename CT, bool MAP, bool BF = false>
__global__ __launch_bounds__(64 * kNumWaves) void k_numeric_owned(NumArgs<T, CT> a, int nbc) {
    constexpr bool NS = !std::is_same_v<T, CT>;  // narrow storage: a_c, b_c, c_c
    static_assert(!BF || std::is_same_v<CT, float>, "BF16 products from FP32 copies");
    extern __shared__ int colmap[];  // nbc entries when MAP
    const int i = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int a0 = a.a_row_ptr[i], a1 = a.a_row_ptr[i + 1];
    const int c0 = a.c_row_ptr[i], c1 = a.c_row_ptr[i + 1];
    if (c0 == c1) return;
    if constexpr (MAP) {
        for (int j = tid; j < nbc; j += 64 * kNumWaves) colmap[j] = -1;
        __syncthreads();
        for (int q = c0 + tid; q < c1; q += 64 * kNumWaves) colmap[a.c_col_ind[q]] = q;
        __syncthreads();
    }
    const int r = a.a_rbs[i];
    const int row_count = a1 - a0;
    const double row_eps = a.thr / double(row_count > 1 ? row_count : 1);
    const int ra = lane >> 3, cb = lane & 7;
    unsigned long long used = 0;
    for (int ka = a0; ka < a1; ++ka) {
        const int k = a.a_col_ind[ka];
        const int m = a.a_cbs[k];
        const double na = a.a_norm[ka];
        const T* A = a.a_vals + a.a_voff[ka];
        CT arow[8];
        if constexpr (NS) {
            const CT* Ac = a.a_c + a.a_voff[ka];
#pragma unroll
            for (int q = 0; q < 8; ++q) arow[q] = (ra < r && q < m) ? Ac[ra * m + q] : Sc<CT>::zero();
            if constexpr (BF)
#pragma unroll
                for (int q = 0; q < 8; ++q) arow[q] = bf16_round(arow[q]);
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q)
                arow[q] = (ra < r && q < m) ? narrow<T, CT>(A[ra * m + q]) : Sc<CT>::zero();
        }
        const int kb0 = a.b_row_ptr[k], kb1 = a.b_row_ptr[k + 1];
        for (int base = kb0; base < kb1; base += 64) {
            const int kb = base + lane;
            int pos = -1, j = 0;
            if (kb < kb1) {
                j = a.b_col_ind[kb];
                if (!(na * a.b_norm[kb] < row_eps)) {
                    if constexpr (MAP) {
                        pos = colmap[j];
                    } else {
                        const int p = find_col(a.c_col_ind, c0, c1, j);
                        pos = (p < c1 && a.c_col_ind[p] == j) ? p : -1;
                    }
                }
            }
            const bool mine = pos >= 0 && (pos - c0) % kNumWaves == wave;
            uint64_t mask = __ballot(mine);
            used += __popcll(mask);
            int64_t boff = 0, coff = 0;
            int cw = 0;
            if (mine) {
                boff = a.b_voff[kb];
                coff = a.c_voff[pos];
                cw = a.c_cbs[j];
            }
            // UP owned pairs per pass: all their loads are issued before the
            // first product (indices clamped into the blocks, so every load
            // is unconditional and in bounds); the pairs of one pass have
            // distinct C blocks, and lane (a, b) always owns element (a, b)
            constexpr int UP = 2;  // measured on McWeeny products: 2 beats 1 (13 %) and 4 (3 %)
            const int rac = ra < r ? ra : r - 1;
            while (mask) {
                int tt[UP];
                tt[0] = __builtin_ctzll(mask);
                mask &= mask - 1;
#pragma unroll
                for (int u = 1; u < UP; ++u) {
                    tt[u] = mask ? __builtin_ctzll(mask) : -1;
                    if (mask) mask &= mask - 1;
                }
                int cc[UP];
                int64_t bo[UP], co[UP];
                CT acc[UP], bv[UP][8];
#pragma unroll
                for (int u = 0; u < UP; ++u) {
                    const int t = tt[u] >= 0 ? tt[u] : tt[0];  // a missing pair repeats pair 0 (not stored)
                    cc[u] = __builtin_amdgcn_readlane(cw, t);
                    bo[u] = int64_t((uint64_t(uint32_t(__builtin_amdgcn_readlane(int(boff >> 32), t))) << 32) |
                                    uint32_t(__builtin_amdgcn_readlane(int(boff), t)));
                    co[u] = int64_t((uint64_t(uint32_t(__builtin_amdgcn_readlane(int(coff >> 32), t))) << 32) |
                                    uint32_t(__builtin_amdgcn_readlane(int(coff), t)));
                    const int cbc = cb < cc[u] ? cb : cc[u] - 1;
                    if constexpr (NS && std::is_same_v<CT, float>) {
                        acc[u] = a.c_c[co[u] + rac * cc[u] + cbc];
                        const int nb = m * cc[u];
                        float ball = a.b_c[bo[u] + (lane < nb ? lane : nb - 1)];
                        if constexpr (BF) ball = bf16_round(ball);
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const int src = q * cc[u] + cbc;
                            const float v = __shfl(ball, src < nb ? src : nb - 1, 64);
                            bv[u][q] = q < m ? v : 0.0f;
                        }
                    } else if constexpr (NS) {
                        acc[u] = a.c_c[co[u] + rac * cc[u] + cbc];
#pragma unroll
                        for (int q = 0; q < 8; ++q)
                            bv[u][q] = q < m ? a.b_c[bo[u] + int64_t(q) * cc[u] + cbc] : Sc<CT>::zero();
                    } else if constexpr (std::is_same_v<T, double>) {
                        // the whole B block in one coalesced load (lane e holds
                        // element e), then each lane takes its column by shuffles
                        acc[u] = a.c_vals[co[u] + rac * cc[u] + cbc];
                        const int nb = m * cc[u];
                        const double ball = a.b_vals[bo[u] + (lane < nb ? lane : nb - 1)];
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const int src = q * cc[u] + cbc;
                            const double v = __shfl(ball, src < nb ? src : nb - 1, 64);
                            bv[u][q] = q < m ? v : 0.0;
                        }
                    } else {
                        acc[u] = narrow<T, CT>(a.c_vals[co[u] + rac * cc[u] + cbc]);
#pragma unroll
                        for (int q = 0; q < 8; ++q)
                            bv[u][q] = q < m ? narrow<T, CT>(a.b_vals[bo[u] + int64_t(q) * cc[u] + cbc]) : Sc<CT>::zero();
                    }
                }
#pragma unroll
                for (int u = 0; u < UP; ++u) {
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (q < m) acc[u] = Sc<CT>::fma(arow[q], bv[u][q], acc[u]);
                    if (tt[u] >= 0 && ra < r && cb < cc[u]) {
                        if constexpr (NS) a.c_c[co[u] + ra * cc[u] + cb] = acc[u];
                        else a.c_vals[co[u] + ra * cc[u] + cb] = widen<T, CT>(acc[u]);
                    }
                }
            }
        }
    }
    if (lane == 0 && used) atomicAdd(&a.npairs[0], used);
    // the wave's own slots: widened into C (narrow storage) and their norms
    if (NS || a.c_norm) {
        __syncthreads();
        row_block_norms<T, CT, NS>(a.c_c, a.c_vals, a.c_voff, c0, c1, wave, kNumWaves, lane, a.c_norm);
    }
}

// ------------------------------------------------------------ BF16 MFMA numeric
// Real matrices only.  The output panel of one C row i is r x (sum_j c_j);
// each A slot k contributes A_ik (r x m) times the row panel of B_k restricted
// to C's columns.  A wave owns 16 output columns of the panel: for every A
// slot it gathers the matching B_kj columns (zero where C_ij is absent or the
// pair is skipped), pads the inner dimension m to 32 and the r rows to 16,
// and issues one v_mfma_f32_16x16x32_bf16.
typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ inline short to_bf16(double x) {
    // round-to-nearest-even on the FP32 value
    return short(__float_as_uint(bf16_round(float(x))) >> 16);
}

struct MfmaArgs {
    const int* a_row_ptr;
    const int* a_col_ind;
    const int64_t* a_voff;
    const double* a_vals;
    const double* a_norm;
    const int* a_rbs;
    const int* a_cbs;
    const int* b_row_ptr;
    const int* b_col_ind;
    const int64_t* b_voff;
    const double* b_vals;
    const double* b_norm;
    const int* c_row_ptr;
    const int* c_col_ind;
    const int64_t* c_voff;
    const int* c_cbs;
    const int* c_colstart;  // per C slot: first panel column of the block within its row
    const int* panel_width; // per C row: total panel columns
    double* c_vals;
    double thr;
    double* c_norm;                // [nnzb(C)] block norms for the filter, or null
    unsigned long long* npairs;    // multiplied (A slot, C block) pairs
};

// Panel column -> (C slot, column within block), found by binary search over
// c_colstart of the row.
__device__ inline int slot_of_panel_col(const int* __restrict__ colstart, int c0, int c1, int pc) {
    int lo = c0, hi = c1 - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (colstart[mid] <= pc) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_numeric_mfma_bf16(MfmaArgs a) {
    const int i = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int a0 = a.a_row_ptr[i], a1 = a.a_row_ptr[i + 1];
    const int c0 = a.c_row_ptr[i], c1 = a.c_row_ptr[i + 1];
    if (c0 == c1) return;
    const int r = a.a_rbs[i];
    const int width = a.panel_width[i];
    const int row_count = a1 - a0;
    const double row_eps = a.thr / double(row_count > 1 ? row_count : 1);
    // MFMA 16x16x32 bf16 operand layout (per lane): A operand lane l holds
    // row (l % 16), k-elements 8*(l / 16) .. +7; B operand lane l holds column
    // (l % 16), k-elements 8*(l / 16) .. +7; result lane l holds column
    // (l % 16), rows 4*(l / 16) .. +3.
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 8;
    unsigned long long used = 0;
    for (int pc0 = wave * 16; pc0 < width; pc0 += 64) {
        const int pc = pc0 + lr;  // this lane's panel column (B operand / result)
        int slot = -1, jcol = -1, cb = 0, cwidth = 0;
        if (pc < width) {
            slot = slot_of_panel_col(a.c_colstart, c0, c1, pc);
            jcol = a.c_col_ind[slot];
            cb = pc - a.c_colstart[slot];
            cwidth = a.c_cbs[jcol];
        }
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int ka = a0; ka < a1; ++ka) {
            const int k = a.a_col_ind[ka];
            const int m = a.a_cbs[k];
            const double na = a.a_norm[ka];
            const double* A = a.a_vals + a.a_voff[ka];
            // this lane's B_kj column, if block (k, jcol) exists and survives the skip
            const double* Bcol = nullptr;
            if (slot >= 0) {
                const int b0 = a.b_row_ptr[k], b1 = a.b_row_ptr[k + 1];
                const int pb = find_col(a.b_col_ind, b0, b1, jcol);
                if (pb < b1 && a.b_col_ind[pb] == jcol && !(na * a.b_norm[pb] < row_eps))
                    Bcol = a.b_vals + a.b_voff[pb] + cb;
            }
            // one lane per (A slot, C block): the block's first column in lane group 0
            used += __popcll(__ballot(Bcol != nullptr && cb == 0 && lane < 16));
            for (int kk = 0; kk < m; kk += 32) {
                bf16x8 av, bv;
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int q = kk + lk + t;
                    av[t] = (lr < r && q < m) ? to_bf16(A[lr * m + q]) : short(0);
                    bv[t] = (Bcol && q < m) ? to_bf16(Bcol[int64_t(q) * cwidth]) : short(0);
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
            }
        }
        if (slot >= 0) {
            double* C = a.c_vals + a.c_voff[slot];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = (lane >> 4) * 4 + t;
                if (row < r) C[row * cwidth + cb] = double(acc[t]);
            }
        }
    }
    if (lane == 0 && used) atomicAdd(a.npairs, used);
    if (a.c_norm) {
        __syncthreads();  // every wave's columns of the row are written
        row_block_norms<double, double, false>(nullptr, a.c_vals, a.c_voff, c0, c1, wave, 4, lane, a.c_norm);
    }
}

// The same products with the B lookups in LDS (C rows with nbc <= kNumMapCols
// block columns): for every A slot k the work-group writes (slot tag, B slot)
// of each kept block of B's row k into a column-indexed map, so a lane finds
// B_kj with one LDS read instead of a binary search over B's row; tags make a
// clearing pass unnecessary.  Each wave keeps kMfmaStrips 16-column strips of
// the panel per pass (strip s of pass p: columns 16 (4 (p S + t) + wave) ..),
// so one map serves 4 S strips.  Products and their order per strip are those
// of k_numeric_mfma_bf16.
constexpr int kMfmaStrips = 4;  // measured on McWeeny products: 4 beats 2 (10 %)

__global__ __launch_bounds__(256) void k_numeric_mfma_bf16_map(MfmaArgs a, int nbc) {
    extern __shared__ int2 bmap[];  // [nbc]: {A slot tag, B slot}
    const int i = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int a0 = a.a_row_ptr[i], a1 = a.a_row_ptr[i + 1];
    const int c0 = a.c_row_ptr[i], c1 = a.c_row_ptr[i + 1];
    if (c0 == c1) return;
    for (int j = tid; j < nbc; j += 256) bmap[j] = make_int2(-1, 0);
    const int r = a.a_rbs[i];
    const int width = a.panel_width[i];
    const int row_count = a1 - a0;
    const double row_eps = a.thr / double(row_count > 1 ? row_count : 1);
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 8;
    constexpr int S = kMfmaStrips;
    const int passes = (width + 64 * S - 1) / (64 * S);  // the same for every wave (barriers below)
    unsigned long long used = 0;
    for (int p = 0; p < passes; ++p) {
        int slot[S], jcol[S], cb[S], cwidth[S];
#pragma unroll
        for (int t = 0; t < S; ++t) {
            const int pc = 16 * (4 * (p * S + t) + wave) + lr;
            slot[t] = -1;
            jcol[t] = 0;
            cb[t] = 0;
            cwidth[t] = 1;
            if (pc < width) {
                slot[t] = slot_of_panel_col(a.c_colstart, c0, c1, pc);
                jcol[t] = a.c_col_ind[slot[t]];
                cb[t] = pc - a.c_colstart[slot[t]];
                cwidth[t] = a.c_cbs[jcol[t]];
            }
        }
        f32x4 acc[S];
#pragma unroll
        for (int t = 0; t < S; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int ka = a0; ka < a1; ++ka) {
            const int k = a.a_col_ind[ka];
            const int m = a.a_cbs[k];
            const double na = a.a_norm[ka];
            const double* A = a.a_vals + a.a_voff[ka];
            __syncthreads();  // every lookup of the previous A slot is done
            for (int kb = a.b_row_ptr[k] + tid; kb < a.b_row_ptr[k + 1]; kb += 256)
                if (!(na * a.b_norm[kb] < row_eps)) bmap[a.b_col_ind[kb]] = make_int2(ka, kb);
            __syncthreads();
            for (int kk = 0; kk < m; kk += 32) {
                bf16x8 av;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int q = kk + lk + e;
                    av[e] = (lr < r && q < m) ? to_bf16(A[lr * m + q]) : short(0);
                }
#pragma unroll
                for (int t = 0; t < S; ++t) {
                    const double* Bcol = nullptr;
                    if (slot[t] >= 0) {
                        const int2 e = bmap[jcol[t]];
                        if (e.x == ka) Bcol = a.b_vals + a.b_voff[e.y] + cb[t];
                    }
                    // one lane per (A slot, C block): the block's first column in lane group 0
                    if (kk == 0) used += __popcll(__ballot(Bcol != nullptr && cb[t] == 0 && lane < 16));
                    // a strip with no B block of this A slot would add exact
                    // zeros: the wave skips its loads and product
                    if (__ballot(Bcol != nullptr) == 0) continue;
                    bf16x8 bv;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int q = kk + lk + e;
                        bv[e] = (Bcol && q < m) ? to_bf16(Bcol[int64_t(q) * cwidth[t]]) : short(0);
                    }
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < S; ++t) {
            if (slot[t] >= 0) {
                double* C = a.c_vals + a.c_voff[slot[t]];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = (lane >> 4) * 4 + e;
                    if (row < r) C[row * cwidth[t] + cb[t]] = double(acc[t][e]);
                }
            }
        }
    }
    if (lane == 0 && used) atomicAdd(a.npairs, used);
    if (a.c_norm) {
        __syncthreads();  // every wave's strips of the row are written
        row_block_norms<double, double, false>(nullptr, a.c_vals, a.c_voff, c0, c1, wave, 4, lane, a.c_norm);
    }
}

// ------------------------------------------------------------ dense-panel numeric
// Products whose C is dense enough (the late McWeeny iterations fill P to
// most of its block pattern) are computed panel by panel from dense copies
// of A and B (zeros where no block is stored).  A work-group owns
// kPanelRows consecutive block rows of C, one per wave, and a window of
// 64 * CPL scalar columns (CPL per lane).  For every inner block k it stages
// the m_k rows of B over the window in LDS -- once for all its waves, so B
// is read once per row tile rather than once per C row -- and each wave adds
// A_ik (staged per wave) times that panel into registers.  A column takes a
// product only when C's block (i, j) is in the symbolic pattern, B_kj is
// stored, and the numeric rule keeps the pair (||A_ik|| ||B_kj|| >= thr /
// nnz_row(A_i), ops/spmm/vbcsr.hpp:371-387); every C element receives its
// products in k order and within a product in inner-index order, as in the
// sparse kernels.  C is written once, and each block's squared Frobenius
// norm is accumulated from the same registers for the filter.
constexpr int kPanelWaves = 8;  // waves per work-group
// block rows per wave: 2 for FP32 products, 1 for FP64 (two FP64 row sets
// of accumulators and two staging sets exceed 128 VGPRs)
template <typename CT>
constexpr int panel_rows_per_wave() {
    return sizeof(CT) == 4 ? 2 : 1;
}
constexpr double kDense16Fill = 0.1;  // minimum block fill of C for the dense BF16 products
constexpr double kPanelFill = 0.8;  // minimum block fill of C (measured: at 0.74 the sparse kernels are 7 % faster)

template <typename CT>
struct PanelArgs {
    int nbr, nbk, nbc;
    int64_t nca, ncb;          // scalar columns of A (= rows of B) and of B (= of C)
    const CT* ad;              // dense A [n_rows(A)][nca]
    const CT* bd;              // dense B [nca][ncb]
    const double* anm;         // [nbr][nbk] ||A_ik|| or -1 (absent)
    const double* bnm;         // [nbk][nbc] ||B_kj|| or -1
    const int* cslot;          // [nbr][nbc] C's slot of (i, j) or -1
    const int* colblk;         // [ncb] block column of each scalar column
    const int* colb;           // [ncb] column within that block
    const int* a_rbs;          // r_i
    const int* a_roff;         // scalar row offset of block row i
    const double* eps;         // [nbr] thr / nnz_row(A_i)
    const int* k_size;         // [nbk] m_k
    const int* k_off;          // [nbk] scalar offset of inner block k
    const int* c_cbs;          // [nbc] c_j
    const int64_t* c_voff;
    double* c_vals;
    double* c_norm2;           // [nnzb(C)] accumulated squared block norms
    unsigned long long* npairs;
    int n_tiles, n_groups;
};

// acc[a][t] += sum_q A[a][q] * B[q][col_t] over the m staged rows of B, all 8
// accumulator rows (A rows past r are zero), columns masked by ok.
template <typename CT, int CPL>
__device__ __forceinline__ void panel_fma(int m, const CT* __restrict__ bt, const CT* __restrict__ as, int lane,
                                          const bool (&ok)[CPL], CT (&acc)[8][CPL]) {
    constexpr int W = 64 * CPL;
#pragma unroll 1
    for (int q = 0; q < m; ++q) {
        CT bq[CPL];
#pragma unroll
        for (int t = 0; t < CPL; ++t) bq[t] = ok[t] ? bt[q * W + lane * CPL + t] : CT(0);
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const CT av = as[a * m + q];
#pragma unroll
            for (int t = 0; t < CPL; ++t) acc[a][t] = Sc<CT>::fma(av, bq[t], acc[a][t]);
        }
    }
}

// At most 128 VGPRs: two work-groups per CU (four waves per SIMD) keep one
// group's staging loads in flight while the other computes.
template <typename CT, int CPL>
__global__ __launch_bounds__(64 * kPanelWaves) __attribute__((amdgpu_waves_per_eu(4)))
void k_numeric_panel(PanelArgs<CT> g) {
    constexpr int W = 64 * CPL;                     // window columns
    constexpr int NT = 64 * kPanelWaves;            // threads
    constexpr int NE = 8 * W / NT;                  // staged B values per thread per k
    constexpr int RW = panel_rows_per_wave<CT>();
    constexpr int kPanelRows = kPanelWaves * RW;  // block rows per work-group
    __shared__ CT bt[2][8 * W];
    __shared__ double bn[2][W];
    __shared__ CT as[kPanelWaves][2][RW][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // the launch is rounded up to a multiple of 8 for the XCD remapping;
    // groups past the panel grid exit (whole work-groups: no barrier is split)
    const int gid = xcd_chunk(int(gridDim.x));
    if (gid >= g.n_groups) return;
    const int win = gid / g.n_tiles, tile = gid - win * g.n_tiles;
    const int64_t c0 = int64_t(win) * W;
    int row[RW], r[RW], cs[RW][CPL];
    int64_t roff[RW];
    double eps[RW];
    bool row_ok[RW];
    int bcol[CPL], cj[CPL];
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
        const int64_t col = c0 + lane * CPL + t;
        const bool v = col < g.ncb;
        cj[t] = v ? g.colblk[col] : 0;
        bcol[t] = v ? g.colb[col] : 1;
    }
#pragma unroll
    for (int h = 0; h < RW; ++h) {
        const int i = tile * kPanelRows + wave * RW + h;
        row[h] = i;
        row_ok[h] = i < g.nbr;
        r[h] = row_ok[h] ? g.a_rbs[i] : 0;
        roff[h] = row_ok[h] ? g.a_roff[i] : 0;
        eps[h] = row_ok[h] ? g.eps[i] : 0.0;
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const bool v = c0 + lane * CPL + t < g.ncb;
            cs[h][t] = (v && row_ok[h]) ? g.cslot[int64_t(i) * g.nbc + cj[t]] : -1;
        }
    }
    // staging of inner block k: B rows k_off + q (q < m_k) over the window,
    // ||B_kj|| per window column, and each wave's A_ik blocks.  Every load of
    // a stage depends only on k (the dense copies hold zeros where no block
    // is stored), and two stages are in flight while one is computed.
    struct Stage {
        CT sv[NE];
        double snorm;
        double an[RW];
        CT av[RW];
    };
    const int jst = (tid < W && c0 + tid < g.ncb) ? g.colblk[c0 + tid] : -1;  // staged norm column's block
    auto load_stage = [&](int k, Stage& st) {
        const int m = g.k_size[k];
        const int64_t ko = g.k_off[k];
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int e = tid + u * NT;
            const int q = e / W, w = e - (e / W) * W;
            const int64_t col = c0 + w;
            const int qc = q < m ? q : m - 1;
            const CT v = g.bd[(ko + qc) * g.ncb + (col < g.ncb ? col : g.ncb - 1)];
            st.sv[u] = (q < m && col < g.ncb) ? v : CT(0);
        }
        st.snorm = jst >= 0 ? g.bnm[int64_t(k) * g.nbc + jst] : -1.0;
        // A_ik element (a, q) = (lane / m, lane % m); rows past r read as zero
        const int ar = lane / m, aq = lane - ar * m;
#pragma unroll
        for (int h = 0; h < RW; ++h) {
            st.an[h] = row_ok[h] ? g.anm[int64_t(row[h]) * g.nbk + k] : -1.0;
            const int arc = ar < r[h] ? ar : (r[h] > 0 ? r[h] - 1 : 0);
            const CT v = g.ad[(roff[h] + arc) * g.nca + ko + (aq < m ? aq : m - 1)];
            st.av[h] = ar < r[h] ? v : CT(0);
        }
    };
    auto store_stage = [&](int buf, const Stage& st) {
#pragma unroll
        for (int u = 0; u < NE; ++u) bt[buf][tid + u * NT] = st.sv[u];
        if (tid < W) bn[buf][tid] = st.snorm;
#pragma unroll
        for (int h = 0; h < RW; ++h) as[wave][buf][h][lane] = st.av[h];
    };
    CT acc[RW][8][CPL];
#pragma unroll
    for (int h = 0; h < RW; ++h)
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int t = 0; t < CPL; ++t) acc[h][a][t] = CT(0);
    unsigned long long used = 0;
    auto compute = [&](int k, int buf, const double (&an)[RW]) {
        const int m = g.k_size[k];
#pragma unroll
        for (int h = 0; h < RW; ++h) {
            const double na = an[h];
            if (!(row_ok[h] && na >= 0.0)) continue;
            bool ok[CPL];
#pragma unroll
            for (int t = 0; t < CPL; ++t) {
                const double nb = bn[buf][lane * CPL + t];
                ok[t] = cs[h][t] >= 0 && nb >= 0.0 && !(na * nb < eps[h]);
                used += (ok[t] && bcol[t] == 0) ? 1ull : 0ull;
            }
            panel_fma<CT, CPL>(m, bt[buf], as[wave][buf][h], lane, ok, acc[h]);
        }
    };
    Stage s0, s1;
    double an0[RW], an1[RW];
    const int nk = g.nbk;
    load_stage(0, s0);
    if (nk > 1) load_stage(1, s1);
    store_stage(0, s0);
#pragma unroll
    for (int h = 0; h < RW; ++h) an0[h] = s0.an[h];
    __syncthreads();
    for (int k = 0; k < nk; k += 2) {
        // buffer 0 holds stage k; s1 (stage k+1) is in flight
        if (k + 2 < nk) load_stage(k + 2, s0);
        compute(k, 0, an0);
        if (k + 1 >= nk) break;
        store_stage(1, s1);
#pragma unroll
        for (int h = 0; h < RW; ++h) an1[h] = s1.an[h];
        __syncthreads();
        if (k + 3 < nk) load_stage(k + 3, s1);
        compute(k + 1, 1, an1);
        if (k + 2 >= nk) break;
        store_stage(0, s0);
#pragma unroll
        for (int h = 0; h < RW; ++h) an0[h] = s0.an[h];
        __syncthreads();
    }
    // C values once, and the blocks' squared norms
#pragma unroll
    for (int h = 0; h < RW; ++h) {
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            if (cs[h][t] < 0) continue;
            const int c = g.c_cbs[cj[t]];
            double* cv = g.c_vals + g.c_voff[cs[h][t]] + bcol[t];
            double n2 = 0.0;
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                if (a < r[h]) {
                    const double v = double(acc[h][a][t]);
                    cv[int64_t(a) * c] = v;
                    n2 += v * v;
                }
            }
            atomicAdd(&g.c_norm2[cs[h][t]], n2);
        }
    }
    for (int s2 = 32; s2 >= 1; s2 >>= 1) used += __shfl_xor(used, s2, 64);
    if (lane == 0 && used) atomicAdd(g.npairs, used);
}

// Storage conversions of the dense copies: FP64, FP32, or BF16 bits
// (round-to-nearest-even of the FP32 value).
template <typename CT>
__device__ inline CT to_dense(double x) {
    if constexpr (std::is_same_v<CT, short>) return to_bf16(x);
    else return CT(x);
}

// Dense copy of a block matrix: out[(roff_i + a) * ld + coff_j + b] = block
// (i, j) element (a, b), or out[(coff_j + b) * ld + roff_i + a] when
// transposed; one wavefront per stored block (out zeroed first).
template <typename T, typename CT>
__global__ __launch_bounds__(256) void k_densify(const T* __restrict__ vals, const int64_t* __restrict__ voff,
                                                 const int* __restrict__ srow, const int* __restrict__ col_ind,
                                                 const int* __restrict__ roff, const int* __restrict__ coff,
                                                 const int* __restrict__ cbs, int64_t nnzb, int64_t ld,
                                                 int transposed, CT* __restrict__ out) {
    const int64_t k = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= nnzb) return;
    const int i = srow[k], j = col_ind[k], c = cbs[j];
    const int64_t n = voff[k + 1] - voff[k];
    const int r = int(n / c);
    for (int64_t e = lane; e < n; e += 64) {
        // consecutive lanes write consecutive output elements: along the
        // block's row (row-major output) or down its column (transposed)
        int a, b;
        if (transposed) {
            b = int(e / r);
            a = int(e - int64_t(b) * r);
        } else {
            a = int(e / c);
            b = int(e - int64_t(a) * c);
        }
        const int64_t row = int64_t(roff[i]) + a, col = int64_t(coff[j]) + b;
        out[transposed ? col * ld + row : row * ld + col] = to_dense<CT>(double(vals[voff[k] + int64_t(a) * c + b]));
    }
}

// srow[k] = block row of slot k.  One wavefront per block row.
__global__ __launch_bounds__(256) void k_slot_rows(const int* __restrict__ row_ptr, int nbr, int* __restrict__ srow) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nbr) return;
    for (int k = row_ptr[i] + lane; k < row_ptr[i + 1]; k += 64) srow[k] = i;
}

// ------------------------------------------------------------ BF16 dense products
// C (FP32, [Mp][Np]) = A (BF16, [Mp][Kp]) times B^T (BF16, [Np][Kp]): both
// operands k-contiguous, every extent padded with zeros to the tile sizes, so
// the kernel has no bounds checks.  A work-group computes a kGemmTM x
// kGemmTN tile of C with one wave per 64 x 64 part, each as 4 x 4 tiles of
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation).  Per k-step of kGemmK the
// group stages A's and B's slices in LDS (row pitch kGemmK + 8 elements so
// the 16 rows an operand read touches fall on different banks); the next
// step's slices are loaded into registers while the current one is
// multiplied, and one barrier separates the steps.  Measured on a 30k x 30k
// x 30k product: 128 x 128 tiles with k-steps of 32, 114 ms; 256 x 256, 59
// ms (half the operand bytes per flop); k-steps of 64, 45 ms (half the
// barriers); grouped tile order, 43.6 ms (1.26 PFLOP/s).  Larger wave parts
// (fewer waves, fewer LDS reads per MFMA) were slower: 128 x 64, 53.6 ms;
// 64 x 128, 49.5 ms; 128 x 128, 70.7 ms.
#ifndef VBH_GEMM_TM
#define VBH_GEMM_TM 256
#endif
#ifndef VBH_GEMM_TN
#define VBH_GEMM_TN 256
#endif
constexpr int kGemmTM = VBH_GEMM_TM;
constexpr int kGemmTN = VBH_GEMM_TN;
#ifndef VBH_GEMM_K
#define VBH_GEMM_K 64
#endif
#ifndef VBH_GEMM_GROUP
#define VBH_GEMM_GROUP 4
#endif
constexpr int kGemmK = VBH_GEMM_K;
constexpr int kGemmPitch = kGemmK + 8;
constexpr int kGemmGroup = VBH_GEMM_GROUP;
#ifndef VBH_GEMM_WTM
#define VBH_GEMM_WTM 64
#endif
#ifndef VBH_GEMM_WTN
#define VBH_GEMM_WTN 64
#endif
constexpr int kGemmWTM = VBH_GEMM_WTM, kGemmWTN = VBH_GEMM_WTN;  // one wave's part of the tile
constexpr int kGemmFI = kGemmWTM / 16, kGemmFJ = kGemmWTN / 16;  // its 16 x 16 MFMA tiles
constexpr int kGemmWaves = (kGemmTM / kGemmWTM) * (kGemmTN / kGemmWTN);
static_assert(kGemmTM % kGemmWTM == 0 && kGemmTN % kGemmWTN == 0 && kGemmWaves <= 16, "work-group tile of wave tiles");

__global__ __launch_bounds__(64 * kGemmWaves) void k_gemm_bf16_nt(const short* __restrict__ A,
                                                                  const short* __restrict__ BT,
                                                                  float* __restrict__ C, int64_t Kp, int64_t ldc,
                                                                  int tiles_n, int n_tiles) {
    __shared__ __attribute__((aligned(16))) short as[2][kGemmTM * kGemmPitch];
    __shared__ __attribute__((aligned(16))) short bs[2][kGemmTN * kGemmPitch];
    // the launch is rounded up to a multiple of 8 for the XCD remapping
    const int gid = xcd_chunk(int(gridDim.x));
    if (gid >= n_tiles) return;
    // consecutive tiles walk kGemmGroup tile rows before the next column, so
    // the work-groups of one XCD share A's and B's slices in its L2
    const int tiles_m = n_tiles / tiles_n;
    const int grp = gid / (kGemmGroup * tiles_n), q = gid - grp * (kGemmGroup * tiles_n);
    const int gm = tiles_m - grp * kGemmGroup < kGemmGroup ? tiles_m - grp * kGemmGroup : kGemmGroup;
    const int tm = grp * kGemmGroup + q % gm, tn = q / gm;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / (kGemmTN / kGemmWTN), wn = wave % (kGemmTN / kGemmWTN);
    const int64_t m0 = int64_t(tm) * kGemmTM, n0 = int64_t(tn) * kGemmTN;
    // staging: RC 16-byte chunks per operand row and k-step (chunk c: row
    // c / RC, k-offset 8 * (c % RC)), CA / CB per thread
    constexpr int NT = 64 * kGemmWaves, RC = kGemmK / 8;
    constexpr int CA = kGemmTM * RC / NT, CB = kGemmTN * RC / NT;
    static_assert(CA * NT == kGemmTM * RC && CB * NT == kGemmTN * RC, "whole staging chunks per thread");
    typedef __attribute__((__vector_size__(4 * sizeof(int)))) int i32x4;
    i32x4 ra[CA], rb[CB];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int u = 0; u < CA; ++u) {
            const int c = tid + NT * u;
            ra[u] = *reinterpret_cast<const i32x4*>(A + (m0 + c / RC) * Kp + k0 + (c % RC) * 8);
        }
#pragma unroll
        for (int u = 0; u < CB; ++u) {
            const int c = tid + NT * u;
            rb[u] = *reinterpret_cast<const i32x4*>(BT + (n0 + c / RC) * Kp + k0 + (c % RC) * 8);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < CA; ++u) {
            const int c = tid + NT * u;
            *reinterpret_cast<i32x4*>(&as[buf][(c / RC) * kGemmPitch + (c % RC) * 8]) = ra[u];
        }
#pragma unroll
        for (int u = 0; u < CB; ++u) {
            const int c = tid + NT * u;
            *reinterpret_cast<i32x4*>(&bs[buf][(c / RC) * kGemmPitch + (c % RC) * 8]) = rb[u];
        }
    };
    f32x4 acc[kGemmFI][kGemmFJ];
#pragma unroll
    for (int i = 0; i < kGemmFI; ++i)
#pragma unroll
        for (int j = 0; j < kGemmFJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int lr = lane & 15, lk = (lane >> 4) * 8;
    const int nsteps = int(Kp / kGemmK);
    load(0);
    store(0);
    __syncthreads();
    for (int st = 0; st < nsteps; ++st) {
        const int buf = st & 1;
        if (st + 1 < nsteps) load(int64_t(st + 1) * kGemmK);
#pragma unroll
        for (int kk = 0; kk < kGemmK; kk += 32) {
            bf16x8 av[kGemmFI], bv[kGemmFJ];
#pragma unroll
            for (int i = 0; i < kGemmFI; ++i)
                av[i] = *reinterpret_cast<const bf16x8*>(&as[buf][(wm * kGemmWTM + i * 16 + lr) * kGemmPitch + kk + lk]);
#pragma unroll
            for (int j = 0; j < kGemmFJ; ++j)
                bv[j] = *reinterpret_cast<const bf16x8*>(&bs[buf][(wn * kGemmWTN + j * 16 + lr) * kGemmPitch + kk + lk]);
#pragma unroll
            for (int i = 0; i < kGemmFI; ++i)
#pragma unroll
                for (int j = 0; j < kGemmFJ; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        if (st + 1 < nsteps) store(buf ^ 1);
        __syncthreads();
    }
    // result lane l: column l % 16, rows 4 (l / 16) .. +3 of each 16 x 16 tile
#pragma unroll
    for (int i = 0; i < kGemmFI; ++i)
#pragma unroll
        for (int j = 0; j < kGemmFJ; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t row = m0 + wm * kGemmWTM + i * 16 + (lane >> 4) * 4 + e;
                const int64_t col = n0 + wn * kGemmWTN + j * 16 + lr;
                C[row * ldc + col] = acc[i][j][e];
            }
}

// C's blocks (its symbolic pattern) from the dense FP32 product, widened to
// FP64, with each block's squared Frobenius norm; one wavefront per block.
__global__ __launch_bounds__(256) void k_scatter_dense(const float* __restrict__ Cd, int64_t ldc,
                                                       const int* __restrict__ srow, const int* __restrict__ col_ind,
                                                       const int64_t* __restrict__ voff, const int* __restrict__ roff,
                                                       const int* __restrict__ coff, const int* __restrict__ cbs,
                                                       int64_t nnzb, double* __restrict__ vals,
                                                       double* __restrict__ norm2) {
    const int64_t k = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= nnzb) return;
    const int lo = srow[k];
    const int j = col_ind[k], c = cbs[j];
    const int64_t n = voff[k + 1] - voff[k];
    double s = 0.0;
    for (int64_t e = lane; e < n; e += 64) {
        const int a = int(e / c), b = int(e - int64_t(a) * c);
        const double v = double(Cd[(int64_t(roff[lo]) + a) * ldc + coff[j] + b]);
        vals[voff[k] + e] = v;
        s += v * v;
    }
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if (lane == 0) norm2[k] = s;
}

// map[i * ld + col_ind[k]] = val[k] (or k itself when val is null), for every
// stored block k of row i.
template <typename V>
__global__ void k_block_map(const int* __restrict__ row_ptr, const int* __restrict__ col_ind, int nbr,
                            const V* __restrict__ val, int64_t ld, V* __restrict__ map) {
    const int i = blockIdx.x;
    if (i >= nbr) return;
    for (int k = row_ptr[i] + threadIdx.x; k < row_ptr[i + 1]; k += blockDim.x)
        map[int64_t(i) * ld + col_ind[k]] = val ? val[k] : V(k);
}

__global__ void k_fill_f64(double* __restrict__ p, int64_t n, double v) {
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x)
        p[e] = v;
}

static inline unsigned nblk4(int64_t n) { return unsigned((n + 3) / 4); }

// Panel offsets of a pattern's slots (cs[k] = scalar columns of the row's
// blocks before slot k, from meta[k].y >> 8) and each row's panel width
// (w[i]).  One wavefront per row.
__global__ __launch_bounds__(256) void k_panel_cols(int nbr, const int* __restrict__ row_ptr,
                                                    const int2* __restrict__ meta, int* __restrict__ cs,
                                                    int* __restrict__ w) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nbr) return;
    const int kb = row_ptr[i], ke = row_ptr[i + 1];
    for (int k = kb + lane; k < ke; k += 64) cs[k] = meta[k].y >> 8;
    if (lane == 0) w[i] = ke > kb ? (meta[ke - 1].y >> 8) + (meta[ke - 1].y & 255) : 0;
}

// Candidate bound of every row of A B: the sum of the lengths of the B rows
// that row i of A reaches.  One wavefront per row.
__global__ __launch_bounds__(256) void k_row_bounds(int nbr, const int* __restrict__ a_row_ptr,
                                                    const int* __restrict__ a_col_ind,
                                                    const int* __restrict__ b_row_ptr, int64_t* __restrict__ bound) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nbr) return;
    int64_t b = 0;
    for (int ka = a_row_ptr[i] + lane; ka < a_row_ptr[i + 1]; ka += 64) {
        const int k = a_col_ind[ka];
        b += b_row_ptr[k + 1] - b_row_ptr[k];
    }
    for (int d = 32; d >= 1; d >>= 1) b += __shfl_xor(b, d, 64);
    if (lane == 0) bound[i] = b;
}

}  // namespace vbh

using namespace vbh;

// C = A B, or its upper block triangle (block columns >= the row) when upper.
static void spgemm_impl(vbh_matrix_t A, vbh_matrix_t B, double thr, int numeric, int upper, vbh_matrix_t* out,
                        int64_t* stats, void* stream) {
    {
        VBH_REQUIRE(A && B && out, "null argument");
        VBH_REQUIRE(!upper || (A->nbr == B->nbc), "the upper-triangle product needs a square result without ghosts");
        VBH_REQUIRE(A->dtype == B->dtype, "A and B must have the same dtype");
        VBH_REQUIRE(A->nbc == B->nbr, "spmm: A's local columns must be B's block rows");
        for (int k = 0; k < B->nbr; ++k)
            VBH_REQUIRE(A->h_cbs[k] == B->h_rbs[k], "spmm: inner block sizes differ");
        VBH_REQUIRE(numeric >= 0 && numeric <= 2, "numeric must be 0, 1 or 2");
        VBH_REQUIRE(!(numeric == 2 && A->dtype != VBH_F64), "BF16 MFMA numeric supports real matrices");
        VBH_REQUIRE(!(numeric == 2 && A->max_block > 16), "BF16 MFMA numeric supports blocks up to 16");
        hipStream_t s = (hipStream_t)stream;
        const int nbr = A->nbr;
        // ---- norms
        DevList tmp;  // every temporary of the call, freed on return or error
        double* an = tmp.own((double*)dalloc(sizeof(double) * std::max<int64_t>(A->nnzb, 1)));
        double* bn = tmp.own((double*)dalloc(sizeof(double) * std::max<int64_t>(B->nnzb, 1)));
        auto run_norms = [&](vbh_matrix* M, double* o) {
            if (M->nnzb == 0) return;
            if (M->dtype == VBH_F64)
                hipLaunchKernelGGL(k_norms<double>, dim3(nblk4(M->nnzb)), dim3(256), 0, s, (const double*)M->vals,
                                   M->voff, M->nnzb, o);
            else
                hipLaunchKernelGGL(k_norms<c128>, dim3(nblk4(M->nnzb)), dim3(256), 0, s, (const c128*)M->vals,
                                   M->voff, M->nnzb, o);
            VBH_CHECK_HIP(hipGetLastError());
        };
        run_norms(A, an);
        run_norms(B, bn);
        // ---- symbolic sizing (host)
        std::vector<int64_t> bound(nbr, 0), out_off(nbr + 1, 0);
        std::vector<int> tsize(nbr, 1);
        if (nbr > 0) {
            int64_t* d_bound = tmp.own((int64_t*)dalloc(sizeof(int64_t) * size_t(nbr)));
            hipLaunchKernelGGL(k_row_bounds, dim3(nblk4(nbr)), dim3(256), 0, s, nbr, A->row_ptr, A->col_ind,
                               B->row_ptr, d_bound);
            VBH_CHECK_HIP(hipGetLastError());
            VBH_CHECK_HIP(hipMemcpyAsync(bound.data(), d_bound, sizeof(int64_t) * size_t(nbr), hipMemcpyDeviceToHost,
                                         s));
            VBH_CHECK_HIP(hipStreamSynchronize(s));
        }
        rows_parallel(nbr, A->h_row_ptr.data(), [&](int r0, int r1) {
            for (int i = r0; i < r1; ++i) {
                const int64_t b = bound[i];
                const int64_t cap = std::min<int64_t>(b, B->nbc);
                out_off[i + 1] = cap;
                int64_t S = 1;
                while (S < 2 * cap + 1) S <<= 1;
                VBH_REQUIRE(S <= (int64_t(1) << 30), "symbolic hash table too large");
                tsize[i] = int(S);
            }
        });
        for (int i = 0; i < nbr; ++i) out_off[i + 1] += out_off[i];
        int* d_tsize = tmp.own(upload(tsize, s));
        int64_t* d_out_off = tmp.own(upload(out_off, s));
        int* d_out_cols = tmp.own((int*)dalloc(sizeof(int) * std::max<int64_t>(out_off[nbr], 1)));
        int* d_count = tmp.own((int*)dalloc(sizeof(int) * std::max(nbr, 1)));
        int64_t* d_pairs = tmp.own((int64_t*)dalloc(sizeof(int64_t) * std::max(nbr, 1)));
        VBH_CHECK_HIP(hipMemsetAsync(d_pairs, 0, sizeof(int64_t) * std::max(nbr, 1), s));
        VBH_CHECK_HIP(hipMemsetAsync(d_count, 0, sizeof(int) * std::max(nbr, 1), s));
        SymArgs sa{};
        sa.a_row_ptr = A->row_ptr;
        sa.a_col_ind = A->col_ind;
        sa.a_norm = an;
        sa.b_row_ptr = B->row_ptr;
        sa.b_col_ind = B->col_ind;
        sa.b_norm = bn;
        sa.tsize = d_tsize;
        sa.out_off = d_out_off;
        sa.out_cols = d_out_cols;
        sa.out_count = d_count;
        sa.pairs = d_pairs;
        sa.thr = thr;
        sa.upper = upper;
        // Dense symbolic phase when the candidate pattern is dense enough
        // (the bound of every row capped at nbc covers >= 30 % of the
        // matrix) and the three dense arrays fit half the free memory:
        // estimates by one product of the block-norm matrices.
        bool dense_sym = false;
        {
            double cover = 0.0;
            for (int i = 0; i < nbr; ++i) cover += double(std::min<int64_t>(bound[i], B->nbc));
            const int64_t Mp = (nbr + kEstTile - 1) / kEstTile * kEstTile;
            const int64_t Np = (B->nbc + kEstTile - 1) / kEstTile * kEstTile;
            const int64_t Kp = (A->nbc + kEstK - 1) / kEstK * kEstK;
            if (thr >= 0.0 && nbr > 0 && B->nbc > 0 && cover >= 0.3 * double(nbr) * double(B->nbc) &&
                (Mp / kEstTile) * (Np / kEstTile) <= INT32_MAX - 8) {
                const size_t need = sizeof(double) * size_t(Mp * Kp + Kp * Np + Mp * Np);
                size_t free_b = 0, total_b = 0;
                VBH_CHECK_HIP(hipMemGetInfo(&free_b, &total_b));
                dense_sym = need <= free_b / 2;
            }
            if (dense_sym) {
                DevList dt;
                double* na = dt.own((double*)dalloc(sizeof(double) * size_t(Mp * Kp)));
                double* nbm = dt.own((double*)dalloc(sizeof(double) * size_t(Kp * Np)));
                double* est = dt.own((double*)dalloc(sizeof(double) * size_t(Mp * Np)));
                VBH_CHECK_HIP(hipMemsetAsync(na, 0, sizeof(double) * size_t(Mp * Kp), s));
                VBH_CHECK_HIP(hipMemsetAsync(nbm, 0, sizeof(double) * size_t(Kp * Np), s));
                hipLaunchKernelGGL(k_block_map<double>, dim3(nbr), dim3(256), 0, s, A->row_ptr, A->col_ind, nbr,
                                   (const double*)an, Kp, na);
                if (B->nbr > 0)
                    hipLaunchKernelGGL(k_block_map<double>, dim3(B->nbr), dim3(256), 0, s, B->row_ptr, B->col_ind,
                                       B->nbr, (const double*)bn, Np, nbm);
                const int tiles_n = int(Np / kEstTile), n_tiles = int(Mp / kEstTile) * tiles_n;
                hipLaunchKernelGGL(k_gemm_f64, dim3((n_tiles + 7) & ~7), dim3(256), 0, s, na, nbm, est, Kp, Np,
                                   tiles_n, n_tiles);
                hipLaunchKernelGGL(k_est_compact, dim3(nbr), dim3(256), 0, s, est, Np, B->nbc, thr, upper,
                                   d_out_off, d_out_cols, d_count);
                VBH_CHECK_HIP(hipGetLastError());
                // candidate pairs per row: the bound (every B slot of every A slot)
                VBH_CHECK_HIP(hipMemcpyAsync(d_pairs, bound.data(), sizeof(int64_t) * size_t(nbr),
                                             hipMemcpyHostToDevice, s));
                VBH_CHECK_HIP(hipStreamSynchronize(s));  // before the temporaries are freed
            }
        }
        // LDS classes: <=256, <=1024, <=4096 slots; wider rows use global tables
        const int classes[3] = {256, 1024, 4096};
        std::vector<int> rows_cls[4];
        for (int i = 0; i < nbr && !dense_sym; ++i) {
            if (bound[i] == 0) continue;
            int c = 3;
            for (int q = 0; q < 3; ++q)
                if (tsize[i] <= classes[q]) { c = q; break; }
            rows_cls[c].push_back(i);
        }
        for (int q = 0; q < 3; ++q) {
            if (rows_cls[q].empty()) continue;
            int* d_rows = tmp.own(upload(rows_cls[q], s));
            sa.rows = d_rows;
            sa.nrows = int(rows_cls[q].size());
            const size_t shm = size_t(classes[q]) * (sizeof(double) + sizeof(int));
            hipLaunchKernelGGL(k_symbolic<true>, dim3(sa.nrows), dim3(256), shm, s, sa);
            VBH_CHECK_HIP(hipGetLastError());
        }
        if (!rows_cls[3].empty() && B->nbc <= kNumMapCols && thr >= 0.0) {
            int* d_rows = tmp.own(upload(rows_cls[3], s));
            sa.rows = d_rows;
            sa.nrows = int(rows_cls[3].size());
            hipLaunchKernelGGL(k_symbolic_dense, dim3(sa.nrows), dim3(256), sizeof(double) * size_t(B->nbc), s, sa,
                               B->nbc);
            VBH_CHECK_HIP(hipGetLastError());
        } else if (!rows_cls[3].empty()) {
            // global tables, processed in batches bounded to ~1 GiB of tables
            const auto& rows = rows_cls[3];
            size_t pos = 0;
            while (pos < rows.size()) {
                std::vector<int64_t> goff(1, 0);
                std::vector<int> batch;
                while (pos < rows.size() && (batch.empty() || goff.back() + tsize[rows[pos]] <= (int64_t(1) << 26))) {
                    batch.push_back(rows[pos]);
                    goff.push_back(goff.back() + tsize[rows[pos]]);
                    ++pos;
                }
                DevList bt;  // this batch's tables
                int* d_rows = bt.own(upload(batch, s));
                int64_t* d_goff = bt.own(upload(goff, s));
                int* gk = bt.own((int*)dalloc(sizeof(int) * goff.back()));
                double* gv = bt.own((double*)dalloc(sizeof(double) * goff.back()));
                sa.rows = d_rows;
                sa.nrows = int(batch.size());
                sa.g_keys = gk;
                sa.g_vals = gv;
                sa.g_off = d_goff;
                hipLaunchKernelGGL(k_symbolic<false>, dim3(sa.nrows), dim3(256), 0, s, sa);
                VBH_CHECK_HIP(hipGetLastError());
                VBH_CHECK_HIP(hipStreamSynchronize(s));
            }
        }
        // ---- pattern of C: the per-row kept columns are compacted on the
        // device (row i's count at crp[i]), so only they cross to the host,
        // which sorts the rows the hash tables left unsorted
        std::vector<int> counts(nbr);
        std::vector<int64_t> pairs(nbr);
        VBH_CHECK_HIP(hipMemcpyAsync(counts.data(), d_count, sizeof(int) * nbr, hipMemcpyDeviceToHost, s));
        VBH_CHECK_HIP(hipMemcpyAsync(pairs.data(), d_pairs, sizeof(int64_t) * nbr, hipMemcpyDeviceToHost, s));
        VBH_CHECK_HIP(hipStreamSynchronize(s));
        std::vector<int> crp(nbr + 1, 0);
        for (int i = 0; i < nbr; ++i) {
            VBH_REQUIRE(int64_t(crp[i]) + counts[i] <= int64_t(INT32_MAX), "SpGEMM result exceeds 2^31 blocks");
            crp[i + 1] = crp[i] + counts[i];
        }
        const int64_t total = crp[nbr];
        DevList ct;
        int* d_cci = ct.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(total, 1))));
        if (total > 0) {
            int* d_crp = ct.own(upload(crp, s));
            hipLaunchKernelGGL(k_compact_cols, dim3(nbr), dim3(256), 0, s, d_out_cols, d_out_off, d_crp, d_cci);
            VBH_CHECK_HIP(hipGetLastError());
            sort_rows_dev(nbr, crp.data(), d_crp, d_cci, s);
        }
        MatrixGuard C(build_matrix_dev(nbr, B->nbc, A->h_rbs.data(), B->h_cbs.data(), crp.data(), d_cci, A->dtype,
                                       s));
        unsigned long long* d_np = tmp.own((unsigned long long*)dalloc(sizeof(unsigned long long)));
        vbh_matrix* Cm = C.m;
        VBH_CHECK_HIP(hipMemsetAsync(d_np, 0, sizeof(unsigned long long), s));
        // ---- numeric
        double* d_cnorm2 = nullptr;  // squared C block norms from the panel kernel (fused filter)
        double* d_cnorm = nullptr;   // C block norms from the sparse kernels (fused filter)
        const double fill = double(Cm->nnzb) / (double(std::max(nbr, 1)) * double(std::max(Cm->nbc, 1)));
        bool panel = numeric != 2 && Cm->nnzb > 0 && A->dtype == VBH_F64 && A->max_block <= 8 &&
                     B->max_block <= 8 && fill >= kPanelFill;
        if (panel) {
            // dense copies of A and B, block maps of norms and C slots: within
            // half the free device memory, else the sparse kernels
            const size_t ct = numeric == 1 ? sizeof(float) : sizeof(double);
            const size_t need = ct * (size_t(A->n_rows) * size_t(A->n_cols) + size_t(B->n_rows) * size_t(B->n_cols)) +
                                sizeof(double) * (size_t(nbr) * size_t(A->nbc) + size_t(B->nbr) * size_t(B->nbc)) +
                                sizeof(int) * size_t(nbr) * size_t(Cm->nbc);
            size_t free_b = 0, total_b = 0;
            VBH_CHECK_HIP(hipMemGetInfo(&free_b, &total_b));
            panel = need <= free_b / 2;
        }
        // BF16 products of dense-enough results: one dense BF16 matrix-core
        // GEMM (FP32 accumulation) over zero-padded dense copies, then C's
        // pattern blocks and their norms.  Every stored pair is multiplied:
        // the numeric skip rule drops products of norm below thr / nnz_row,
        // under BF16 rounding.
        const int64_t Mp = (A->n_rows + kGemmTM - 1) / kGemmTM * kGemmTM;
        const int64_t Np = (B->n_cols + kGemmTN - 1) / kGemmTN * kGemmTN;
        const int64_t Kp = (A->n_cols + kGemmK - 1) / kGemmK * kGemmK;
        bool dense16 = numeric == 2 && Cm->nnzb > 0 && A->dtype == VBH_F64 && fill >= kDense16Fill &&
                       Mp / kGemmTM * (Np / kGemmTN) <= INT32_MAX - 8;
        if (dense16) {
            const size_t need = size_t(2) * size_t(Mp + Np) * size_t(Kp) + size_t(4) * size_t(Mp) * size_t(Np);
            size_t free_b = 0, total_b = 0;
            VBH_CHECK_HIP(hipMemGetInfo(&free_b, &total_b));
            dense16 = need <= free_b / 2;
        }
        // block row of every slot, for the dense copies and the scatter
        // (one array per matrix, A's reused when B is A)
        auto slot_rows = [&](const vbh_matrix* M) {
            int* d = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(M->nnzb, 1))));
            if (M->nbr > 0)
                hipLaunchKernelGGL(k_slot_rows, dim3(nblk4(M->nbr)), dim3(256), 0, s, M->row_ptr, M->nbr, d);
            VBH_CHECK_HIP(hipGetLastError());
            return d;
        };
        if (dense16) {
            d_cnorm2 = tmp.own((double*)dalloc(sizeof(double) * size_t(Cm->nnzb)));
            const int* a_srow = slot_rows(A);
            const int* b_srow = B == A ? a_srow : slot_rows(B);
            const int* c_srow = slot_rows(Cm);
            DevList dt;
            short* a16 = dt.own((short*)dalloc(sizeof(short) * size_t(Mp) * size_t(Kp)));
            short* b16 = dt.own((short*)dalloc(sizeof(short) * size_t(Np) * size_t(Kp)));
            float* cd = dt.own((float*)dalloc(sizeof(float) * size_t(Mp) * size_t(Np)));
            VBH_CHECK_HIP(hipMemsetAsync(a16, 0, sizeof(short) * size_t(Mp) * size_t(Kp), s));
            VBH_CHECK_HIP(hipMemsetAsync(b16, 0, sizeof(short) * size_t(Np) * size_t(Kp), s));
            if (A->nnzb)
                hipLaunchKernelGGL((k_densify<double, short>), dim3(nblk4(A->nnzb)), dim3(256), 0, s,
                                   (const double*)A->vals, A->voff, a_srow, A->col_ind, A->roff, A->coff, A->cbs,
                                   A->nnzb, Kp, 0, a16);
            if (B->nnzb)
                hipLaunchKernelGGL((k_densify<double, short>), dim3(nblk4(B->nnzb)), dim3(256), 0, s,
                                   (const double*)B->vals, B->voff, b_srow, B->col_ind, B->roff, B->coff, B->cbs,
                                   B->nnzb, Kp, 1, b16);
            VBH_CHECK_HIP(hipGetLastError());
            const int tiles_n = int(Np / kGemmTN);
            const int n_tiles = int(Mp / kGemmTM) * tiles_n;
            hipLaunchKernelGGL(k_gemm_bf16_nt, dim3((n_tiles + 7) & ~7), dim3(64 * kGemmWaves), 0, s, a16, b16, cd,
                               Kp, Np, tiles_n, n_tiles);
            VBH_CHECK_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_scatter_dense, dim3(nblk4(Cm->nnzb)), dim3(256), 0, s, cd, Np, c_srow,
                               Cm->col_ind, Cm->voff, Cm->roff, Cm->coff, Cm->cbs, Cm->nnzb, (double*)Cm->vals,
                               d_cnorm2);
            VBH_CHECK_HIP(hipGetLastError());
            VBH_CHECK_HIP(hipStreamSynchronize(s));  // before the temporaries are freed
        } else if (panel) {
            d_cnorm2 = tmp.own((double*)dalloc(sizeof(double) * size_t(Cm->nnzb)));
            VBH_CHECK_HIP(hipMemsetAsync(d_cnorm2, 0, sizeof(double) * size_t(Cm->nnzb), s));
            auto go = [&](auto* ctag) {
                using CT = std::remove_pointer_t<decltype(ctag)>;
                constexpr int CPL = 2;
                DevList pt;  // the panel's temporaries
                CT* ad = pt.own((CT*)dalloc(sizeof(CT) * size_t(A->n_rows) * size_t(A->n_cols)));
                CT* bd = pt.own((CT*)dalloc(sizeof(CT) * size_t(B->n_rows) * size_t(B->n_cols)));
                VBH_CHECK_HIP(hipMemsetAsync(ad, 0, sizeof(CT) * size_t(A->n_rows) * size_t(A->n_cols), s));
                VBH_CHECK_HIP(hipMemsetAsync(bd, 0, sizeof(CT) * size_t(B->n_rows) * size_t(B->n_cols), s));
                const int* a_srow = slot_rows(A);
                const int* b_srow = B == A ? a_srow : slot_rows(B);
                if (A->nnzb)
                    hipLaunchKernelGGL((k_densify<double, CT>), dim3(nblk4(A->nnzb)), dim3(256), 0, s,
                                       (const double*)A->vals, A->voff, a_srow, A->col_ind, A->roff, A->coff,
                                       A->cbs, A->nnzb, int64_t(A->n_cols), 0, ad);
                if (B->nnzb)
                    hipLaunchKernelGGL((k_densify<double, CT>), dim3(nblk4(B->nnzb)), dim3(256), 0, s,
                                       (const double*)B->vals, B->voff, b_srow, B->col_ind, B->roff, B->coff,
                                       B->cbs, B->nnzb, int64_t(B->n_cols), 0, bd);
                VBH_CHECK_HIP(hipGetLastError());
                double* anm = pt.own((double*)dalloc(sizeof(double) * size_t(nbr) * size_t(A->nbc)));
                double* bnm = pt.own((double*)dalloc(sizeof(double) * size_t(B->nbr) * size_t(B->nbc)));
                int* csl = pt.own((int*)dalloc(sizeof(int) * size_t(nbr) * size_t(Cm->nbc)));
                hipLaunchKernelGGL(k_fill_f64, dim3(2048), dim3(256), 0, s, anm, int64_t(nbr) * A->nbc, -1.0);
                hipLaunchKernelGGL(k_fill_f64, dim3(2048), dim3(256), 0, s, bnm, int64_t(B->nbr) * B->nbc, -1.0);
                VBH_CHECK_HIP(hipMemsetAsync(csl, 0xff, sizeof(int) * size_t(nbr) * size_t(Cm->nbc), s));
                hipLaunchKernelGGL(k_block_map<double>, dim3(nbr), dim3(256), 0, s, A->row_ptr, A->col_ind, nbr,
                                   (const double*)an, int64_t(A->nbc), anm);
                hipLaunchKernelGGL(k_block_map<double>, dim3(B->nbr), dim3(256), 0, s, B->row_ptr, B->col_ind, B->nbr,
                                   (const double*)bn, int64_t(B->nbc), bnm);
                hipLaunchKernelGGL(k_block_map<int>, dim3(nbr), dim3(256), 0, s, Cm->row_ptr, Cm->col_ind, nbr,
                                   (const int*)nullptr, int64_t(Cm->nbc), csl);
                VBH_CHECK_HIP(hipGetLastError());
                std::vector<int> colblk(size_t(B->n_cols)), colb(size_t(B->n_cols));
                for (int j = 0; j < B->nbc; ++j)
                    for (int b = 0; b < B->h_cbs[j]; ++b) {
                        colblk[size_t(B->h_coff[j] + b)] = j;
                        colb[size_t(B->h_coff[j] + b)] = b;
                    }
                std::vector<double> eps(nbr);
                for (int i = 0; i < nbr; ++i) {
                    const int cnt = A->h_row_ptr[i + 1] - A->h_row_ptr[i];
                    eps[i] = thr / double(cnt > 1 ? cnt : 1);
                }
                PanelArgs<CT> pa{};
                pa.nbr = nbr;
                pa.nbk = A->nbc;
                pa.nbc = Cm->nbc;
                pa.nca = A->n_cols;
                pa.ncb = B->n_cols;
                pa.ad = ad;
                pa.bd = bd;
                pa.anm = anm;
                pa.bnm = bnm;
                pa.cslot = csl;
                pa.colblk = pt.own(upload(colblk, s));
                pa.colb = pt.own(upload(colb, s));
                pa.a_rbs = A->rbs;
                pa.a_roff = A->roff;
                pa.eps = pt.own(upload(eps, s));
                pa.k_size = B->rbs;
                pa.k_off = B->roff;
                pa.c_cbs = Cm->cbs;
                pa.c_voff = Cm->voff;
                pa.c_vals = (double*)Cm->vals;
                pa.c_norm2 = d_cnorm2;
                pa.npairs = d_np;
                constexpr int rows_per_group = kPanelWaves * panel_rows_per_wave<CT>();
                pa.n_tiles = (nbr + rows_per_group - 1) / rows_per_group;
                const int64_t nwin = (B->n_cols + 64 * CPL - 1) / (64 * CPL);
                VBH_REQUIRE(int64_t(pa.n_tiles) * nwin <= INT32_MAX - 8, "internal: panel grid too large");
                pa.n_groups = int(pa.n_tiles * nwin);
                hipLaunchKernelGGL((k_numeric_panel<CT, CPL>), dim3((pa.n_groups + 7) & ~7), dim3(64 * kPanelWaves), 0,
                                   s, pa);
                VBH_CHECK_HIP(hipGetLastError());
                VBH_CHECK_HIP(hipStreamSynchronize(s));  // before the temporaries are freed
            };
            if (numeric == 1) go((float*)nullptr);
            else go((double*)nullptr);
        } else if (Cm->nnzb > 0) {
            // the sparse kernels write C's block norms for the filter (fused)
            if (thr > 0.0) d_cnorm = tmp.own((double*)dalloc(sizeof(double) * size_t(Cm->nnzb)));
            // BF16 products of blocks up to 8 x 8 run on the vector ALUs (the
            // owned kernel on BF16-rounded FP32 copies): 16 x 16 x 32 matrix-core
            // tiles would be at most an eighth full.  Larger blocks take the
            // matrix cores.
            const bool bf_owned = numeric == 2 && A->max_block <= 8 && B->max_block <= 8;
            if (numeric == 2 && !bf_owned) {
                // each C slot's first scalar column within its row's panel and
                // the panel widths, from C's slot metadata
                int* d_cs = tmp.own((int*)dalloc(sizeof(int) * size_t(Cm->nnzb)));
                int* d_w = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max(nbr, 1))));
                hipLaunchKernelGGL(k_panel_cols, dim3(nblk4(nbr)), dim3(256), 0, s, nbr, Cm->row_ptr, Cm->meta, d_cs,
                                   d_w);
                VBH_CHECK_HIP(hipGetLastError());
                MfmaArgs ma{A->row_ptr, A->col_ind, A->voff, (const double*)A->vals, an, A->rbs, A->cbs,
                            B->row_ptr, B->col_ind, B->voff, (const double*)B->vals, bn, Cm->row_ptr, Cm->col_ind,
                            Cm->voff, Cm->cbs, d_cs, d_w, (double*)Cm->vals, thr, d_cnorm, d_np};
                if (Cm->nbc <= kNumMapCols)
                    hipLaunchKernelGGL(k_numeric_mfma_bf16_map, dim3(nbr), dim3(256), sizeof(int2) * size_t(Cm->nbc), s,
                                       ma, Cm->nbc);
                else
                    hipLaunchKernelGGL(k_numeric_mfma_bf16, dim3(nbr), dim3(256), 0, s, ma);
                VBH_CHECK_HIP(hipGetLastError());
                VBH_CHECK_HIP(hipStreamSynchronize(s));
            } else {
                auto launch = [&](auto* tag, auto* ctag, auto bf_) {
                    using T = std::remove_pointer_t<decltype(tag)>;
                    using CT = std::remove_pointer_t<decltype(ctag)>;
                    constexpr bool BF = decltype(bf_)::value;
                    NumArgs<T, CT> na{A->row_ptr, A->col_ind, A->voff, (const T*)A->vals, an, A->rbs, A->cbs,
                                      B->row_ptr, B->col_ind, B->voff, (const T*)B->vals, bn, Cm->row_ptr,
                                      Cm->col_ind, Cm->voff, Cm->cbs, (T*)Cm->vals, d_np, thr,
                                      nullptr, nullptr, nullptr, d_cnorm};
                    if (A->max_block <= 8 && B->max_block <= 8) {
                        constexpr bool narrow_products = !std::is_same_v<T, CT>;
                        if constexpr (narrow_products) {
                            // FP32 products from the FP32 value copies (rounded as
                            // the products round them), accumulated in FP32
                            ensure_vals32(A, s);
                            ensure_vals32(B, s);
                            na.a_c = (const CT*)A->vals32;
                            na.b_c = (const CT*)B->vals32;
                            na.c_c = tmp.own((CT*)dalloc(sizeof(CT) * size_t(std::max<int64_t>(Cm->nnz, 1))));
                            VBH_CHECK_HIP(hipMemsetAsync(na.c_c, 0, sizeof(CT) * size_t(Cm->nnz), s));
                        }
                        if (Cm->nbc <= kNumMapCols)
                            hipLaunchKernelGGL((k_numeric_owned<T, CT, true, BF>), dim3(nbr), dim3(64 * kNumWaves),
                                               sizeof(int) * size_t(Cm->nbc), s, na, Cm->nbc);
                        else
                            hipLaunchKernelGGL((k_numeric_owned<T, CT, false, BF>), dim3(nbr), dim3(64 * kNumWaves), 0,
                                               s, na, Cm->nbc);
                        if constexpr (narrow_products) {
                            VBH_CHECK_HIP(hipGetLastError());
                            VBH_CHECK_HIP(hipStreamSynchronize(s));  // before c_c is freed
                        }
                    } else if constexpr (!BF) {
                        hipLaunchKernelGGL((k_numeric<T, CT>), dim3(nbr), dim3(256), 0, s, na);
                    }
                    VBH_CHECK_HIP(hipGetLastError());
                };
                using no_bf = std::false_type;
                if (A->dtype == VBH_F64) {
                    if (numeric == 0) launch((double*)nullptr, (double*)nullptr, no_bf{});
                    else if (numeric == 1) launch((double*)nullptr, (float*)nullptr, no_bf{});
                    else launch((double*)nullptr, (float*)nullptr, std::true_type{});
                } else {
                    if (numeric == 0) launch((c128*)nullptr, (c128*)nullptr, no_bf{});
                    else launch((c128*)nullptr, (c64*)nullptr, no_bf{});
                }
            }
        }
        unsigned long long np = 0;
        VBH_CHECK_HIP(hipMemcpyAsync(&np, d_np, sizeof(np), hipMemcpyDeviceToHost, s));
        VBH_CHECK_HIP(hipStreamSynchronize(s));
        const int64_t before = C.m->nnzb;
        // ---- filter
        vbh_matrix* F = nullptr;
        if (thr > 0.0 && C.m->nnzb > 0) {
            if (d_cnorm2 || d_cnorm) {
                // norms from the numeric phase: no separate pass
                F = d_cnorm2 ? filter_by_norms(C.m, thr, d_cnorm2, true, s) : filter_by_norms(C.m, thr, d_cnorm, false, s);
            } else {
                const int rc = vbh_matrix_filter(C.m, thr, &F, stream);
                if (rc != VBH_OK) throw Error(rc, vbh_last_error());
            }
        }
        if (F) C.reset(F);
        if (stats) {
            int64_t sp = 0;
            for (int64_t p : pairs) sp += p;
            stats[0] = sp;
            stats[1] = dense16 ? sp : int64_t(np);  // dense BF16 products multiply every candidate pair
            stats[2] = before;
            stats[3] = C.m->nnzb;
        }
        *out = C.release();
    }
}

// Hermitian completion of an upper-triangle matrix U: C keeps U's blocks on
// and above the diagonal, gains C_ji = U_ij^H for every strict-upper block, and
// each diagonal block D becomes (D + D^H)/2 (transpose.hpp:139-200,
// block_csr.hpp:1913-1975).  src[k]: U's slot for C's slot k; mode[k]: 0 copy,
// 1 conjugate transpose, 2 Hermitian average.  One thread per C value.
template <typename T>
__global__ void k_complete_hermitian(int64_t nvals, const int64_t* __restrict__ cvoff, int nnzb,
                                     const int* __restrict__ crow, const int* __restrict__ ccol,
                                     const int* __restrict__ rbs, const int64_t* __restrict__ uvoff,
                                     const int* __restrict__ src, const int* __restrict__ mode,
                                     const T* __restrict__ U, T* __restrict__ C) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nvals) return;
    // slot of value e: last k with cvoff[k] <= e
    int lo = 0, hi = nnzb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cvoff[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    const int k = lo;
    const int c = rbs[ccol[k]];
    const int64_t in = e - cvoff[k];
    const int a = int(in / c), b = int(in - int64_t(a) * c);
    const T* u = U + uvoff[src[k]];
    const int m = mode[k];
    if (m == 0) {
        C[e] = u[int64_t(a) * c + b];
    } else if (m == 1) {
        // U block (j, i) is c x r with r = rbs[crow[k]]: element (b, a)
        C[e] = Sc<T>::conj(u[int64_t(b) * rbs[crow[k]] + a]);
    } else {
        const T x = u[int64_t(a) * c + b], y = Sc<T>::conj(u[int64_t(b) * c + a]);
        C[e] = rscale(0.5, x + y);
    }
}

// Pattern of the Hermitian completion, on the device.  Counts: U block
// (j, c) with c > j adds one mirrored block to row c of C (low[c]); row j
// keeps its blocks with c >= j (own[j], a suffix of the sorted row).  One
// wavefront per row of U.
__global__ __launch_bounds__(256) void k_herm_counts(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                     int* __restrict__ low, int* __restrict__ own) {
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= n) return;
    int o = 0;
    for (int k = rp[j] + lane; k < rp[j + 1]; k += 64) {
        const int c = ci[k];
        if (c > j) atomicAdd(&low[c], 1);
        o += c >= j ? 1 : 0;
    }
    for (int d = 32; d >= 1; d >>= 1) o += __shfl_xor(o, d, 64);
    if (lane == 0) own[j] = o;
}

// Mirrored blocks into the lower segments (lseg[fill[c]++] = U slot, any
// order; the segments are sorted afterwards).  One wavefront per row of U.
__global__ __launch_bounds__(256) void k_herm_scatter(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                      int* __restrict__ fill, int* __restrict__ lseg) {
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= n) return;
    for (int k = rp[j] + lane; k < rp[j + 1]; k += 64) {
        const int c = ci[k];
        if (c > j) lseg[atomicAdd(&fill[c], 1)] = k;
    }
}

// Row i of C: its mirrored blocks (sorted U slots of segment lp[i].., so
// ascending source rows; column = the source row, mode 1), then U's blocks
// (i, c >= i) (mode 2 on the diagonal, else 0); src = U slot, crow = i.
// One wavefront per row.
__global__ __launch_bounds__(256) void k_herm_fill(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const int* __restrict__ crp, const int* __restrict__ lp,
                                                   const int* __restrict__ lseg, int* __restrict__ cci,
                                                   int* __restrict__ src, int* __restrict__ mode,
                                                   int* __restrict__ crow) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const int nl = lp[i + 1] - lp[i];
    for (int t = lane; t < nl; t += 64) {
        const int k = lseg[lp[i] + t];
        int lo = 0, hi = n - 1;  // the row of slot k
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (rp[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        const int pos = crp[i] + t;
        cci[pos] = lo;
        src[pos] = k;
        mode[pos] = 1;
        crow[pos] = i;
    }
    const int kb = rp[i + 1] - (crp[i + 1] - crp[i] - nl);  // first block with c >= i
    for (int k = kb + lane; k < rp[i + 1]; k += 64) {
        const int pos = crp[i] + nl + (k - kb);
        const int c = ci[k];
        cci[pos] = c;
        src[pos] = k;
        mode[pos] = c == i ? 2 : 0;
        crow[pos] = i;
    }
}

static void complete_hermitian_impl(vbh_matrix_t U, vbh_matrix_t* out, hipStream_t s) {
    VBH_REQUIRE(U && out, "null argument");
    VBH_REQUIRE(U->nbr == U->nbc, "Hermitian completion needs a square pattern without ghosts");
    const int n = U->nbr;
    for (int i = 0; i < n; ++i)
        VBH_REQUIRE(U->h_rbs[i] == U->h_cbs[i], "Hermitian completion needs equal row and column block sizes");
    // rows of C: mirrored blocks (column j < i, from U's row j) then U's own
    // blocks with j >= i; U blocks below the diagonal are ignored.  The
    // pattern is built on the device; the host receives the row counts.
    std::vector<int> low(n), own(n), crp(n + 1, 0), lp(n + 1, 0);
    DevList tmp;
    int* d_low = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max(n, 1))));
    int* d_own = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max(n, 1))));
    const unsigned nb4 = unsigned((n + 3) / 4);
    if (n > 0) {
        VBH_CHECK_HIP(hipMemsetAsync(d_low, 0, sizeof(int) * size_t(n), s));
        hipLaunchKernelGGL(k_herm_counts, dim3(nb4), dim3(256), 0, s, n, U->row_ptr, U->col_ind, d_low, d_own);
        VBH_CHECK_HIP(hipGetLastError());
        VBH_CHECK_HIP(hipMemcpyAsync(low.data(), d_low, sizeof(int) * size_t(n), hipMemcpyDeviceToHost, s));
        VBH_CHECK_HIP(hipMemcpyAsync(own.data(), d_own, sizeof(int) * size_t(n), hipMemcpyDeviceToHost, s));
        VBH_CHECK_HIP(hipStreamSynchronize(s));
    }
    for (int i = 0; i < n; ++i) {
        VBH_REQUIRE(int64_t(crp[i]) + low[i] + own[i] <= int64_t(INT32_MAX), "Hermitian completion exceeds 2^31 blocks");
        crp[i + 1] = crp[i] + low[i] + own[i];
        lp[i + 1] = lp[i] + low[i];
    }
    const int64_t total = crp[n];
    int* d_cci = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(total, 1))));
    int* d_src = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(total, 1))));
    int* d_mode = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(total, 1))));
    int* d_crow = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(total, 1))));
    if (total > 0) {
        int* d_crp = tmp.own(upload(crp, s));
        int* d_lp = tmp.own(upload(lp, s));
        int* d_lseg = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max(lp[n], 1))));
        VBH_CHECK_HIP(hipMemcpyAsync(d_low, d_lp, sizeof(int) * size_t(n), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_herm_scatter, dim3(nb4), dim3(256), 0, s, n, U->row_ptr, U->col_ind, d_low, d_lseg);
        VBH_CHECK_HIP(hipGetLastError());
        sort_rows_dev(n, lp.data(), d_lp, d_lseg, s);
        hipLaunchKernelGGL(k_herm_fill, dim3(nb4), dim3(256), 0, s, n, U->row_ptr, U->col_ind, d_crp, d_lp, d_lseg,
                           d_cci, d_src, d_mode, d_crow);
        VBH_CHECK_HIP(hipGetLastError());
    }
    vbh_matrix* C = build_matrix_dev(n, n, U->h_rbs.data(), U->h_cbs.data(), crp.data(), d_cci, U->dtype, s);
    try {
        if (C->nnz > 0) {
            const unsigned nb = unsigned((C->nnz + 255) / 256);
            if (U->dtype == VBH_F64)
                hipLaunchKernelGGL(k_complete_hermitian<double>, dim3(nb), dim3(256), 0, s, C->nnz, C->voff,
                                   int(C->nnzb), d_crow, C->col_ind, C->rbs, U->voff, d_src, d_mode,
                                   (const double*)U->vals, (double*)C->vals);
            else
                hipLaunchKernelGGL(k_complete_hermitian<c128>, dim3(nb), dim3(256), 0, s, C->nnz, C->voff,
                                   int(C->nnzb), d_crow, C->col_ind, C->rbs, U->voff, d_src, d_mode,
                                   (const c128*)U->vals, (c128*)C->vals);
            VBH_CHECK_HIP(hipGetLastError());
            VBH_CHECK_HIP(hipStreamSynchronize(s));
        }
    } catch (...) {
        destroy_matrix(C);
        throw;
    }
    *out = C;
}

extern "C" VBH_API int vbh_spgemm(vbh_matrix_t A, vbh_matrix_t B, double thr, int numeric, vbh_matrix_t* out,
                                  int64_t* stats, void* stream) {
    return guard_on(stream, [&] { spgemm_impl(A, B, thr, numeric, 0, out, stats, stream); });
}

extern "C" VBH_API int vbh_spgemm_upper(vbh_matrix_t A, vbh_matrix_t B, double thr, int numeric, vbh_matrix_t* out,
                                        int64_t* stats, void* stream) {
    return guard_on(stream, [&] { spgemm_impl(A, B, thr, numeric, 1, out, stats, stream); });
}

extern "C" VBH_API int vbh_matrix_complete_hermitian(vbh_matrix_t U, vbh_matrix_t* out, void* stream) {
    return guard_on(stream, [&] { complete_hermitian_impl(U, out, (hipStream_t)stream); });
}

extern "C" VBH_API int vbh_spgemm_hermitian(vbh_matrix_t A, vbh_matrix_t B, double thr, int numeric,
                                            vbh_matrix_t* out, int64_t* stats, void* stream) {
    return guard_on(stream, [&] {
        vbh_matrix_t U = nullptr;
        spgemm_impl(A, B, thr, numeric, 1, &U, stats, stream);
        try {
            complete_hermitian_impl(U, out, (hipStream_t)stream);
        } catch (...) {
            destroy_matrix(U);
            throw;
        }
        destroy_matrix(U);
    });
}
