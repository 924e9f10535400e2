// Shared definitions of the gfx950 variable-block sparse library: scalar
// types, the device view of a block pattern, the matrix handle, launch and
// error helpers.  Written for CDNA4 (64-wide wavefronts, 8 XCDs).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vbcsr_hip.h"

namespace vbh {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string& msg);

#define VBH_CHECK_HIP(expr)                                                                       \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            throw ::vbh::Error(VBH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define VBH_REQUIRE(cond, msg)                                    \
    do {                                                          \
        if (!(cond)) throw ::vbh::Error(VBH_ERR_ARG, (msg));      \
    } while (0)

#define VBH_REQUIRE_STATE(cond, msg)                              \
    do {                                                          \
        if (!(cond)) throw ::vbh::Error(VBH_ERR_STATE, (msg));    \
    } while (0)

template <typename F>
int guard(F&& f) {
    try {
        f();
        return VBH_OK;
    } catch (const Error& e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return VBH_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return VBH_ERR_ARG;
    }
}

// ---------------------------------------------------------------- scalars
template <typename R>
struct cplx {
    R re, im;
};
using c64 = cplx<float>;
using c128 = cplx<double>;

// a if c else b, member-wise for complex values: a select of a whole struct
// compiles to a round trip through scratch memory.
template <typename T>
__host__ __device__ inline T pick(bool c, T a, T b) {
    return c ? a : b;
}
template <typename R>
__host__ __device__ inline cplx<R> pick(bool c, cplx<R> a, cplx<R> b) {
    return {c ? a.re : b.re, c ? a.im : b.im};
}

template <typename T>
struct Sc;  // scalar traits

template <>
struct Sc<double> {
    using real = double;
    static constexpr bool is_complex = false;
    __host__ __device__ static double zero() { return 0.0; }
    __host__ __device__ static double make(double re, double) { return re; }
    __host__ __device__ static double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
    __host__ __device__ static double conj(double a) { return a; }
    __host__ __device__ static double re(double a) { return a; }
    __host__ __device__ static double im(double) { return 0.0; }
};
template <>
struct Sc<float> {
    using real = float;
    static constexpr bool is_complex = false;
    __host__ __device__ static float zero() { return 0.0f; }
    __host__ __device__ static float make(double re, double) { return float(re); }
    __host__ __device__ static float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    __host__ __device__ static float conj(float a) { return a; }
    __host__ __device__ static double re(float a) { return a; }
    __host__ __device__ static double im(float) { return 0.0; }
};
template <typename R>
struct Sc<cplx<R>> {
    using T = cplx<R>;
    using real = R;
    static constexpr bool is_complex = true;
    __host__ __device__ static T zero() { return T{R(0), R(0)}; }
    __host__ __device__ static T make(double re, double im) { return T{R(re), R(im)}; }
    // c + a*b
    __host__ __device__ static T fma(T a, T b, T c) {
        return T{c.re + a.re * b.re - a.im * b.im, c.im + a.re * b.im + a.im * b.re};
    }
    __host__ __device__ static T conj(T a) { return T{a.re, -a.im}; }
    __host__ __device__ static double re(T a) { return a.re; }
    __host__ __device__ static double im(T a) { return a.im; }
};

template <typename R>
__host__ __device__ inline cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__host__ __device__ inline cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
__host__ __device__ inline cplx<R> operator*(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__host__ __device__ inline cplx<R> operator*(R s, cplx<R> b) { return {s * b.re, s * b.im}; }

// real scale of any scalar
__host__ __device__ inline double rscale(double s, double a) { return s * a; }
__host__ __device__ inline float rscale(double s, float a) { return float(s) * a; }
template <typename R>
__host__ __device__ inline cplx<R> rscale(double s, cplx<R> a) { return {R(s) * a.re, R(s) * a.im}; }
// Re(conj(a) * b) in double
__host__ __device__ inline double redot(double a, double b) { return a * b; }
__host__ __device__ inline double redot(float a, float b) { return double(a) * double(b); }
template <typename R>
__host__ __device__ inline double redot(cplx<R> a, cplx<R> b) {
    return double(a.re) * double(b.re) + double(a.im) * double(b.im);
}
__host__ __device__ inline double sqabs(double a) { return a * a; }
__host__ __device__ inline double sqabs(float a) { return double(a) * a; }
template <typename R>
__host__ __device__ inline double sqabs(cplx<R> a) { return double(a.re) * a.re + double(a.im) * a.im; }

// Single-precision partner of a storage type.
template <typename T>
struct Low;
template <>
struct Low<double> { using type = float; };
template <>
struct Low<c128> { using type = c64; };

// ---------------------------------------------------------------- wave ops
__device__ inline double shfl_xor(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ inline float shfl_xor(float v, int m) { return __shfl_xor(v, m, 64); }
template <typename R>
__device__ inline cplx<R> shfl_xor(cplx<R> v, int m) {
    return {__shfl_xor(v.re, m, 64), __shfl_xor(v.im, m, 64)};
}

__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__host__ __device__ inline int pow2ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Work-group id remapped so that consecutive logical chunks land on the same
// XCD: the dispatcher hands work-groups to the 8 XCDs round-robin, and each
// XCD has its own L2, so a contiguous run of block rows (and the x rows they
// gather) stays in one L2.  Requires G % 8 == 0.
__device__ inline int xcd_chunk(int G) {
    const int b = blockIdx.x;
    const int per = G >> 3;
    return (b & 7) * per + (b >> 3);
}

// ---------------------------------------------------------------- pattern
// Device view of one block pattern.  Values of block k start at voff[k];
// element (a, b) of an r x c block is at vals[voff[k] + a*c + b].
struct DevPattern {
    int nbr, nbc;
    int bs;  // uniform block size (0 when variable)
    const int* rbs;
    const int* cbs;
    const int* row_ptr;
    const int* col_ind;
    const int* roff;  // nbr+1 scalar row offsets
    const int* coff;  // nbc+1 scalar column offsets
    const int64_t* voff;  // nnzb+1
    // per slot k of row i: {coff[col_ind[k]], (panel << 8) | c} where c is the
    // block's column size and panel = sum of c over the row's earlier slots,
    // so the block's values start at voff[row_ptr[i]] + rbs[i] * panel.  One
    // lane-parallel load yields every block's x offset, width and values.
    const int2* meta;
    const int* row_of;  // n_rows: block row of every scalar row
    // row classes of the real-type SpMM: block rows with kMfmaMinRows <= r <= 16
    // (matrix-core tiles, big_desc) and the scalar rows of all other block
    // rows (lane family, small_rows); n_big == 0 when there are none.  Each
    // matrix-core row has a 32-byte descriptor, so one load yields all it
    // needs: [0] = {row_ptr[i], row_ptr[i+1], rbs[i], roff[i]},
    // [1] = {coff[i], voff[row_ptr[i]] (low, high word), i}
    const int* small_rows;
    const int4* big_desc;
    int64_t n_small;
    int n_big;
};

constexpr int kWaves = 4;  // waves per work-group in the row kernels
constexpr int kMfmaRowsPerWave = 1;  // matrix-core block rows per wave (the next row's header is prefetched)
#ifndef VBH_MFMA_MIN_ROWS
#define VBH_MFMA_MIN_ROWS 4
#endif
constexpr int kMfmaMinRows = VBH_MFMA_MIN_ROWS;  // block rows of kMfmaMinRows..16 rows run on 16x16x4 FP64 matrix-core tiles
constexpr int kThreads = 64 * kWaves;
constexpr int kVecPad = 8;  // padding elements after every value array

}  // namespace vbh

// ---------------------------------------------------------------- handle
struct vbh_matrix {
    int nbr = 0, nbc = 0;
    int64_t nnzb = 0, nnz = 0, n_rows = 0, n_cols = 0;
    int dtype = VBH_F64, kind = 2, max_block = 0, n_classes = 0, uniform_bs = 0;
    // host pattern: block sizes and row pointers always; the column indices
    // and value offsets of a matrix whose pattern was made on the device
    // (products, sums, filtered copies) are downloaded on first host use
    // (host_col_ind / host_voff), under host_mu
    std::vector<int> h_rbs, h_cbs, h_row_ptr;
    mutable std::vector<int> h_col_ind;
    mutable std::vector<int64_t> h_voff;
    mutable bool h_cols_valid = true, h_voff_valid = true;
    mutable std::mutex host_mu;
    // device pattern
    int *rbs = nullptr, *cbs = nullptr, *row_ptr = nullptr, *col_ind = nullptr;
    int *roff = nullptr, *coff = nullptr;
    int64_t* voff = nullptr;
    int2* meta = nullptr;
    int* row_of = nullptr;
    // row-class lists, one set per matrix-core threshold t (row_classes()):
    // the scalar rows of the lane family and the matrix-core descriptors
    struct RowLists {
        int* small_rows = nullptr;
        int4* big_desc = nullptr;
        int64_t n_small = 0;
        int n_big = 0;
    };
    std::map<int, RowLists> row_lists;
    std::vector<int> h_roff, h_coff, h_bfs;  // host copies for building the lists
    // values
    void* vals = nullptr;
    void* vals32 = nullptr;  // FP32 copy, built on demand
    bool vals32_valid = false;
    // block-row wave kernel, built on demand: FP32 values with every block
    // column-major (same offsets voff), valid while the values are unchanged,
    // and the X row of every stored scalar column of each block row
    // (entries eoff[i] .. eoff[i+1]; pattern only)
    float* vt = nullptr;
    bool vt_valid = false;
    int64_t* eoff = nullptr;
    int* ecol = nullptr;
    // apply plan: work-balanced chunks of block rows, G chunks (G % 8 == 0)
    int* split = nullptr;
    int G = 0;
    // adjoint plan (CSC of the pattern), built on demand
    // incoming block q of column j (t_ptr[j] <= q < t_ptr[j+1]): its value
    // offset and (first scalar row, row count) of its block row
    int *t_ptr = nullptr, *t_split = nullptr;
    int64_t* t_voff = nullptr;
    int2* t_rows = nullptr;
    int tG = 0;
    bool t_valid = false;
    // Guards the structures built on demand (FP32 copy, row-class lists,
    // adjoint plan): calls on several streams may share one matrix.  Each is
    // complete on the device (its stream synchronised) before it is marked
    // valid, so a kernel on any stream may read it.
    std::mutex lazy_mu;

    vbh::DevPattern pattern(const RowLists& L) const {
        return vbh::DevPattern{nbr,  nbc,  uniform_bs, rbs,          cbs,        row_ptr,   col_ind, roff,
                               coff, voff, meta,       row_of,       L.small_rows, L.big_desc, L.n_small, L.n_big};
    }
    vbh::DevPattern pattern() const { return pattern(RowLists{}); }
    size_t elem_size() const { return dtype == VBH_C128 ? 16 : 8; }
};

namespace vbh {
// Host work over the block rows [0, nbr) of a CSR pattern (row_ptr) in
// contiguous row ranges of about equal slot counts, on up to 16 threads
// (serial below ~2^18 slots, where threads do not pay): fn(r0, r1) per
// range.  The first exception of any range is rethrown.
template <typename F>
void rows_parallel(int nbr, const int* row_ptr, F&& fn) {
    const int64_t nnzb = nbr > 0 ? row_ptr[nbr] : 0;
    const int hw = int(std::max(1u, std::thread::hardware_concurrency()));
    const int nt = int(std::min<int64_t>(std::min(hw, 16), std::max<int64_t>(1, nnzb >> 18)));
    if (nt <= 1 || nbr < 2 * nt) {
        fn(0, nbr);
        return;
    }
    std::vector<int> cut(nt + 1, nbr);
    cut[0] = 0;
    for (int t = 1; t < nt; ++t)
        cut[t] = int(std::upper_bound(row_ptr, row_ptr + nbr + 1, int(nnzb * t / nt)) - row_ptr) - 1;
    std::exception_ptr err;
    std::mutex mu;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            try {
                if (cut[t] < cut[t + 1]) fn(cut[t], cut[t + 1]);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu);
                if (!err) err = std::current_exception();
            }
        });
    for (auto& x : th) x.join();
    if (err) std::rethrow_exception(err);
}

size_t dtype_size(int dtype);
void* dalloc(size_t bytes);
void dfree(void* p);
template <typename T>
T* upload(const std::vector<T>& h, hipStream_t s = nullptr) {
    T* d = static_cast<T*>(dalloc(sizeof(T) * (h.empty() ? 1 : h.size())));
    if (!h.empty()) VBH_CHECK_HIP(hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
    return d;
}
// Ensure the FP32 copy of the values exists and is current.
void ensure_vals32(vbh_matrix* m, hipStream_t s);
// Mark the device values as changed: the FP32 copies are rebuilt on next use.
void values_changed(vbh_matrix* m);
// Ensure the block-row wave kernel's column-major FP32 values and entry
// columns exist and are current (real matrices).
void ensure_brw(vbh_matrix* m, hipStream_t s);
// Whether the FP32 copy is current / the row-class lists of threshold t exist
// (so using them allocates nothing).
bool has_vals32(vbh_matrix* m);
bool has_row_classes(vbh_matrix* m, int t);
// Ensure the adjoint (column-major incoming) plan exists.
void ensure_adjoint_plan(vbh_matrix* m, hipStream_t s);
// Row-class lists of matrix-core threshold t (block rows of t..16 rows on the
// matrix cores), built on first use and kept until the matrix is destroyed.
const vbh_matrix::RowLists& row_classes(vbh_matrix* m, int t, hipStream_t s);
// Build a matrix handle from host pattern arrays (no value upload).
vbh_matrix* build_matrix(int nbr, int nbc, const int* rbs, const int* cbs, const int* row_ptr,
                         const int* col_ind, int dtype, hipStream_t s);
// Build a matrix handle from a pattern whose column indices are on the
// device (d_col_ind, strictly ascending per row; copied, not adopted) and
// whose row pointers and block sizes are on the host: the value offsets and
// slot metadata are computed on the device, and the host mirror of the
// columns and offsets is left to first use.
vbh_matrix* build_matrix_dev(int nbr, int nbc, const int* rbs, const int* cbs, const int* row_ptr,
                             const int* d_col_ind, int dtype, hipStream_t s);
// Host mirror of the column indices / value offsets (downloaded on first use).
const std::vector<int>& host_col_ind(const vbh_matrix* m, hipStream_t s);
const std::vector<int64_t>& host_voff(const vbh_matrix* m, hipStream_t s);
// Whether two matrices have the same pattern (rows, columns, block sizes);
// compares the device column arrays.
bool same_pattern(const vbh_matrix* a, const vbh_matrix* b, hipStream_t s);
// Sort the columns of every row of a device pattern (row i: d_col_ind
// [row_ptr[i], row_ptr[i+1]); row_ptr on the host, d_row_ptr its device copy).
void sort_rows_dev(int nbr, const int* row_ptr, const int* d_row_ptr, int* d_col_ind, hipStream_t s);
void destroy_matrix(vbh_matrix* m);
// Filtered copy of m: blocks with norm < thr dropped (norm = sqrt(d_norm[k])
// when squared); null when every block is kept.  Device buffers only.
vbh_matrix* filter_by_norms(vbh_matrix* m, double thr, const double* d_norm, bool squared, hipStream_t s);

// Device buffers of one call, freed when it returns or throws.
struct DevList {
    std::vector<void*> bufs;
    template <typename P>
    P* own(P* p) {
        bufs.push_back((void*)p);
        return p;
    }
    DevList() = default;
    DevList(const DevList&) = delete;
    DevList& operator=(const DevList&) = delete;
    ~DevList() {
        for (void* p : bufs) dfree(p);
    }
};
// A matrix under construction: destroyed unless released to the caller.
struct MatrixGuard {
    vbh_matrix* m = nullptr;
    explicit MatrixGuard(vbh_matrix* x) : m(x) {}
    MatrixGuard(const MatrixGuard&) = delete;
    MatrixGuard& operator=(const MatrixGuard&) = delete;
    ~MatrixGuard() { destroy_matrix(m); }
    vbh_matrix* release() {
        vbh_matrix* r = m;
        m = nullptr;
        return r;
    }
    void reset(vbh_matrix* x) {
        destroy_matrix(m);
        m = x;
    }
};

// ---------------------------------------------------------------- per-stream state
// A second stream (default priority) and two events for fork/join of
// independent launches, eager or under graph capture.
struct StreamFork {
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    void init();
};
// A captured batched-KPM graph, the stream it is launched on and the key of
// the batch it was captured for.
struct GraphCache {
    hipStream_t cs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uintptr_t> key;
};
// Library state of one (device, stream): grow-only scratch slots and cached
// launch objects.  Every C-ABI call holds the state of its stream for the
// whole call (StreamScope), so calls on different streams never share a
// scratch buffer, and calls on one stream run one after another on the host
// as their work does on the device.  A slot only grows after its stream is
// synchronised, so no queued kernel still reads the buffer it frees.
struct StreamState {
    std::recursive_mutex mu;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<std::pair<void*, size_t>> slots;
    StreamFork fork;
    GraphCache graph;
};
// RAII: locks the state of (current device, s) and makes it the current
// state of this thread (scratch() uses the innermost scope).
class StreamScope {
public:
    explicit StreamScope(hipStream_t s);
    ~StreamScope();
    StreamScope(const StreamScope&) = delete;
    StreamScope& operator=(const StreamScope&) = delete;
    StreamState& state() { return *st_; }

private:
    StreamState* st_;
    StreamState* prev_;
};
// Scratch slot of the current StreamScope (grown on demand).
void* scratch(size_t bytes, int slot);
// State of the current StreamScope.
StreamState& current_stream_state();
// guard() with the stream's state held for the call.
template <typename F>
int guard_on(void* stream, F&& f) {
    return guard([&] {
        StreamScope scope((hipStream_t)stream);
        f();
    });
}
}  // namespace vbh
