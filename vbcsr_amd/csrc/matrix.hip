// Matrix handles: construction of the device pattern, value transfer, the
// FP32 value copy, the adjoint plan, and the value-level matrix algebra
// (scale, shift, block norms, filter, union axpby, trace).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>

#include "common.hpp"

namespace vbh {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

size_t dtype_size(int dtype) {
    switch (dtype) {
        case VBH_F64: return 8;
        case VBH_C128: return 16;
        case VBH_F32: return 4;
        case VBH_C64: return 8;
    }
    throw Error(VBH_ERR_ARG, "unknown dtype");
}

// Large device buffers (>= 64 MiB) are kept for reuse when freed, per
// device, up to 16 GiB: releasing a multi-GB buffer took 130-175 ms here
// (the two largest frees of a late McWeeny iteration), and products and
// sums allocate buffers of similar sizes again.  A kept buffer is handed out
// for a request of at least 4/5 of its size.  dfree synchronises the device
// before keeping a buffer, as hipFree does, so no queued kernel still uses
// it when it is handed out again; a failed allocation releases the kept
// buffers and retries.
namespace {
struct KeptBlock {
    void* p;
    size_t bytes;
    int dev;
};
constexpr size_t kKeepMin = size_t(64) << 20;
constexpr size_t kKeepCap = size_t(16) << 30;
std::mutex g_keep_mu;
std::vector<KeptBlock> g_kept;                      // free, kept for reuse
std::map<void*, std::pair<size_t, int>> g_large;     // live large blocks: size, device
size_t g_kept_bytes = 0;

void release_kept(int dev) {  // g_keep_mu held
    for (size_t q = 0; q < g_kept.size();) {
        if (g_kept[q].dev == dev) {
            (void)hipFree(g_kept[q].p);
            g_kept_bytes -= g_kept[q].bytes;
            g_kept[q] = g_kept.back();
            g_kept.pop_back();
        } else {
            ++q;
        }
    }
}
}  // namespace

void* dalloc(size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 16;
    if (bytes >= kKeepMin) {
        int dev = 0;
        VBH_CHECK_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lk(g_keep_mu);
        size_t best = SIZE_MAX, bq = 0;
        for (size_t q = 0; q < g_kept.size(); ++q)
            if (g_kept[q].dev == dev && g_kept[q].bytes >= bytes && g_kept[q].bytes / 5 * 4 <= bytes &&
                g_kept[q].bytes < best) {
                best = g_kept[q].bytes;
                bq = q;
            }
        if (best != SIZE_MAX) {
            p = g_kept[bq].p;
            g_kept_bytes -= best;
            g_kept[bq] = g_kept.back();
            g_kept.pop_back();
            g_large[p] = {best, dev};
            return p;
        }
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess && !g_kept.empty()) {
            (void)hipGetLastError();
            release_kept(dev);
            e = hipMalloc(&p, bytes);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(VBH_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
        }
        g_large[p] = {bytes, dev};
        return p;
    }
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(VBH_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
    }
    return p;
}

void dfree(void* p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_keep_mu);
        auto it = g_large.find(p);
        if (it != g_large.end()) {
            const size_t bytes = it->second.first;
            const int dev = it->second.second;
            g_large.erase(it);
            int cur = -1;
            if (hipGetDevice(&cur) == hipSuccess && cur == dev && g_kept_bytes + bytes <= kKeepCap &&
                hipDeviceSynchronize() == hipSuccess) {
                g_kept.push_back({p, bytes, dev});
                g_kept_bytes += bytes;
                return;
            }
            (void)hipGetLastError();
        }
    }
    (void)hipFree(p);
}

// ---------------------------------------------------------------- stream state
// One StreamState per (device, stream), created on first use and kept until
// vbh_stream_release or the end of the process.
static std::mutex g_states_mu;
static std::map<std::pair<int, uintptr_t>, std::unique_ptr<StreamState>> g_states;
static thread_local StreamState* t_state = nullptr;
constexpr int kScratchSlots = 12;

static StreamState* state_of(hipStream_t s) {
    int dev = 0;
    VBH_CHECK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_states_mu);
    auto& st = g_states[{dev, reinterpret_cast<uintptr_t>(s)}];
    if (!st) {
        st.reset(new StreamState());
        st->device = dev;
        st->stream = s;
        st->slots.assign(kScratchSlots, {nullptr, 0});
    }
    return st.get();
}

StreamScope::StreamScope(hipStream_t s) : st_(state_of(s)), prev_(t_state) {
    st_->mu.lock();
    t_state = st_;
}
StreamScope::~StreamScope() {
    t_state = prev_;
    st_->mu.unlock();
}

void* scratch(size_t bytes, int slot) {
    StreamState* st = t_state;
    if (!st) throw Error(VBH_ERR_STATE, "internal: scratch used outside a stream scope");
    auto& sl = st->slots.at(slot);
    if (sl.second < bytes) {
        if (sl.first) {
            // kernels queued on this stream may still read the old buffer
            VBH_CHECK_HIP(hipStreamSynchronize(st->stream));
            dfree(sl.first);
            sl = {nullptr, 0};
        }
        sl.first = dalloc(bytes);
        sl.second = bytes;
    }
    return sl.first;
}

StreamState& current_stream_state() {
    if (!t_state) throw Error(VBH_ERR_STATE, "internal: no stream scope");
    return *t_state;
}

void StreamFork::init() {
    if (side) return;
    VBH_CHECK_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    VBH_CHECK_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    VBH_CHECK_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
}

// Frees everything a stream state holds (its stream synchronised first).
static void release_state(StreamState* st) {
    (void)hipStreamSynchronize(st->stream);
    if (st->graph.cs) (void)hipStreamSynchronize(st->graph.cs);
    if (st->fork.side) (void)hipStreamSynchronize(st->fork.side);
    for (auto& sl : st->slots) dfree(sl.first);
    if (st->graph.exec) (void)hipGraphExecDestroy(st->graph.exec);
    if (st->graph.graph) (void)hipGraphDestroy(st->graph.graph);
    if (st->graph.cs) (void)hipStreamDestroy(st->graph.cs);
    if (st->fork.side) (void)hipStreamDestroy(st->fork.side);
    if (st->fork.fork) (void)hipEventDestroy(st->fork.fork);
    if (st->fork.join) (void)hipEventDestroy(st->fork.join);
}

// ---------------------------------------------------------------- kernels
template <typename T, typename L>
__global__ void k_convert(const T* __restrict__ in, L* __restrict__ out, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (Sc<T>::is_complex) out[i] = L{(float)in[i].re, (float)in[i].im};
    else out[i] = (L)in[i];
}

template <typename T>
__global__ void k_scale(T* __restrict__ v, int64_t n, T alpha) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = v[i] * alpha;
}
__global__ void k_scale_d(double* __restrict__ v, int64_t n, double alpha) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] *= alpha;
}

// Frobenius norm per block: one wave per block, FP64 accumulation.
template <typename T>
__global__ __launch_bounds__(256) void k_block_norms(const T* __restrict__ vals, const int64_t* __restrict__ voff,
                                                     int64_t nnzb, double* __restrict__ out) {
    const int64_t k = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= nnzb) return;
    const int64_t b = voff[k], e = voff[k + 1];
    double s = 0.0;
    for (int64_t i = b + lane; i < e; i += 64) s += sqabs(vals[i]);
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if (lane == 0) out[k] = sqrt(s);
}

// shift: alpha added to the leading min(r,c) diagonal of row i's own block.
template <typename T>
__global__ void k_shift(T* __restrict__ vals, const int64_t* __restrict__ voff, const int* __restrict__ diag_slot,
                        const int* __restrict__ rbs, const int* __restrict__ cbs, const int* __restrict__ col_ind,
                        int nbr, T alpha) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbr) return;
    const int k = diag_slot[i];
    if (k < 0) return;
    const int r = rbs[i], c = cbs[col_ind[k]];
    const int d = r < c ? r : c;
    T* blk = vals + voff[k];
    for (int t = 0; t < d; ++t) blk[int64_t(t) * c + t] = blk[int64_t(t) * c + t] + alpha;
}

// Copy selected blocks: dst block q takes src block src_slot[q] (whole block).
template <typename T>
__global__ __launch_bounds__(256) void k_gather_blocks(const T* __restrict__ src, const int64_t* __restrict__ src_voff,
                                                       const int* __restrict__ src_slot, T* __restrict__ dst,
                                                       const int64_t* __restrict__ dst_voff, int64_t nq) {
    const int64_t q = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= nq) return;
    const int64_t sb = src_voff[src_slot[q]];
    const int64_t db = dst_voff[q], de = dst_voff[q + 1];
    for (int64_t i = lane; i < de - db; i += 64) dst[db + i] = src[sb + i];
}

// Union axpby: dst block q = beta*Y[ys[q]] (if ys >= 0) + alpha*X[xs[q]] (if xs >= 0).
template <typename T>
__global__ __launch_bounds__(256) void k_union_axpby(const T* __restrict__ Y, const int64_t* __restrict__ yvoff,
                                                     const int* __restrict__ ys, const T* __restrict__ X,
                                                     const int64_t* __restrict__ xvoff, const int* __restrict__ xs,
                                                     T* __restrict__ dst, const int64_t* __restrict__ dvoff,
                                                     int64_t nq, T alpha, T beta) {
    const int64_t q = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= nq) return;
    const int64_t db = dvoff[q], n = dvoff[q + 1] - db;
    const int yq = ys[q], xq = xs[q];
    for (int64_t i = lane; i < n; i += 64) {
        T acc = Sc<T>::zero();
        if (yq >= 0) acc = beta * Y[yvoff[yq] + i];
        if (xq >= 0) acc = Sc<T>::fma(alpha, X[xvoff[xq] + i], acc);
        dst[db + i] = acc;
    }
}
// Union of the sorted columns of row i of two patterns (Y: yp/yc, X:
// xp/xc).  Pass 0 (out_ptr null): counts[i] = the merged length.  Pass 1:
// the merged columns from out_ptr[i] on, with the source slot of each side
// (oy / ox, -1: absent).  One wavefront per row.  The merged position of a
// column is (Y columns below it) + (X-only columns below it): for Y's
// element k, (k - yb) + (X columns below, by binary search) - (Y columns
// before k that X also holds); for an X-only element l, (Y columns below,
// by binary search) + (X-only elements before l).  The two running counts
// are ballot prefix sums over the wave's chunks.  A column in both rows is
// written by Y's side.
__global__ __launch_bounds__(256) void k_union_rows(int nbr, const int* __restrict__ yp, const int* __restrict__ yc,
                                                    const int* __restrict__ xp, const int* __restrict__ xc,
                                                    const int* __restrict__ out_ptr, int* __restrict__ counts,
                                                    int* __restrict__ oc, int* __restrict__ oy, int* __restrict__ ox) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nbr) return;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    const int yb = yp[i], ye = yp[i + 1], xb = xp[i], xe = xp[i + 1];
    auto lower = [](const int* c, int lo, int hi, int v) {
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (c[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    int common = 0;  // Y columns so far that X also holds
    for (int k0 = yb; k0 < ye; k0 += 64) {
        const int k = k0 + lane;
        int c = 0, l = xb;
        bool match = false;
        if (k < ye) {
            c = yc[k];
            l = lower(xc, xb, xe, c);
            match = l < xe && xc[l] == c;
        }
        const uint64_t mm = __ballot(match);
        if (out_ptr && k < ye) {
            const int pos = out_ptr[i] + (k - yb) + (l - xb) - (common + __popcll(mm & below));
            oc[pos] = c;
            oy[pos] = k;
            ox[pos] = match ? l : -1;
        }
        common += __popcll(mm);
    }
    if (!out_ptr) {
        if (lane == 0) counts[i] = (ye - yb) + (xe - xb) - common;
        return;
    }
    int xonly = 0;  // X-only columns so far
    for (int l0 = xb; l0 < xe; l0 += 64) {
        const int l = l0 + lane;
        int c = 0, k = yb;
        bool only = false;
        if (l < xe) {
            c = xc[l];
            k = lower(yc, yb, ye, c);
            only = !(k < ye && yc[k] == c);
        }
        const uint64_t mo = __ballot(only);
        if (only) {
            const int pos = out_ptr[i] + (k - yb) + xonly + __popcll(mo & below);
            oc[pos] = c;
            oy[pos] = -1;
            ox[pos] = l;
        }
        xonly += __popcll(mo);
    }
}

template <typename T>
__global__ void k_axpby_flat(const T* __restrict__ X, T* __restrict__ Y, int64_t n, T alpha, T beta) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) Y[i] = Sc<T>::fma(alpha, X[i], beta * Y[i]);
}
__global__ void k_axpby_flat_d(const double* __restrict__ X, double* __restrict__ Y, int64_t n, double alpha,
                               double beta) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) Y[i] = __builtin_fma(alpha, X[i], beta * Y[i]);
}

// Diagonal trace: sum over rows of the diagonal of the row's own block.
template <typename T>
__global__ __launch_bounds__(256) void k_trace(const T* __restrict__ vals, const int64_t* __restrict__ voff,
                                               const int* __restrict__ diag_slot, const int* __restrict__ rbs,
                                               int nbr, double* __restrict__ partial) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nbr; i += gridDim.x * blockDim.x) {
        const int k = diag_slot[i];
        if (k < 0) continue;
        const int r = rbs[i];
        for (int t = 0; t < r; ++t) s += Sc<T>::re(vals[voff[k] + int64_t(t) * r + t]);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int m = 128; m >= 1; m >>= 1) {
        if (threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------- helpers
static inline unsigned blocks_for(int64_t n, int per) { return unsigned((n + per - 1) / per); }

template <typename F>
static void dispatch_dtype(int dtype, F&& f) {
    if (dtype == VBH_F64) f(double{});
    else if (dtype == VBH_C128) f(c128{});
    else throw Error(VBH_ERR_NOTSUP, "matrix dtype must be VBH_F64 or VBH_C128");
}

template <typename T>
static T scalar_from(const double* a) {
    return Sc<T>::make(a[0], a[1]);
}

// Breadth-first order of block rows over the local square part of the
// pattern (columns < nbr), restarted at the lowest unvisited row for each
// connected component.
static std::vector<int> bfs_order(int nbr, const int* row_ptr, const int* col_ind) {
    std::vector<int> order;
    order.reserve(nbr);
    std::vector<char> seen(nbr, 0);
    for (int start = 0; start < nbr; ++start) {
        if (seen[start]) continue;
        seen[start] = 1;
        size_t head = order.size();
        order.push_back(start);
        while (head < order.size()) {
            const int i = order[head++];
            for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
                const int j = col_ind[k];
                if (j < nbr && !seen[j]) {
                    seen[j] = 1;
                    order.push_back(j);
                }
            }
        }
    }
    return order;
}

// Row-class lists for matrix-core threshold t (built once per threshold and
// kept): block rows of t..16 rows of a variable-block matrix run on the
// matrix cores (uniform 4 x 4 matrices keep the quad-shared lane path), the
// scalar rows of all others on the lane family.  Both lists follow a
// breadth-first order of the block graph: rows that run at the same time on
// one XCD are then graph neighbours, so the X rows they gather overlap and
// stay in that XCD's L2.  Each row is computed exactly as in index order; only
// the order of the rows changes.
const vbh_matrix::RowLists& row_classes(vbh_matrix* m, int t, hipStream_t s) {
    std::lock_guard<std::mutex> lk(m->lazy_mu);
    auto found = m->row_lists.find(t);
    if (found != m->row_lists.end()) return found->second;
    const int nbr = m->nbr;
    const std::vector<int>& rbs = m->h_rbs;
    auto is_big = [&](int i) { return m->uniform_bs != 4 && rbs[i] >= t && rbs[i] <= 16; };
    bool any_big = false;
    for (int i = 0; i < nbr && !any_big; ++i) any_big = is_big(i);
    vbh_matrix::RowLists L;
    if (any_big) {
        if (m->h_bfs.empty()) m->h_bfs = bfs_order(nbr, m->h_row_ptr.data(), host_col_ind(m, s).data());
        const std::vector<int64_t>& hv = host_voff(m, s);
        std::vector<int> small, big;
        for (int i : m->h_bfs) {
            if (is_big(i)) big.push_back(i);
            else
                for (int r = m->h_roff[i]; r < m->h_roff[i + 1]; ++r) small.push_back(r);
        }
        std::vector<int4> desc(2 * big.size());
        for (size_t q = 0; q < big.size(); ++q) {
            const int i = big[q];
            const int kb = m->h_row_ptr[i];
            const uint64_t v = uint64_t(hv[kb]);
            desc[2 * q] = make_int4(kb, m->h_row_ptr[i + 1], rbs[i], m->h_roff[i]);
            desc[2 * q + 1] = make_int4(m->h_coff[i], int(uint32_t(v)), int(uint32_t(v >> 32)), i);
        }
        L.n_big = int(big.size());
        L.n_small = int64_t(small.size());
        L.big_desc = upload(desc, s);
        try {
            L.small_rows = upload(small, s);
            // read by kernels on any stream: complete before it is published
            VBH_CHECK_HIP(hipStreamSynchronize(s));
        } catch (...) {
            dfree(L.big_desc);
            dfree(L.small_rows);
            throw;
        }
    }
    return m->row_lists.emplace(t, L).first->second;
}

// Block sizes of a new matrix (rows rbs, columns cbs): checks, extents,
// kind, uniform size, and the scalar offsets of block rows and columns.
static void set_block_layout(vbh_matrix* m, const int* rbs, const int* cbs, std::vector<int>& roff,
                             std::vector<int>& coff) {
    const int nbr = m->nbr, nbc = m->nbc;
    m->h_rbs.assign(rbs, rbs + nbr);
    m->h_cbs.assign(cbs, cbs + nbc);
    int mn = 1 << 30, mx = 0;
    for (int i = 0; i < nbr; ++i) {
        VBH_REQUIRE(rbs[i] > 0, "block sizes must be positive");
        mn = std::min(mn, rbs[i]);
        mx = std::max(mx, rbs[i]);
    }
    for (int j = 0; j < nbc; ++j) {
        VBH_REQUIRE(cbs[j] > 0, "block sizes must be positive");
        mn = std::min(mn, cbs[j]);
        mx = std::max(mx, cbs[j]);
    }
    VBH_REQUIRE(mx <= 255, "block sizes above 255 are not supported");
    m->max_block = mx;
    // kind detection: block_csr.hpp:2770-2798 (on the owned row sizes)
    int rmn = 1 << 30, rmx = 0;
    for (int i = 0; i < nbr; ++i) { rmn = std::min(rmn, rbs[i]); rmx = std::max(rmx, rbs[i]); }
    if (nbr == 0) { rmn = 1; rmx = 1; }
    m->kind = rmx <= 1 ? 0 : (rmn == rmx ? 1 : 2);
    m->uniform_bs = (mn == mx && nbc > 0) ? mx : 0;
    roff.assign(nbr + 1, 0);
    coff.assign(nbc + 1, 0);
    for (int i = 0; i < nbr; ++i) roff[i + 1] = roff[i] + rbs[i];
    for (int j = 0; j < nbc; ++j) coff[j + 1] = coff[j] + cbs[j];
    m->n_rows = roff[nbr];
    m->n_cols = coff[nbc];
}

static void set_row_ptr(vbh_matrix* m, const int* row_ptr) {
    m->h_row_ptr.assign(row_ptr, row_ptr + m->nbr + 1);
    VBH_REQUIRE(m->h_row_ptr[0] == 0, "row_ptr must start at 0");
    for (int i = 0; i < m->nbr; ++i) VBH_REQUIRE(row_ptr[i + 1] >= row_ptr[i], "row_ptr not monotone");
    m->nnzb = row_ptr[m->nbr];
}

// The rest of a new matrix once its slots are laid out (col_ind, voff, meta
// on the device): work-balanced chunks of rows from the inclusive prefix of
// the row work weights (work[i + 1] = rows 0..i), the small device arrays,
// and zeroed values.  Synchronises the stream, so host arrays the uploads
// read may go out of scope afterwards.
static void finish_matrix(vbh_matrix* m, const std::vector<int64_t>& work, const std::vector<int>& roff,
                          const std::vector<int>& coff, hipStream_t s) {
    const int nbr = m->nbr;
    // work-balanced chunks of rows, G a multiple of 8 (one run per XCD)
#ifndef VBH_MAX_GROUPS
#define VBH_MAX_GROUPS 8192
#endif
    int G = int(std::min<int64_t>(std::max<int64_t>((nbr + kWaves - 1) / kWaves, 8), VBH_MAX_GROUPS));
    G = (G + 7) & ~7;
    std::vector<int> split(G + 1, nbr);
    split[0] = 0;
    const int64_t total = work[nbr];
    int row = 0;
    for (int g = 1; g < G; ++g) {
        const int64_t target = total * g / G;
        while (row < nbr && work[row] < target) ++row;
        split[g] = row;
    }
    split[G] = nbr;
    m->G = G;
    if (!m->rbs) m->rbs = upload(m->h_rbs, s);
    if (!m->cbs) m->cbs = upload(m->h_cbs, s);
    if (!m->row_ptr) m->row_ptr = upload(m->h_row_ptr, s);
    if (!m->roff) m->roff = upload(roff, s);
    if (!m->coff) m->coff = upload(coff, s);
    std::vector<int> row_of(std::max<int64_t>(m->n_rows, 1), 0);
    for (int i = 0; i < nbr; ++i)
        for (int t = roff[i]; t < roff[i + 1]; ++t) row_of[t] = i;
    m->row_of = upload(row_of, s);
    m->h_roff = roff;
    m->h_coff = coff;
    // the row-class lists (breadth-first order, matrix-core descriptors) are
    // built by the first apply (make_args): products, filtered copies and
    // sums that are never applied to vectors do not pay for them
    m->split = upload(split, s);
    // kVecPad trailing elements: the streaming kernels read whole 16-byte
    // words, which may reach up to one word past the last value
    m->vals = dalloc(m->elem_size() * (m->nnz + kVecPad));
    VBH_CHECK_HIP(hipMemsetAsync(m->vals, 0, m->elem_size() * (m->nnz + kVecPad), s));
    // the pageable-memory uploads read the host arrays until they complete
    VBH_CHECK_HIP(hipStreamSynchronize(s));
}

vbh_matrix* build_matrix(int nbr, int nbc, const int* rbs, const int* cbs, const int* row_ptr, const int* col_ind,
                         int dtype, hipStream_t s) {
    VBH_REQUIRE(nbr >= 0 && nbc >= 0, "negative block counts");
    VBH_REQUIRE(dtype == VBH_F64 || dtype == VBH_C128, "matrix dtype must be VBH_F64 or VBH_C128");
    auto* m = new vbh_matrix();
    try {
        m->nbr = nbr;
        m->nbc = nbc;
        m->dtype = dtype;
        set_row_ptr(m, row_ptr);
        m->h_col_ind.assign(col_ind, col_ind + m->nnzb);
        std::vector<int> roff, coff;
        set_block_layout(m, rbs, cbs, roff, coff);
        const int mx = m->max_block;
        m->h_voff.assign(m->nnzb + 1, 0);
        std::vector<int2> meta(m->nnzb);
        std::vector<int64_t> work(nbr + 1, 0), rowsz(nbr + 1, 0);
        std::vector<char> cls(size_t(mx + 1) * (mx + 1), 0);
        std::mutex cls_mu;
        // per row (in parallel row ranges): checks, slot metadata, the row's
        // value count and work weight; then the row prefixes; then the value
        // offsets of every slot
        rows_parallel(nbr, row_ptr, [&](int r0, int r1) {
            std::vector<char> mine(cls.size(), 0);
            for (int i = r0; i < r1; ++i) {
                int64_t w = 64, sz_row = 0;
                int panel = 0;
                for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
                    const int j = col_ind[k];
                    VBH_REQUIRE(j >= 0 && j < nbc, "column index out of range");
                    if (k > row_ptr[i]) VBH_REQUIRE(col_ind[k - 1] < j, "columns must be strictly ascending per row");
                    const int64_t sz = int64_t(rbs[i]) * cbs[j];
                    meta[k] = make_int2(coff[j], (panel << 8) | cbs[j]);
                    panel += cbs[j];
                    VBH_REQUIRE(panel < (1 << 23), "block row wider than 2^23 scalar columns");
                    sz_row += sz;
                    w += sz + 16;
                    mine[size_t(rbs[i]) * (mx + 1) + cbs[j]] = 1;
                }
                rowsz[i + 1] = sz_row;
                work[i + 1] = w;
            }
            std::lock_guard<std::mutex> lk(cls_mu);
            for (size_t c = 0; c < cls.size(); ++c) cls[c] |= mine[c];
        });
        for (int i = 0; i < nbr; ++i) {
            rowsz[i + 1] += rowsz[i];
            work[i + 1] += work[i];
        }
        rows_parallel(nbr, row_ptr, [&](int r0, int r1) {
            for (int i = r0; i < r1; ++i) {
                int64_t v = rowsz[i];
                for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
                    v += int64_t(rbs[i]) * cbs[col_ind[k]];
                    m->h_voff[k + 1] = v;
                }
            }
        });
        m->n_classes = int(std::count(cls.begin(), cls.end(), 1));
        m->nnz = m->h_voff[m->nnzb];
        m->col_ind = upload(m->h_col_ind, s);
        m->voff = upload(m->h_voff, s);
        m->meta = upload(meta, s);
        finish_matrix(m, work, roff, coff, s);
    } catch (...) {
        destroy_matrix(m);
        throw;
    }
    return m;
}

// Per block row (one wavefront each): the row's value count and work
// weight, the (r, c) block classes present (a flag per class), and a check
// of the columns (in range, strictly ascending, row narrower than 2^23 scalar
// columns): bad[0] is set to 1 when a check fails.
__global__ __launch_bounds__(256) void k_row_sizes(int nbr, int nbc, const int* __restrict__ row_ptr,
                                                   const int* __restrict__ col_ind, const int* __restrict__ rbs,
                                                   const int* __restrict__ cbs, int mx, int64_t* __restrict__ rowsz,
                                                   int64_t* __restrict__ work, unsigned char* __restrict__ cls,
                                                   int* __restrict__ bad) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nbr) return;
    const int r = rbs[i], kb = row_ptr[i], ke = row_ptr[i + 1];
    int64_t width = 0;
    bool ok = true;
    for (int k = kb + lane; k < ke; k += 64) {
        const int j = col_ind[k];
        if (j < 0 || j >= nbc || (k > kb && col_ind[k - 1] >= j)) {
            ok = false;
            continue;
        }
        const int c = cbs[j];
        width += c;
        cls[r * (mx + 1) + c] = 1;
    }
    for (int d = 32; d >= 1; d >>= 1) width += __shfl_xor(width, d, 64);
    const bool any_bad = __ballot(!ok) != 0;
    if (lane == 0) {
        rowsz[i] = int64_t(r) * width;
        work[i] = 64 + int64_t(r) * width + 16 * int64_t(ke - kb);
        if (any_bad || width >= (int64_t(1) << 23)) bad[0] = 1;
    }
}

// Value offsets and slot metadata: voff[k + 1] = the row's first value +
// r * (block widths of the row up to and including slot k); meta[k] =
// (first scalar column of the block, offset of the block in the row's panel
// << 8 | block width).  One wavefront per block row, an in-wave scan of the
// block widths; voff[0] is set by the caller.
__global__ __launch_bounds__(256) void k_slot_layout(int nbr, const int* __restrict__ row_ptr,
                                                     const int* __restrict__ col_ind, const int* __restrict__ rbs,
                                                     const int* __restrict__ cbs, const int* __restrict__ coff,
                                                     const int64_t* __restrict__ rowstart, int64_t* __restrict__ voff,
                                                     int2* __restrict__ meta) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nbr) return;
    const int r = rbs[i], ke = row_ptr[i + 1];
    int64_t base = rowstart[i];
    int panel = 0;
    for (int k0 = row_ptr[i]; k0 < ke; k0 += 64) {
        const int k = k0 + lane;
        const int j = k < ke ? col_ind[k] : 0;
        const int c = k < ke ? cbs[j] : 0;
        int inc = c;
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(inc, d, 64);
            if (lane >= d) inc += t;
        }
        if (k < ke) {
            voff[k + 1] = base + int64_t(r) * inc;
            meta[k] = make_int2(coff[j], ((panel + inc - c) << 8) | c);
        }
        const int tot = __shfl(inc, 63, 64);
        base += int64_t(r) * tot;
        panel += tot;
    }
}

vbh_matrix* build_matrix_dev(int nbr, int nbc, const int* rbs, const int* cbs, const int* row_ptr,
                             const int* d_col_ind, int dtype, hipStream_t s) {
    VBH_REQUIRE(nbr >= 0 && nbc >= 0, "negative block counts");
    VBH_REQUIRE(dtype == VBH_F64 || dtype == VBH_C128, "matrix dtype must be VBH_F64 or VBH_C128");
    auto* m = new vbh_matrix();
    try {
        m->nbr = nbr;
        m->nbc = nbc;
        m->dtype = dtype;
        set_row_ptr(m, row_ptr);
        std::vector<int> roff, coff;
        set_block_layout(m, rbs, cbs, roff, coff);
        const int mx = m->max_block;
        m->h_cols_valid = m->h_voff_valid = false;
        m->rbs = upload(m->h_rbs, s);
        m->cbs = upload(m->h_cbs, s);
        m->row_ptr = upload(m->h_row_ptr, s);
        m->roff = upload(roff, s);
        m->coff = upload(coff, s);
        m->col_ind = (int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(m->nnzb, 1)));
        m->voff = (int64_t*)dalloc(sizeof(int64_t) * size_t(m->nnzb + 1));
        m->meta = (int2*)dalloc(sizeof(int2) * size_t(std::max<int64_t>(m->nnzb, 1)));
        VBH_CHECK_HIP(hipMemsetAsync(m->voff, 0, sizeof(int64_t), s));
        std::vector<int64_t> work(nbr + 1, 0), rowstart(nbr + 1, 0);
        const size_t ncls = size_t(mx + 1) * (mx + 1);
        std::vector<unsigned char> cls(ncls, 0);
        int bad = 0;
        if (nbr > 0) {
            DevList tmp;
            int64_t* d_rowsz = tmp.own((int64_t*)dalloc(sizeof(int64_t) * size_t(nbr)));
            int64_t* d_work = tmp.own((int64_t*)dalloc(sizeof(int64_t) * size_t(nbr)));
            unsigned char* d_cls = tmp.own((unsigned char*)dalloc(ncls));
            int* d_bad = tmp.own((int*)dalloc(sizeof(int)));
            VBH_CHECK_HIP(hipMemsetAsync(d_cls, 0, ncls, s));
            VBH_CHECK_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), s));
            if (m->nnzb > 0)
                VBH_CHECK_HIP(hipMemcpyAsync(m->col_ind, d_col_ind, sizeof(int) * size_t(m->nnzb),
                                             hipMemcpyDeviceToDevice, s));
            const unsigned nb = unsigned((nbr + 3) / 4);
            hipLaunchKernelGGL(k_row_sizes, dim3(nb), dim3(256), 0, s, nbr, nbc, m->row_ptr, m->col_ind, m->rbs,
                               m->cbs, mx, d_rowsz, d_work, d_cls, d_bad);
            VBH_CHECK_HIP(hipGetLastError());
            VBH_CHECK_HIP(hipMemcpyAsync(rowstart.data() + 1, d_rowsz, sizeof(int64_t) * size_t(nbr),
                                         hipMemcpyDeviceToHost, s));
            VBH_CHECK_HIP(hipMemcpyAsync(work.data() + 1, d_work, sizeof(int64_t) * size_t(nbr),
                                         hipMemcpyDeviceToHost, s));
            VBH_CHECK_HIP(hipMemcpyAsync(cls.data(), d_cls, ncls, hipMemcpyDeviceToHost, s));
            VBH_CHECK_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
            VBH_CHECK_HIP(hipStreamSynchronize(s));
            VBH_REQUIRE(!bad, "device pattern: column index out of range, columns not strictly ascending, or a "
                              "block row wider than 2^23 scalar columns");
            for (int i = 0; i < nbr; ++i) {
                rowstart[i + 1] += rowstart[i];
                work[i + 1] += work[i];
            }
            int64_t* d_rowstart = tmp.own(upload(rowstart, s));
            hipLaunchKernelGGL(k_slot_layout, dim3(nb), dim3(256), 0, s, nbr, m->row_ptr, m->col_ind, m->rbs,
                               m->cbs, m->coff, d_rowstart, m->voff, m->meta);
            VBH_CHECK_HIP(hipGetLastError());
            VBH_CHECK_HIP(hipStreamSynchronize(s));  // before the temporaries are freed
        }
        m->n_classes = int(std::count(cls.begin(), cls.end(), (unsigned char)1));
        m->nnz = rowstart[nbr];
        finish_matrix(m, work, roff, coff, s);
    } catch (...) {
        destroy_matrix(m);
        throw;
    }
    return m;
}

const std::vector<int>& host_col_ind(const vbh_matrix* m, hipStream_t s) {
    std::lock_guard<std::mutex> lk(m->host_mu);
    if (!m->h_cols_valid) {
        std::vector<int> h(size_t(m->nnzb));
        if (m->nnzb > 0) {
            VBH_CHECK_HIP(hipMemcpyAsync(h.data(), m->col_ind, sizeof(int) * h.size(), hipMemcpyDeviceToHost, s));
            VBH_CHECK_HIP(hipStreamSynchronize(s));
        }
        m->h_col_ind.swap(h);
        m->h_cols_valid = true;
    }
    return m->h_col_ind;
}

const std::vector<int64_t>& host_voff(const vbh_matrix* m, hipStream_t s) {
    std::lock_guard<std::mutex> lk(m->host_mu);
    if (!m->h_voff_valid) {
        std::vector<int64_t> h(size_t(m->nnzb + 1));
        VBH_CHECK_HIP(hipMemcpyAsync(h.data(), m->voff, sizeof(int64_t) * h.size(), hipMemcpyDeviceToHost, s));
        VBH_CHECK_HIP(hipStreamSynchronize(s));
        m->h_voff.swap(h);
        m->h_voff_valid = true;
    }
    return m->h_voff;
}

// flag[0] = 1 when a[i] != b[i] for some i < n.
__global__ void k_any_diff(const int* __restrict__ a, const int* __restrict__ b, int64_t n, int* __restrict__ flag) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        if (a[i] != b[i]) flag[0] = 1;
}

bool same_pattern(const vbh_matrix* a, const vbh_matrix* b, hipStream_t s) {
    if (a == b) return true;
    if (a->nbr != b->nbr || a->nbc != b->nbc || a->nnzb != b->nnzb || a->h_row_ptr != b->h_row_ptr ||
        a->h_rbs != b->h_rbs || a->h_cbs != b->h_cbs)
        return false;
    if (a->nnzb == 0) return true;
    DevList tmp;
    int* d_flag = tmp.own((int*)dalloc(sizeof(int)));
    VBH_CHECK_HIP(hipMemsetAsync(d_flag, 0, sizeof(int), s));
    const unsigned nb = unsigned(std::min<int64_t>((a->nnzb + 255) / 256, 8192));
    hipLaunchKernelGGL(k_any_diff, dim3(nb), dim3(256), 0, s, a->col_ind, b->col_ind, a->nnzb, d_flag);
    VBH_CHECK_HIP(hipGetLastError());
    int flag = 0;
    VBH_CHECK_HIP(hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, s));
    VBH_CHECK_HIP(hipStreamSynchronize(s));
    return flag == 0;
}

// Rows of at most kSortMax columns are sorted in LDS (bitonic network over
// the next power of two, INT_MAX padding); already sorted rows are left as
// they are; a longer unsorted row sets long_rows[i] = 1 for the host.  One
// work-group per row.
constexpr int kSortMax = 4096;
__global__ __launch_bounds__(256) void k_sort_rows(const int* __restrict__ row_ptr, int* __restrict__ col,
                                                   int* __restrict__ long_rows) {
    __shared__ int sv[kSortMax];
    __shared__ int unsorted;
    const int i = blockIdx.x;
    const int kb = row_ptr[i], n = row_ptr[i + 1] - kb;
    if (n < 2) return;
    if (threadIdx.x == 0) unsorted = 0;
    __syncthreads();
    for (int t = threadIdx.x; t + 1 < n; t += 256)
        if (col[kb + t] > col[kb + t + 1]) unsorted = 1;
    __syncthreads();
    if (!unsorted) return;
    if (n > kSortMax) {
        if (threadIdx.x == 0) long_rows[i] = 1;
        return;
    }
    int P = 2;
    while (P < n) P <<= 1;
    for (int t = threadIdx.x; t < P; t += 256) sv[t] = t < n ? col[kb + t] : INT_MAX;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < P; t += 256) {
                const int u = t ^ j;
                if (u > t) {
                    const int a = sv[t], b = sv[u];
                    if ((a > b) == ((t & k) == 0)) {
                        sv[t] = b;
                        sv[u] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int t = threadIdx.x; t < n; t += 256) col[kb + t] = sv[t];
}

void sort_rows_dev(int nbr, const int* row_ptr, const int* d_row_ptr, int* d_col_ind, hipStream_t s) {
    if (nbr <= 0 || row_ptr[nbr] == 0) return;
    DevList tmp;
    int* d_long = tmp.own((int*)dalloc(sizeof(int) * size_t(nbr)));
    VBH_CHECK_HIP(hipMemsetAsync(d_long, 0, sizeof(int) * size_t(nbr), s));
    hipLaunchKernelGGL(k_sort_rows, dim3(nbr), dim3(256), 0, s, d_row_ptr, d_col_ind, d_long);
    VBH_CHECK_HIP(hipGetLastError());
    std::vector<int> lr(nbr);
    VBH_CHECK_HIP(hipMemcpyAsync(lr.data(), d_long, sizeof(int) * size_t(nbr), hipMemcpyDeviceToHost, s));
    VBH_CHECK_HIP(hipStreamSynchronize(s));
    if (std::find(lr.begin(), lr.end(), 1) == lr.end()) return;
    // rows longer than the LDS sort: on the host
    const int64_t nnzb = row_ptr[nbr];
    std::vector<int> ci(static_cast<size_t>(nnzb));
    VBH_CHECK_HIP(hipMemcpyAsync(ci.data(), d_col_ind, sizeof(int) * size_t(nnzb), hipMemcpyDeviceToHost, s));
    VBH_CHECK_HIP(hipStreamSynchronize(s));
    rows_parallel(nbr, row_ptr, [&](int r0, int r1) {
        for (int i = r0; i < r1; ++i)
            if (lr[i]) std::sort(ci.begin() + row_ptr[i], ci.begin() + row_ptr[i + 1]);
    });
    VBH_CHECK_HIP(hipMemcpyAsync(d_col_ind, ci.data(), sizeof(int) * size_t(nnzb), hipMemcpyHostToDevice, s));
    VBH_CHECK_HIP(hipStreamSynchronize(s));
}

void destroy_matrix(vbh_matrix* m) {
    if (!m) return;
    for (void* p : {(void*)m->rbs, (void*)m->cbs, (void*)m->row_ptr, (void*)m->col_ind, (void*)m->roff,
                    (void*)m->coff, (void*)m->voff, (void*)m->meta, (void*)m->row_of, m->vals, m->vals32,
                    (void*)m->split, (void*)m->t_ptr, (void*)m->vt, (void*)m->eoff, (void*)m->ecol,
                    (void*)m->t_voff, (void*)m->t_rows, (void*)m->t_split})
        dfree(p);
    for (auto& kv : m->row_lists) {
        dfree(kv.second.small_rows);
        dfree(kv.second.big_desc);
    }
    delete m;
}

void ensure_vals32(vbh_matrix* m, hipStream_t s) {
    std::lock_guard<std::mutex> lk(m->lazy_mu);
    if (m->vals32_valid) return;
    if (!m->vals32) {
        m->vals32 = dalloc((m->elem_size() / 2) * (m->nnz + kVecPad));
        VBH_CHECK_HIP(hipMemsetAsync(m->vals32, 0, (m->elem_size() / 2) * (m->nnz + kVecPad), s));
    }
    const int64_t n = m->nnz;
    if (n > 0) {
        if (m->dtype == VBH_F64)
            hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks_for(n, 256)), dim3(256), 0, s,
                               (const double*)m->vals, (float*)m->vals32, n);
        else
            hipLaunchKernelGGL((k_convert<c128, c64>), dim3(blocks_for(n, 256)), dim3(256), 0, s,
                               (const c128*)m->vals, (c64*)m->vals32, n);
        VBH_CHECK_HIP(hipGetLastError());
    }
    // read by kernels on any stream: complete before it is marked valid
    VBH_CHECK_HIP(hipStreamSynchronize(s));
    m->vals32_valid = true;
}

void values_changed(vbh_matrix* m) {
    std::lock_guard<std::mutex> lk(m->lazy_mu);
    m->vals32_valid = false;
    m->vt_valid = false;
}

// Entry columns of block row i (one wave per row): the stored scalar columns
// of the row in slot order, entry panel + b of a block -> its X row coff + b.
__global__ __launch_bounds__(256) void k_entry_cols(const int* __restrict__ row_ptr, const int2* __restrict__ meta,
                                                    const int64_t* __restrict__ eoff, int nbr, int* __restrict__ ecol) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nbr) return;
    int* out = ecol + eoff[i];
    for (int k = row_ptr[i] + lane; k < row_ptr[i + 1]; k += 64) {
        const int2 mm = meta[k];
        const int c = mm.y & 255, panel = mm.y >> 8;
        for (int b = 0; b < c; ++b) out[panel + b] = mm.x + b;
    }
}

// FP32 values with every block column-major (one wave per block row): block
// k's element (a, b) at vt[voff[k] + b * r + a].
__global__ __launch_bounds__(256) void k_values_colmajor(const double* __restrict__ vals,
                                                         const int64_t* __restrict__ voff, const int* __restrict__ row_ptr,
                                                         const int* __restrict__ rbs, const int* __restrict__ cbs,
                                                         const int* __restrict__ col_ind, int nbr, float* __restrict__ vt) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nbr) return;
    const int r = rbs[i];
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
        const int c = cbs[col_ind[k]];
        const int64_t base = voff[k];
        for (int v = lane; v < r * c; v += 64) {
            const int a = v / c, b = v - a * c;
            vt[base + int64_t(b) * r + a] = float(vals[base + v]);
        }
    }
}

void ensure_brw(vbh_matrix* m, hipStream_t s) {
    std::lock_guard<std::mutex> lk(m->lazy_mu);
    VBH_REQUIRE(m->dtype == VBH_F64, "internal: the block-row wave kernel takes real matrices");
    if (m->vt_valid && m->ecol) return;
    const unsigned grid = unsigned((m->nbr + 3) / 4);
    if (!m->ecol) {
        std::vector<int64_t> eo(size_t(m->nbr) + 1, 0);
        const std::vector<int64_t>& hv = host_voff(m, s);
        for (int i = 0; i < m->nbr; ++i) {
            const int64_t nv = hv[m->h_row_ptr[i + 1]] - hv[m->h_row_ptr[i]];
            eo[i + 1] = eo[i] + (m->h_rbs[i] > 0 ? nv / m->h_rbs[i] : 0);
        }
        int64_t* d_eo = upload(eo, s);
        int* d_ec = nullptr;
        try {
            d_ec = (int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(eo[m->nbr], 1)));
            if (m->nbr > 0) {
                hipLaunchKernelGGL(k_entry_cols, dim3(grid), dim3(256), 0, s, m->row_ptr, m->meta, d_eo, m->nbr,
                                   d_ec);
                VBH_CHECK_HIP(hipGetLastError());
            }
            VBH_CHECK_HIP(hipStreamSynchronize(s));
        } catch (...) {
            dfree(d_eo);
            dfree(d_ec);
            throw;
        }
        m->eoff = d_eo;
        m->ecol = d_ec;
    }
    if (!m->vt) {
        m->vt = (float*)dalloc(sizeof(float) * size_t(m->nnz + kVecPad));
        VBH_CHECK_HIP(hipMemsetAsync(m->vt, 0, sizeof(float) * size_t(m->nnz + kVecPad), s));
    }
    if (m->nbr > 0) {
        hipLaunchKernelGGL(k_values_colmajor, dim3(grid), dim3(256), 0, s, (const double*)m->vals, m->voff,
                           m->row_ptr, m->rbs, m->cbs, m->col_ind, m->nbr, m->vt);
        VBH_CHECK_HIP(hipGetLastError());
    }
    // read by kernels on any stream: complete before it is marked valid
    VBH_CHECK_HIP(hipStreamSynchronize(s));
    m->vt_valid = true;
}

bool has_vals32(vbh_matrix* m) {
    std::lock_guard<std::mutex> lk(m->lazy_mu);
    return m->vals32_valid;
}

bool has_row_classes(vbh_matrix* m, int t) {
    std::lock_guard<std::mutex> lk(m->lazy_mu);
    return m->row_lists.count(t) != 0;
}

// Column-major "incoming" plan for the adjoint: for every local column j the
// (row, slot) pairs in ascending row order -- one writer per output block, as
// in the reference's adjoint plan (vbcsr_backend.hpp:543-625).
void ensure_adjoint_plan(vbh_matrix* m, hipStream_t s) {
    std::lock_guard<std::mutex> lk(m->lazy_mu);
    if (m->t_valid) return;
    std::vector<int> tp(m->nbc + 1, 0), tr(m->nnzb), ts(m->nnzb);
    const std::vector<int>& hc = host_col_ind(m, s);
    for (int64_t k = 0; k < m->nnzb; ++k) tp[hc[k] + 1]++;
    for (int j = 0; j < m->nbc; ++j) tp[j + 1] += tp[j];
    std::vector<int> fill(tp.begin(), tp.end() - 1);
    for (int i = 0; i < m->nbr; ++i)
        for (int k = m->h_row_ptr[i]; k < m->h_row_ptr[i + 1]; ++k) {
            const int j = hc[k];
            tr[fill[j]] = i;
            ts[fill[j]] = k;
            ++fill[j];
        }
    std::vector<int64_t> work(m->nbc + 1, 0);
    for (int j = 0; j < m->nbc; ++j) {
        int64_t w = 64;
        for (int q = tp[j]; q < tp[j + 1]; ++q) w += int64_t(m->h_rbs[tr[q]]) * m->h_cbs[j] + 16;
        work[j + 1] = work[j] + w;
    }
    int G = int(std::min<int64_t>(std::max<int64_t>((m->nbc + kWaves - 1) / kWaves, 8), 4096));
    G = (G + 7) & ~7;
    std::vector<int> split(G + 1, m->nbc);
    split[0] = 0;
    int col = 0;
    for (int g = 1; g < G; ++g) {
        const int64_t target = work[m->nbc] * g / G;
        while (col < m->nbc && work[col] < target) ++col;
        split[g] = col;
    }
    // per incoming block: its value offset and its block row's (first
    // scalar row, row count), so the adjoint kernel reads them in one step
    const std::vector<int64_t>& hv = host_voff(m, s);
    std::vector<int64_t> tvo(m->nnzb);
    std::vector<int2> trr(m->nnzb);
    for (int64_t q = 0; q < m->nnzb; ++q) {
        tvo[q] = hv[ts[q]];
        trr[q] = make_int2(m->h_roff[tr[q]], m->h_rbs[tr[q]]);
    }
    m->t_ptr = upload(tp, s);
    m->t_voff = upload(tvo, s);
    m->t_rows = upload(trr, s);
    m->t_split = upload(split, s);
    m->tG = G;
    VBH_CHECK_HIP(hipStreamSynchronize(s));
    m->t_valid = true;
}

// Slot of the block (i, want[i]) of every row (want = i when null; -1 when
// absent or want[i] < 0).  One wavefront per row.
__global__ __launch_bounds__(256) void k_diag_slots(const int* __restrict__ row_ptr, const int* __restrict__ col_ind,
                                                    int nbr, const int* __restrict__ want, int* __restrict__ out) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nbr) return;
    const int w = want ? want[i] : i;
    int found = -1;
    if (w >= 0)
        for (int k = row_ptr[i] + lane; k < row_ptr[i + 1]; k += 64)
            if (col_ind[k] == w) found = k;
    for (int d = 32; d >= 1; d >>= 1) found = max(found, __shfl_xor(found, d, 64));
    if (lane == 0) out[i] = found;
}

// Device array of the diagonal slots (k_diag_slots), owned by tmp.
static int* diag_slots(const vbh_matrix* m, const int* diag_col_host, DevList& tmp, hipStream_t s) {
    int* d = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max(m->nbr, 1))));
    const int* want = nullptr;
    std::vector<int> h;
    if (diag_col_host) {
        h.assign(diag_col_host, diag_col_host + m->nbr);
        want = tmp.own(upload(h, s));
    }
    hipLaunchKernelGGL(k_diag_slots, dim3(blocks_for(m->nbr, 4)), dim3(256), 0, s, m->row_ptr, m->col_ind, m->nbr,
                       want, d);
    VBH_CHECK_HIP(hipGetLastError());
    VBH_CHECK_HIP(hipStreamSynchronize(s));  // the upload reads h
    return d;
}

// Filter compaction on the device: block k of row i is kept when its norm
// is >= thr (norm = sqrt(norm2[k]) when squared); counts[i] = kept blocks of
// row i.  One wavefront per row.
__global__ __launch_bounds__(256) void k_keep_count(const int* __restrict__ row_ptr, int nbr,
                                                    const double* __restrict__ norm, int squared, double thr,
                                                    int* __restrict__ counts) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nbr) return;
    int n = 0;
    for (int k = row_ptr[i] + lane; k < row_ptr[i + 1]; k += 64) {
        const double v = squared ? sqrt(norm[k]) : norm[k];
        n += v < thr ? 0 : 1;  // keep >= thr: block_csr.hpp:1411
    }
    for (int m = 32; m >= 1; m >>= 1) n += __shfl_xor(n, m, 64);
    if (lane == 0) counts[i] = n;
}

// Kept blocks of every row, in order: new column indices and source slots.
__global__ __launch_bounds__(256) void k_keep_compact(const int* __restrict__ row_ptr, const int* __restrict__ col_ind,
                                                      int nbr, const double* __restrict__ norm, int squared, double thr,
                                                      const int* __restrict__ new_ptr, int* __restrict__ new_ci,
                                                      int* __restrict__ src) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nbr) return;
    int out = new_ptr[i];
    for (int k0 = row_ptr[i]; k0 < row_ptr[i + 1]; k0 += 64) {
        const int k = k0 + lane;
        bool keep = false;
        if (k < row_ptr[i + 1]) {
            const double v = squared ? sqrt(norm[k]) : norm[k];
            keep = !(v < thr);
        }
        const uint64_t mask = __ballot(keep);
        if (keep) {
            const int pos = out + __popcll(mask & ((uint64_t(1) << lane) - 1));
            new_ci[pos] = col_ind[k];
            src[pos] = k;
        }
        out += __popcll(mask);
    }
}

// Filtered copy of m (blocks whose norm is below thr dropped), or null when
// every block is kept.  The keep test, the compaction and the new pattern's
// layout run on the device; the host receives the per-row counts.
vbh_matrix* filter_by_norms(vbh_matrix* m, double thr, const double* d_norm, bool squared, hipStream_t s) {
    if (m->nnzb == 0) return nullptr;
    DevList tmp;
    int* d_counts = tmp.own((int*)dalloc(sizeof(int) * size_t(m->nbr)));
    const unsigned nb = unsigned((m->nbr + 3) / 4);
    hipLaunchKernelGGL(k_keep_count, dim3(nb), dim3(256), 0, s, m->row_ptr, m->nbr, d_norm, squared ? 1 : 0, thr,
                       d_counts);
    VBH_CHECK_HIP(hipGetLastError());
    std::vector<int> rp(m->nbr + 1, 0);
    VBH_CHECK_HIP(hipMemcpyAsync(rp.data() + 1, d_counts, sizeof(int) * size_t(m->nbr), hipMemcpyDeviceToHost, s));
    VBH_CHECK_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < m->nbr; ++i) rp[i + 1] += rp[i];
    const int64_t total = rp[m->nbr];
    if (total == m->nnzb) return nullptr;
    int* d_src = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(total, 1))));
    int* d_ci = tmp.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(total, 1))));
    if (total > 0) {
        int* d_rp = tmp.own(upload(rp, s));
        hipLaunchKernelGGL(k_keep_compact, dim3(nb), dim3(256), 0, s, m->row_ptr, m->col_ind, m->nbr, d_norm,
                           squared ? 1 : 0, thr, d_rp, d_ci, d_src);
        VBH_CHECK_HIP(hipGetLastError());
    }
    MatrixGuard g(build_matrix_dev(m->nbr, m->nbc, m->h_rbs.data(), m->h_cbs.data(), rp.data(), d_ci, m->dtype, s));
    vbh_matrix* c = g.m;
    if (c->nnzb) {
        dispatch_dtype(m->dtype, [&](auto tag) {
            using T = decltype(tag);
            hipLaunchKernelGGL(k_gather_blocks<T>, dim3(blocks_for(c->nnzb, 4)), dim3(256), 0, s,
                               (const T*)m->vals, m->voff, d_src, (T*)c->vals, c->voff, c->nnzb);
        });
        VBH_CHECK_HIP(hipGetLastError());
        VBH_CHECK_HIP(hipStreamSynchronize(s));
    }
    return g.release();
}

// Transposed pattern: blocks per block column (cnt[j] += 1 per slot of
// column j); each slot placed in its column's segment at fill[j]++ (any
// order: the segments are sorted afterwards); row and column of the slot
// src[q] of every transposed entry (the row by binary search in row_ptr).
__global__ void k_col_counts(const int* __restrict__ col_ind, int64_t nnzb, int* __restrict__ cnt) {
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nnzb; k += int64_t(gridDim.x) * blockDim.x)
        atomicAdd(&cnt[col_ind[k]], 1);
}
__global__ void k_col_scatter(const int* __restrict__ col_ind, int64_t nnzb, int* __restrict__ fill,
                              int* __restrict__ src) {
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nnzb; k += int64_t(gridDim.x) * blockDim.x)
        src[atomicAdd(&fill[col_ind[k]], 1)] = int(k);
}
__global__ void k_slot_row_col(const int* __restrict__ row_ptr, int nbr, const int* __restrict__ col_ind,
                               const int* __restrict__ src, int64_t n, int* __restrict__ row,
                               int* __restrict__ col) {
    for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x) {
        const int k = src[q];
        int lo = 0, hi = nbr - 1;  // last row with row_ptr[row] <= k
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (row_ptr[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        row[q] = lo;
        col[q] = col_ind[k];
    }
}

// Block q of the transpose is source slot src[q] (r x c) transposed (and
// conjugated): out[b * r + a] = conj(in[a * c + b]); one wavefront per block.
template <typename T>
__global__ __launch_bounds__(256) void k_transpose_blocks(const T* __restrict__ in, const int64_t* __restrict__ in_voff,
                                                          const int* __restrict__ src, const int* __restrict__ src_cols,
                                                          const int* __restrict__ in_cbs, T* __restrict__ out,
                                                          const int64_t* __restrict__ out_voff, int64_t nq, int conj) {
    const int64_t q = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= nq) return;
    const int k = src[q];
    const int c = in_cbs[src_cols[q]];
    const int64_t ib = in_voff[k], ob = out_voff[q];
    const int64_t n = out_voff[q + 1] - ob;
    const int r = int(n / c);
    for (int64_t e = lane; e < n; e += 64) {
        const int b = int(e / r), a = int(e - int64_t(b) * r);  // out element (b, a)
        const T v = in[ib + int64_t(a) * c + b];
        out[ob + e] = conj ? Sc<T>::conj(v) : v;
    }
}

}  // namespace vbh

using namespace vbh;

// ============================================================== C ABI
extern "C" {

VBH_API const char* vbh_last_error(void) { return g_last_error.c_str(); }
VBH_API int vbh_version(void) { return 100; }

VBH_API int vbh_device_count(int* count) {
    return guard([&] { VBH_CHECK_HIP(hipGetDeviceCount(count)); });
}
VBH_API int vbh_set_device(int device) {
    return guard([&] { VBH_CHECK_HIP(hipSetDevice(device)); });
}
VBH_API int vbh_malloc(void** ptr, size_t bytes) {
    return guard([&] { *ptr = dalloc(bytes); });
}
VBH_API int vbh_free(void* ptr) {
    return guard([&] { dfree(ptr); });
}
VBH_API int vbh_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream) {
    return guard([&] {
        if (bytes == 0) return;
        const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                : kind == 1 ? hipMemcpyDeviceToHost
                                            : hipMemcpyDeviceToDevice;
        VBH_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream));
        if (kind == 1) VBH_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}
VBH_API int vbh_memset(void* dst, int value, size_t bytes, void* stream) {
    return guard([&] {
        if (bytes) VBH_CHECK_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
    });
}
VBH_API int vbh_stream_sync(void* stream) {
    return guard([&] { VBH_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream)); });
}
VBH_API int vbh_stream_release(void* stream) {
    return guard([&] {
        int dev = 0;
        VBH_CHECK_HIP(hipGetDevice(&dev));
        std::unique_ptr<StreamState> st;
        {
            std::lock_guard<std::mutex> lk(g_states_mu);
            auto it = g_states.find({dev, reinterpret_cast<uintptr_t>(stream)});
            if (it == g_states.end()) return;
            VBH_REQUIRE(t_state != it->second.get(), "vbh_stream_release inside a call on that stream");
            std::lock_guard<std::recursive_mutex> busy(it->second->mu);  // no call in flight on it
            st = std::move(it->second);
            g_states.erase(it);
        }
        release_state(st.get());
    });
}
VBH_API int vbh_matrix_values_changed(vbh_matrix_t m) {
    return guard([&] {
        VBH_REQUIRE(m, "null matrix");
        values_changed(m);
    });
}

VBH_API int vbh_matrix_create(int nbr, int nbc, const int* rbs, const int* cbs, const int* row_ptr,
                              const int* col_ind, int dtype, vbh_matrix_t* out) {
    return guard([&] {
        VBH_REQUIRE(out, "null output handle");
        *out = build_matrix(nbr, nbc, rbs, cbs, row_ptr, col_ind, dtype, nullptr);
    });
}

VBH_API int vbh_matrix_destroy(vbh_matrix_t m) {
    return guard([&] { destroy_matrix(m); });
}

VBH_API int vbh_matrix_info_get(vbh_matrix_t m, vbh_matrix_info* info) {
    return guard([&] {
        VBH_REQUIRE(m && info, "null argument");
        info->nbr = m->nbr;
        info->nbc = m->nbc;
        info->nnzb = m->nnzb;
        info->nnz = m->nnz;
        info->n_rows = m->n_rows;
        info->n_cols = m->n_cols;
        info->dtype = m->dtype;
        info->kind = m->kind;
        info->max_block = m->max_block;
        info->shape_classes = m->n_classes;
    });
}

VBH_API int vbh_matrix_get_pattern(vbh_matrix_t m, int* row_ptr, int* col_ind) {
    return guard([&] {
        VBH_REQUIRE(m, "null matrix");
        std::copy(m->h_row_ptr.begin(), m->h_row_ptr.end(), row_ptr);
        {
            std::lock_guard<std::mutex> lk(m->host_mu);
            if (m->h_cols_valid) {
                std::copy(m->h_col_ind.begin(), m->h_col_ind.end(), col_ind);
                return;
            }
        }
        // a device-made pattern: straight from the device (the pattern is
        // complete there once the handle exists), without keeping a mirror
        if (m->nnzb > 0)
            VBH_CHECK_HIP(hipMemcpy(col_ind, m->col_ind, sizeof(int) * size_t(m->nnzb), hipMemcpyDeviceToHost));
    });
}

VBH_API int vbh_matrix_set_values(vbh_matrix_t m, const void* vals, int location, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m, "null matrix");
        const size_t bytes = m->elem_size() * m->nnz;
        if (bytes)
            VBH_CHECK_HIP(hipMemcpyAsync(m->vals, vals, bytes,
                                         location ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                         (hipStream_t)stream));
        if (!location) VBH_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
        values_changed(m);
    });
}

VBH_API int vbh_matrix_get_values(vbh_matrix_t m, void* vals, int location, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m, "null matrix");
        const size_t bytes = m->elem_size() * m->nnz;
        if (bytes)
            VBH_CHECK_HIP(hipMemcpyAsync(vals, m->vals, bytes,
                                         location ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                         (hipStream_t)stream));
        if (!location) VBH_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}

VBH_API void* vbh_matrix_values(vbh_matrix_t m) { return m ? m->vals : nullptr; }

// A copy of m: the host pattern, device-to-device copies of the device
// pattern arrays and values (nothing is recomputed on the host); the FP32
// copy, row-class lists and adjoint plan are rebuilt on demand.
static vbh_matrix* clone_matrix(const vbh_matrix* m, hipStream_t s) {
    auto* c = new vbh_matrix();
    try {
        c->nbr = m->nbr;
        c->nbc = m->nbc;
        c->nnzb = m->nnzb;
        c->nnz = m->nnz;
        c->n_rows = m->n_rows;
        c->n_cols = m->n_cols;
        c->dtype = m->dtype;
        c->kind = m->kind;
        c->max_block = m->max_block;
        c->n_classes = m->n_classes;
        c->uniform_bs = m->uniform_bs;
        c->h_rbs = m->h_rbs;
        c->h_cbs = m->h_cbs;
        c->h_row_ptr = m->h_row_ptr;
        {
            std::lock_guard<std::mutex> lk(m->host_mu);
            c->h_cols_valid = m->h_cols_valid;
            c->h_voff_valid = m->h_voff_valid;
            if (m->h_cols_valid) c->h_col_ind = m->h_col_ind;
            if (m->h_voff_valid) c->h_voff = m->h_voff;
        }
        c->h_roff = m->h_roff;
        c->h_coff = m->h_coff;
        c->h_bfs = m->h_bfs;
        c->G = m->G;
        auto copy = [&](auto* src, int64_t count) {
            using E = std::remove_pointer_t<decltype(src)>;
            E* d = static_cast<E*>(dalloc(sizeof(E) * size_t(count > 0 ? count : 1)));
            if (count > 0) VBH_CHECK_HIP(hipMemcpyAsync(d, src, sizeof(E) * size_t(count), hipMemcpyDeviceToDevice, s));
            return d;
        };
        c->rbs = copy(m->rbs, m->nbr);
        c->cbs = copy(m->cbs, m->nbc);
        c->row_ptr = copy(m->row_ptr, int64_t(m->nbr) + 1);
        c->col_ind = copy(m->col_ind, m->nnzb);
        c->roff = copy(m->roff, int64_t(m->nbr) + 1);
        c->coff = copy(m->coff, int64_t(m->nbc) + 1);
        c->voff = copy(m->voff, m->nnzb + 1);
        c->meta = copy(m->meta, m->nnzb);
        c->row_of = copy(m->row_of, m->n_rows > 0 ? m->n_rows : 1);
        c->split = copy(m->split, int64_t(m->G) + 1);
        const size_t vbytes = m->elem_size() * size_t(m->nnz + kVecPad);
        c->vals = dalloc(vbytes);
        VBH_CHECK_HIP(hipMemcpyAsync(c->vals, m->vals, vbytes, hipMemcpyDeviceToDevice, s));  // padding included
        VBH_CHECK_HIP(hipStreamSynchronize(s));
    } catch (...) {
        destroy_matrix(c);
        throw;
    }
    return c;
}

VBH_API int vbh_matrix_transpose(vbh_matrix_t m, int conj, vbh_matrix_t* out, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m && out, "null argument");
        hipStream_t s = (hipStream_t)stream;
        // pattern on the device: per-column counts, every slot placed in its
        // column's segment, the segments sorted by slot index (slots ascend
        // with the rows), then each entry's row and block column from its
        // slot; the host only sees the column counts
        const int nbr = m->nbr, nbc = m->nbc;
        const int64_t nnzb = m->nnzb;
        std::vector<int> tp(nbc + 1, 0);
        DevList tt;
        int* d_cnt = tt.own((int*)dalloc(sizeof(int) * size_t(std::max(nbc, 1))));
        int* d_src = tt.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(nnzb, 1))));
        int* d_tr = tt.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(nnzb, 1))));
        int* d_scol = tt.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(nnzb, 1))));
        const unsigned nb = unsigned(std::min<int64_t>((nnzb + 255) / 256, 8192));
        if (nnzb > 0 && nbc > 0) {
            VBH_CHECK_HIP(hipMemsetAsync(d_cnt, 0, sizeof(int) * size_t(nbc), s));
            hipLaunchKernelGGL(k_col_counts, dim3(nb), dim3(256), 0, s, m->col_ind, nnzb, d_cnt);
            VBH_CHECK_HIP(hipGetLastError());
            VBH_CHECK_HIP(hipMemcpyAsync(tp.data() + 1, d_cnt, sizeof(int) * size_t(nbc), hipMemcpyDeviceToHost, s));
            VBH_CHECK_HIP(hipStreamSynchronize(s));
        }
        for (int j = 0; j < nbc; ++j) tp[j + 1] += tp[j];
        if (nnzb > 0) {
            int* d_tp = tt.own(upload(tp, s));
            VBH_CHECK_HIP(hipMemcpyAsync(d_cnt, d_tp, sizeof(int) * size_t(nbc), hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(k_col_scatter, dim3(nb), dim3(256), 0, s, m->col_ind, nnzb, d_cnt, d_src);
            VBH_CHECK_HIP(hipGetLastError());
            sort_rows_dev(nbc, tp.data(), d_tp, d_src, s);
            hipLaunchKernelGGL(k_slot_row_col, dim3(nb), dim3(256), 0, s, m->row_ptr, nbr, m->col_ind, d_src, nnzb,
                               d_tr, d_scol);
            VBH_CHECK_HIP(hipGetLastError());
        }
        MatrixGuard g(build_matrix_dev(nbc, nbr, m->h_cbs.data(), m->h_rbs.data(), tp.data(), d_tr, m->dtype, s));
        vbh_matrix* t = g.m;
        if (t->nnzb) {
            dispatch_dtype(m->dtype, [&](auto tag) {
                using T = decltype(tag);
                hipLaunchKernelGGL(k_transpose_blocks<T>, dim3(blocks_for(t->nnzb, 4)), dim3(256), 0, s,
                                   (const T*)m->vals, m->voff, d_src, d_scol, m->cbs, (T*)t->vals, t->voff, t->nnzb,
                                   conj);
            });
            VBH_CHECK_HIP(hipGetLastError());
            VBH_CHECK_HIP(hipStreamSynchronize(s));
        }
        *out = g.release();
    });
}

VBH_API int vbh_matrix_duplicate(vbh_matrix_t m, vbh_matrix_t* out, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m && out, "null argument");
        *out = clone_matrix(m, (hipStream_t)stream);
    });
}

VBH_API int vbh_matrix_scale(vbh_matrix_t m, const double* alpha, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m && alpha, "null argument");
        hipStream_t s = (hipStream_t)stream;
        if (m->nnz == 0) return;
        dispatch_dtype(m->dtype, [&](auto tag) {
            using T = decltype(tag);
            if constexpr (std::is_same_v<T, double>)
                hipLaunchKernelGGL(k_scale_d, dim3(blocks_for(m->nnz, 256)), dim3(256), 0, s, (double*)m->vals,
                                   m->nnz, alpha[0]);
            else
                hipLaunchKernelGGL(k_scale<T>, dim3(blocks_for(m->nnz, 256)), dim3(256), 0, s, (T*)m->vals,
                                   m->nnz, scalar_from<T>(alpha));
        });
        VBH_CHECK_HIP(hipGetLastError());
        values_changed(m);
    });
}

VBH_API int vbh_matrix_shift(vbh_matrix_t m, const double* alpha, const int* diag_col_host, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m && alpha, "null argument");
        hipStream_t s = (hipStream_t)stream;
        if (m->nbr == 0) return;
        DevList tmp;
        int* d = diag_slots(m, diag_col_host, tmp, s);
        dispatch_dtype(m->dtype, [&](auto tag) {
            using T = decltype(tag);
            hipLaunchKernelGGL(k_shift<T>, dim3(blocks_for(m->nbr, 256)), dim3(256), 0, s, (T*)m->vals, m->voff, d,
                               m->rbs, m->cbs, m->col_ind, m->nbr, scalar_from<T>(alpha));
        });
        VBH_CHECK_HIP(hipGetLastError());
        VBH_CHECK_HIP(hipStreamSynchronize(s));
        values_changed(m);
    });
}

VBH_API int vbh_matrix_block_norms(vbh_matrix_t m, double* norms_dev, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m && norms_dev, "null argument");
        if (m->nnzb == 0) return;
        hipStream_t s = (hipStream_t)stream;
        dispatch_dtype(m->dtype, [&](auto tag) {
            using T = decltype(tag);
            hipLaunchKernelGGL(k_block_norms<T>, dim3(blocks_for(m->nnzb, 4)), dim3(256), 0, s, (const T*)m->vals,
                               m->voff, m->nnzb, norms_dev);
        });
        VBH_CHECK_HIP(hipGetLastError());
    });
}

VBH_API int vbh_matrix_filter(vbh_matrix_t m, double thr, vbh_matrix_t* out, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m && out, "null argument");
        *out = nullptr;
        if (!(thr > 0.0) || m->nnzb == 0) return;  // block_csr.hpp:1400
        hipStream_t s = (hipStream_t)stream;
        double* d_norms = (double*)scratch(sizeof(double) * m->nnzb, 0);
        dispatch_dtype(m->dtype, [&](auto tag) {
            using T = decltype(tag);
            hipLaunchKernelGGL(k_block_norms<T>, dim3(blocks_for(m->nnzb, 4)), dim3(256), 0, s, (const T*)m->vals,
                               m->voff, m->nnzb, d_norms);
        });
        VBH_CHECK_HIP(hipGetLastError());
        *out = filter_by_norms(m, thr, d_norms, false, s);
    });
}

VBH_API int vbh_matrix_axpby(vbh_matrix_t self, const double* alpha, vbh_matrix_t X, const double* beta,
                             vbh_matrix_t* out, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(self && X && alpha && beta && out, "null argument");
        VBH_REQUIRE(self->dtype == X->dtype, "axpby: dtype mismatch");
        VBH_REQUIRE(self->nbr == X->nbr, "Matrix row count mismatch in axpby");
        *out = nullptr;
        hipStream_t s = (hipStream_t)stream;
        const bool same = same_pattern(self, X, s);
        if (same) {
            if (self->nnz == 0) return;
            dispatch_dtype(self->dtype, [&](auto tag) {
                using T = decltype(tag);
                if constexpr (std::is_same_v<T, double>)
                    hipLaunchKernelGGL(k_axpby_flat_d, dim3(blocks_for(self->nnz, 256)), dim3(256), 0, s,
                                       (const double*)X->vals, (double*)self->vals, self->nnz, alpha[0], beta[0]);
                else
                    hipLaunchKernelGGL(k_axpby_flat<T>, dim3(blocks_for(self->nnz, 256)), dim3(256), 0, s,
                                       (const T*)X->vals, (T*)self->vals, self->nnz, scalar_from<T>(alpha),
                                       scalar_from<T>(beta));
            });
            VBH_CHECK_HIP(hipGetLastError());
            values_changed(self);
            return;
        }
        VBH_REQUIRE(self->h_rbs == X->h_rbs, "axpby: block row sizes differ");
        // union pattern, sorted columns (ops/axpby.hpp:417-527)
        const int nbc = std::max(self->nbc, X->nbc);
        std::vector<int> cbs(nbc);
        for (int j = 0; j < nbc; ++j) {
            const int a = j < self->nbc ? self->h_cbs[j] : 0;
            const int b = j < X->nbc ? X->h_cbs[j] : 0;
            VBH_REQUIRE(a == 0 || b == 0 || a == b, "axpby: column block sizes differ");
            cbs[j] = a ? a : b;
        }
        // on the device, two passes: merged lengths, then the merged columns
        // with the source slot of each side (-1: absent)
        const int nbr = self->nbr;
        std::vector<int> rp(nbr + 1, 0);
        DevList ut;
        int* d_cnt = ut.own((int*)dalloc(sizeof(int) * size_t(std::max(nbr, 1))));
        if (nbr > 0) {
            hipLaunchKernelGGL(k_union_rows, dim3(blocks_for(nbr, 4)), dim3(256), 0, s, nbr, self->row_ptr,
                               self->col_ind, X->row_ptr, X->col_ind, (const int*)nullptr, d_cnt, (int*)nullptr,
                               (int*)nullptr, (int*)nullptr);
            VBH_CHECK_HIP(hipGetLastError());
            VBH_CHECK_HIP(hipMemcpyAsync(rp.data() + 1, d_cnt, sizeof(int) * size_t(nbr), hipMemcpyDeviceToHost, s));
            VBH_CHECK_HIP(hipStreamSynchronize(s));
        }
        for (int i = 0; i < nbr; ++i) {
            VBH_REQUIRE(int64_t(rp[i]) + rp[i + 1] <= int64_t(INT32_MAX), "axpby: union exceeds 2^31 blocks");
            rp[i + 1] += rp[i];
        }
        const int64_t total = rp[nbr];
        int* d_ci = ut.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(total, 1))));
        int* d_ys = ut.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(total, 1))));
        int* d_xs = ut.own((int*)dalloc(sizeof(int) * size_t(std::max<int64_t>(total, 1))));
        if (total > 0) {
            int* d_rp = ut.own(upload(rp, s));
            hipLaunchKernelGGL(k_union_rows, dim3(blocks_for(nbr, 4)), dim3(256), 0, s, nbr, self->row_ptr,
                               self->col_ind, X->row_ptr, X->col_ind, (const int*)d_rp, (int*)nullptr, d_ci, d_ys,
                               d_xs);
            VBH_CHECK_HIP(hipGetLastError());
        }
        MatrixGuard g(build_matrix_dev(nbr, nbc, self->h_rbs.data(), cbs.data(), rp.data(), d_ci, self->dtype, s));
        vbh_matrix* c = g.m;
        if (c->nnzb) {
            dispatch_dtype(self->dtype, [&](auto tag) {
                using T = decltype(tag);
                hipLaunchKernelGGL(k_union_axpby<T>, dim3(blocks_for(c->nnzb, 4)), dim3(256), 0, s,
                                   (const T*)self->vals, self->voff, d_ys, (const T*)X->vals, X->voff, d_xs,
                                   (T*)c->vals, c->voff, c->nnzb, scalar_from<T>(alpha), scalar_from<T>(beta));
            });
            VBH_CHECK_HIP(hipGetLastError());
            VBH_CHECK_HIP(hipStreamSynchronize(s));
        }
        *out = g.release();
    });
}

VBH_API int vbh_matrix_trace(vbh_matrix_t m, double* out, void* stream) {
    return guard_on(stream, [&] {
        VBH_REQUIRE(m && out, "null argument");
        hipStream_t s = (hipStream_t)stream;
        out[0] = 0.0;
        if (m->nbr == 0) return;
        DevList tmp;
        int* d = diag_slots(m, nullptr, tmp, s);
        const int nb = 64;
        double* part = (double*)scratch(sizeof(double) * nb, 1);
        dispatch_dtype(m->dtype, [&](auto tag) {
            using T = decltype(tag);
            hipLaunchKernelGGL(k_trace<T>, dim3(nb), dim3(256), 0, s, (const T*)m->vals, m->voff, d, m->rbs, m->nbr,
                               part);
        });
        VBH_CHECK_HIP(hipGetLastError());
        std::vector<double> h(nb);
        VBH_CHECK_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
        VBH_CHECK_HIP(hipStreamSynchronize(s));
        double t = 0.0;
        for (double v : h) t += v;
        out[0] = t;
    });
}

}  // extern "C"
