// Test-infrastructure only: a thin C driver around the reference vbcsr C++
// core (header-only, compiled from /root/reference/vbcsr/core by
// oracle/Makefile into oracle/_ref/libvbcsr_ref.so).  It exists so the tests
// can compare this repository's CPU restatement (oracle/vbcsr_oracle.cpp) and
// its HIP path against the reference's own CPU output, and so bench.py can time
// the reference CPU path as the baseline.  Nothing in the product links it.
//
// Every entry point runs the reference in serial mode (MPI_COMM_NULL, one
// rank), which is the single-rank local share the GPU path models.
//
// Reference interfaces driven here:
//   DistGraph::construct_serial      vbcsr/core/dist_graph.hpp:167
//   BlockSpMat::add_block / assemble vbcsr/core/block_csr.hpp:613, :861
//   BlockSpMat::mult / mult_dense    vbcsr/core/block_csr.hpp:985, :1011
//   BlockSpMat::mult_adjoint(_dense) vbcsr/core/block_csr.hpp:1024, :1040
//   BlockSpMat::spmm                 vbcsr/core/block_csr.hpp:1530
//   BlockSpMat::filter_blocks        vbcsr/core/block_csr.hpp:1399
//   BlockSpMat::axpby/scale/shift    vbcsr/core/block_csr.hpp:1250, :1053, :1289
//   DistMultiVector::set_random      vbcsr/core/dist_multivector.hpp:335
#include "block_csr.hpp"
#include "dist_multivector.hpp"
#include "dist_vector.hpp"
#include "random_vectors.hpp"

#include <chrono>
#include <complex>
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

using namespace vbcsr;

namespace {

thread_local std::string g_err;

template <typename T>
struct RefMat {
    DistGraph* graph = nullptr;  // owned when non-null (input graphs)
    BlockSpMat<T> mat;
    ~RefMat() {
        mat.reset();  // release the matrix's graph reference first
        delete graph;
    }
};

template <typename T>
RefMat<T>* create(int nbr, const int* bs, const int* row_ptr, const int* col_ind) {
    auto* h = new RefMat<T>();
    h->graph = new DistGraph(MPI_COMM_NULL);
    std::vector<int> sizes(bs, bs + nbr);
    std::vector<std::vector<int>> adj(nbr);
    for (int i = 0; i < nbr; ++i)
        adj[i].assign(col_ind + row_ptr[i], col_ind + row_ptr[i + 1]);
    h->graph->construct_serial(nbr, sizes, adj);
    h->mat = BlockSpMat<T>(h->graph);
    return h;
}

template <typename T>
void set_values(RefMat<T>* h, const T* vals) {
    auto& g = *h->mat.graph;
    const int nbr = static_cast<int>(g.adj_ptr.size()) - 1;
    size_t off = 0;
    for (int i = 0; i < nbr; ++i) {
        const int r = g.block_sizes[i];
        for (int k = g.adj_ptr[i]; k < g.adj_ptr[i + 1]; ++k) {
            const int lc = g.adj_ind[k];
            const int c = g.block_sizes[lc];
            h->mat.add_block(g.get_global_index(i), g.get_global_index(lc), vals + off, r, c,
                             AssemblyMode::INSERT, MatrixLayout::RowMajor);
            off += static_cast<size_t>(r) * c;
        }
    }
    h->mat.assemble();
}

template <typename T>
void mult(RefMat<T>* h, const T* x, T* y) {
    DistVector<T> vx(h->mat.graph), vy(h->mat.graph);
    std::memcpy(vx.data.data(), x, sizeof(T) * vx.local_size);
    h->mat.mult(vx, vy);
    std::memcpy(y, vy.data.data(), sizeof(T) * vy.local_size);
}

template <typename T>
void mult_adjoint(RefMat<T>* h, const T* x, T* y) {
    DistVector<T> vx(h->mat.graph), vy(h->mat.graph);
    std::memcpy(vx.data.data(), x, sizeof(T) * vx.local_size);
    h->mat.mult_adjoint(vx, vy);
    std::memcpy(y, vy.data.data(), sizeof(T) * vy.local_size);
}

template <typename T>
void load_mv(DistMultiVector<T>& mv, const T* X, int K) {
    for (int r = 0; r < mv.local_rows; ++r)
        for (int v = 0; v < K; ++v) mv(r, v) = X[static_cast<size_t>(r) * K + v];
}

template <typename T>
void store_mv(const DistMultiVector<T>& mv, T* Y, int K) {
    for (int r = 0; r < mv.local_rows; ++r)
        for (int v = 0; v < K; ++v) Y[static_cast<size_t>(r) * K + v] = mv(r, v);
}

template <typename T>
void mult_dense(RefMat<T>* h, int K, const T* X, T* Y, bool adjoint) {
    DistMultiVector<T> mx(h->mat.graph, K), my(h->mat.graph, K);
    load_mv(mx, X, K);
    if (adjoint) h->mat.mult_dense_adjoint(mx, my);
    else h->mat.mult_dense(mx, my);
    store_mv(my, Y, K);
}

template <typename T>
double time_mult_dense(RefMat<T>* h, int K, const T* X, int iters) {
    DistMultiVector<T> mx(h->mat.graph, K), my(h->mat.graph, K);
    load_mv(mx, X, K);
    h->mat.mult_dense(mx, my);  // warm-up: builds the cached apply plan
    const auto t0 = std::chrono::steady_clock::now();
    for (int it = 0; it < iters; ++it) h->mat.mult_dense(mx, my);
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count() / iters;
}

template <typename T>
double time_mult(RefMat<T>* h, const T* x, int iters) {
    DistVector<T> vx(h->mat.graph), vy(h->mat.graph);
    std::memcpy(vx.data.data(), x, sizeof(T) * vx.local_size);
    h->mat.mult(vx, vy);
    const auto t0 = std::chrono::steady_clock::now();
    for (int it = 0; it < iters; ++it) h->mat.mult(vx, vy);
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count() / iters;
}

// The KPM recursion of examples/ex08_graphene_dos.py:95-114 composed from the
// reference's own C++ ops (mult_dense, scale, axpy, bdot), one moment per
// matvec.  moments[n] = sum_v Re <v0_v, T_n(H) v0_v> (not yet normalised).
template <typename T>
double kpm_moments(RefMat<T>* h, int K, const T* V0, int M, double* moments) {
    auto* g = h->mat.graph;
    DistMultiVector<T> v0(g, K), vprev(g, K), vcur(g, K), vnext(g, K);
    load_mv(v0, V0, K);
    const auto t0 = std::chrono::steady_clock::now();
    auto sum_re = [&](const std::vector<T>& d) {
        double s = 0;
        for (auto& x : d) s += std::real(x);
        return s;
    };
    vprev.copy_from(v0);
    moments[0] = sum_re(v0.bdot(v0));
    if (M > 1) {
        h->mat.mult_dense(v0, vcur);
        moments[1] = sum_re(v0.bdot(vcur));
    }
    for (int n = 2; n < M; ++n) {
        h->mat.mult_dense(vcur, vnext);
        vnext.scale(T(2.0));
        vnext.axpy(T(-1.0), vprev);
        moments[n] = sum_re(v0.bdot(vnext));
        vprev.swap(vcur);
        vcur.swap(vnext);
    }
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

template <typename T>
RefMat<T>* spmm(RefMat<T>* a, RefMat<T>* b, double thr) {
    auto* h = new RefMat<T>();
    h->mat = a->mat.spmm(b->mat, thr, false, false);
    return h;
}

template <typename T>
RefMat<T>* duplicate(RefMat<T>* a) {
    auto* h = new RefMat<T>();
    h->mat = a->mat.duplicate(true);
    return h;
}

template <typename T>
void get_pattern(RefMat<T>* h, int* row_ptr, int* col_ind) {
    auto& g = *h->mat.graph;
    const int nbr = static_cast<int>(g.adj_ptr.size()) - 1;
    for (int i = 0; i <= nbr; ++i) row_ptr[i] = g.adj_ptr[i];
    for (int i = 0; i < nbr; ++i)
        for (int k = g.adj_ptr[i]; k < g.adj_ptr[i + 1]; ++k)
            col_ind[k] = g.get_global_index(g.adj_ind[k]);
}

template <typename T>
void get_values(RefMat<T>* h, T* out) {
    auto v = h->mat.get_values(MatrixLayout::RowMajor);
    std::memcpy(out, v.data(), sizeof(T) * v.size());
}

template <typename T>
void random_mv(RefMat<T>* h, int K, uint64_t seed, int normalize, T* out) {
    DistMultiVector<T> mv(h->mat.graph, K);
    mv.set_random(RandomKind::Rademacher, seed, normalize != 0);
    store_mv(mv, out, K);
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // namespace

using cplx = std::complex<double>;

#define REF_API extern "C" __attribute__((visibility("default")))

REF_API const char* ref_last_error() { return g_err.c_str(); }

REF_API uint64_t ref_entry_hash(uint64_t seed, uint64_t block, uint64_t row, uint64_t col) {
    return vbcsr::detail::entry_hash(seed, block, row, col);
}

#define REF_TYPED(SFX, T)                                                                       \
    REF_API void* ref_create_##SFX(int nbr, const int* bs, const int* rp, const int* ci) {      \
        void* out = nullptr;                                                                    \
        if (guarded([&] { out = create<T>(nbr, bs, rp, ci); })) return nullptr;                 \
        return out;                                                                             \
    }                                                                                           \
    REF_API void ref_free_##SFX(void* h) { delete static_cast<RefMat<T>*>(h); }                 \
    REF_API int ref_set_values_##SFX(void* h, const void* v) {                                  \
        return guarded([&] { set_values(static_cast<RefMat<T>*>(h), static_cast<const T*>(v)); }); \
    }                                                                                           \
    REF_API long ref_nnzb_##SFX(void* h) {                                                      \
        return static_cast<long>(static_cast<RefMat<T>*>(h)->mat.local_block_nnz());            \
    }                                                                                           \
    REF_API long ref_nnz_##SFX(void* h) {                                                       \
        return static_cast<long>(static_cast<RefMat<T>*>(h)->mat.local_scalar_nnz());          \
    }                                                                                           \
    REF_API const char* ref_kind_##SFX(void* h) {                                               \
        static thread_local std::string s;                                                      \
        s = static_cast<RefMat<T>*>(h)->mat.matrix_kind_string();                               \
        return s.c_str();                                                                       \
    }                                                                                           \
    REF_API void ref_get_pattern_##SFX(void* h, int* rp, int* ci) {                             \
        get_pattern(static_cast<RefMat<T>*>(h), rp, ci);                                        \
    }                                                                                           \
    REF_API void ref_get_values_##SFX(void* h, void* out) {                                     \
        get_values(static_cast<RefMat<T>*>(h), static_cast<T*>(out));                           \
    }                                                                                           \
    REF_API int ref_mult_##SFX(void* h, const void* x, void* y) {                               \
        return guarded([&] { mult(static_cast<RefMat<T>*>(h), static_cast<const T*>(x), static_cast<T*>(y)); }); \
    }                                                                                           \
    REF_API int ref_mult_adjoint_##SFX(void* h, const void* x, void* y) {                       \
        return guarded([&] { mult_adjoint(static_cast<RefMat<T>*>(h), static_cast<const T*>(x), static_cast<T*>(y)); }); \
    }                                                                                           \
    REF_API int ref_mult_dense_##SFX(void* h, int K, const void* X, void* Y) {                  \
        return guarded([&] { mult_dense(static_cast<RefMat<T>*>(h), K, static_cast<const T*>(X), static_cast<T*>(Y), false); }); \
    }                                                                                           \
    REF_API int ref_mult_dense_adjoint_##SFX(void* h, int K, const void* X, void* Y) {          \
        return guarded([&] { mult_dense(static_cast<RefMat<T>*>(h), K, static_cast<const T*>(X), static_cast<T*>(Y), true); }); \
    }                                                                                           \
    REF_API double ref_time_mult_dense_##SFX(void* h, int K, const void* X, int iters) {        \
        double t = -1;                                                                          \
        guarded([&] { t = time_mult_dense(static_cast<RefMat<T>*>(h), K, static_cast<const T*>(X), iters); }); \
        return t;                                                                               \
    }                                                                                           \
    REF_API double ref_time_mult_##SFX(void* h, const void* x, int iters) {                     \
        double t = -1;                                                                          \
        guarded([&] { t = time_mult(static_cast<RefMat<T>*>(h), static_cast<const T*>(x), iters); }); \
        return t;                                                                               \
    }                                                                                           \
    REF_API double ref_kpm_moments_##SFX(void* h, int K, const void* V0, int M, double* mom) {  \
        double t = -1;                                                                          \
        guarded([&] { t = kpm_moments(static_cast<RefMat<T>*>(h), K, static_cast<const T*>(V0), M, mom); }); \
        return t;                                                                               \
    }                                                                                           \
    REF_API void* ref_spmm_##SFX(void* a, void* b, double thr) {                                \
        void* out = nullptr;                                                                    \
        if (guarded([&] { out = spmm(static_cast<RefMat<T>*>(a), static_cast<RefMat<T>*>(b), thr); })) return nullptr; \
        return out;                                                                             \
    }                                                                                           \
    REF_API void* ref_duplicate_##SFX(void* a) {                                                \
        void* out = nullptr;                                                                    \
        if (guarded([&] { out = duplicate(static_cast<RefMat<T>*>(a)); })) return nullptr;      \
        return out;                                                                             \
    }                                                                                           \
    REF_API int ref_filter_##SFX(void* h, double thr) {                                         \
        return guarded([&] { static_cast<RefMat<T>*>(h)->mat.filter_blocks(thr); });            \
    }                                                                                           \
    REF_API int ref_scale_##SFX(void* h, const void* a) {                                       \
        return guarded([&] { static_cast<RefMat<T>*>(h)->mat.scale(*static_cast<const T*>(a)); }); \
    }                                                                                           \
    REF_API int ref_shift_##SFX(void* h, const void* a) {                                       \
        return guarded([&] { static_cast<RefMat<T>*>(h)->mat.shift(*static_cast<const T*>(a)); }); \
    }                                                                                           \
    REF_API int ref_axpby_##SFX(void* h, const void* alpha, void* x, const void* beta) {        \
        return guarded([&] {                                                                    \
            static_cast<RefMat<T>*>(h)->mat.axpby(*static_cast<const T*>(alpha),                \
                                                  static_cast<RefMat<T>*>(x)->mat,              \
                                                  *static_cast<const T*>(beta));                \
        });                                                                                     \
    }                                                                                           \
    REF_API int ref_block_norms_##SFX(void* h, double* out) {                                   \
        return guarded([&] {                                                                    \
            const auto& n = static_cast<RefMat<T>*>(h)->mat.get_block_norms();                  \
            std::memcpy(out, n.data(), sizeof(double) * n.size());                              \
        });                                                                                     \
    }                                                                                           \
    REF_API int ref_random_mv_##SFX(void* h, int K, uint64_t seed, int normalize, void* out) {  \
        return guarded([&] { random_mv(static_cast<RefMat<T>*>(h), K, seed, normalize, static_cast<T*>(out)); }); \
    }

REF_TYPED(d, double)
REF_TYPED(z, cplx)
