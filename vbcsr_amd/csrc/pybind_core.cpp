// vbcsr_amd._core -- the host side of the drop-in: a pybind11 module with the
// class and method names of the reference's vbcsr_core module
// (vbcsr/pybind_vbcsr.cpp:407-505) whose objects keep their data in HBM and
// call the HIP library through its C ABI (include/vbcsr_hip.h).
//
// One rank's local share: DistGraph holds owned block rows [0, n_owned) and
// ghost columns after them (dist_graph.hpp:50-54); there is no communicator,
// so sync_ghosts / reduce_ghosts are no-ops exactly as the reference's are at
// one rank (dist_vector.hpp:225, dist_multivector.hpp:400).
//
// Data residency: every vector, multivector and matrix owns a device buffer
// and a host mirror with a three-state flag (device newer / host newer /
// both current).  Device operations run on the device copy, uploading first
// when the host copy is newer; a buffer-protocol export (numpy view,
// to_numpy) synchronises the host copy and marks it newer, so writes through
// the view reach the device at the next device operation.
#include <fstream>
#include <iomanip>
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <iostream>
#include <memory>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/vbcsr_hip.h"

namespace py = pybind11;
using cplx = std::complex<double>;

namespace {

void check(int rc, const char* what) {
    if (rc == VBH_OK) return;
    std::string msg = std::string(what) + ": " + vbh_last_error();
    if (rc == VBH_ERR_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

template <typename T>
constexpr int dtype_of() {
    return std::is_same_v<T, double> ? VBH_F64 : VBH_C128;
}

template <typename T>
void scalar2(T v, double out[2]) {
    if constexpr (std::is_same_v<T, double>) {
        out[0] = v;
        out[1] = 0.0;
    } else {
        out[0] = v.real();
        out[1] = v.imag();
    }
}

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t b) : bytes(b) {
        check(vbh_malloc(&ptr, std::max<size_t>(b, 16)), "device allocation");
        check(vbh_memset(ptr, 0, std::max<size_t>(b, 16), nullptr), "device memset");
    }
    ~DeviceBuffer() {
        if (ptr) vbh_free(ptr);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

// ------------------------------------------------------------------ DistGraph
struct Graph {
    int rank = 0, size = 1;
    std::vector<int> owned_global_indices, ghost_global_indices, block_sizes, adj_ptr, adj_ind, block_offsets;
    std::unordered_map<int, int> g2l;
    // device metadata for the counter-based random fill (owned blocks)
    std::shared_ptr<DeviceBuffer> d_bs, d_gids, d_offs;

    int n_owned() const { return int(owned_global_indices.size()); }
    int owned_scalar_rows() const { return block_offsets.empty() ? 0 : block_offsets[n_owned()]; }
    int local_scalar_cols() const { return block_offsets.empty() ? 0 : block_offsets.back(); }

    void finish() {
        block_offsets.assign(block_sizes.size() + 1, 0);
        for (size_t i = 0; i < block_sizes.size(); ++i) block_offsets[i + 1] = block_offsets[i] + block_sizes[i];
        d_bs.reset();
        d_gids.reset();
        d_offs.reset();
    }

    // Local construction shared by the public constructors: owned blocks in
    // the given (sorted) order, then ghosts sorted by global id, rows sorted
    // and de-duplicated.
    void build(const std::vector<int>& owned, const std::vector<int>& sizes,
               const std::vector<std::vector<int>>& adj, const std::vector<int>& ghosts,
               const std::vector<int>& ghost_sizes) {
        if (!std::is_sorted(owned.begin(), owned.end()))
            throw std::runtime_error("DistGraph::construct_distributed requires sorted owned_global_indices");
        if (sizes.size() != owned.size() || adj.size() != owned.size())
            throw std::runtime_error("DistGraph: owned indices, block sizes and adjacency must have equal length");
        owned_global_indices = owned;
        g2l.clear();
        block_sizes = sizes;
        for (int i = 0; i < int(owned.size()); ++i) g2l[owned[i]] = i;
        std::vector<std::pair<int, int>> gh;
        for (size_t q = 0; q < ghosts.size(); ++q) gh.push_back({ghosts[q], ghost_sizes[q]});
        std::sort(gh.begin(), gh.end());
        ghost_global_indices.clear();
        for (auto& [g, s] : gh) {
            if (g2l.count(g)) continue;
            g2l[g] = int(block_sizes.size());
            ghost_global_indices.push_back(g);
            block_sizes.push_back(s);
        }
        adj_ptr.assign(owned.size() + 1, 0);
        adj_ind.clear();
        std::vector<int> row;
        for (size_t i = 0; i < owned.size(); ++i) {
            row.clear();
            for (int g : adj[i]) {
                auto it = g2l.find(g);
                if (it == g2l.end())
                    throw std::runtime_error("DistGraph: block column " + std::to_string(g) +
                                             " is neither owned nor a declared ghost (one rank: pass its size "
                                             "with construct_local)");
                row.push_back(it->second);
            }
            std::sort(row.begin(), row.end());
            row.erase(std::unique(row.begin(), row.end()), row.end());
            adj_ind.insert(adj_ind.end(), row.begin(), row.end());
            adj_ptr[i + 1] = int(adj_ind.size());
        }
        for (int s : block_sizes)
            if (s <= 0) throw std::runtime_error("DistGraph: block sizes must be positive");
        finish();
    }

    void construct_serial(int n, const std::vector<int>& sizes, const std::vector<std::vector<int>>& adj) {
        if (int(sizes.size()) != n)
            throw std::runtime_error("DistGraph::construct_serial received invalid block size metadata on root");
        if (int(adj.size()) != n)
            throw std::runtime_error("DistGraph::construct_serial received invalid adjacency metadata on root");
        std::vector<int> owned(n);
        std::iota(owned.begin(), owned.end(), 0);
        build(owned, sizes, adj, {}, {});
    }

    int local_index(int gid) const {
        auto it = g2l.find(gid);
        return it == g2l.end() ? -1 : it->second;
    }
    int global_index(int lid) const {
        if (lid < 0 || lid >= int(block_sizes.size())) throw std::out_of_range("local index out of range");
        return lid < n_owned() ? owned_global_indices[lid] : ghost_global_indices[lid - n_owned()];
    }
    bool same_owned(const Graph& o) const {
        return owned_global_indices == o.owned_global_indices &&
               std::equal(block_sizes.begin(), block_sizes.begin() + n_owned(), o.block_sizes.begin());
    }

    void ensure_device_meta() {
        if (d_bs) return;
        const int n = n_owned();
        std::vector<int64_t> offs(n + 1);
        for (int i = 0; i <= n; ++i) offs[i] = block_offsets[i];
        d_bs = std::make_shared<DeviceBuffer>(sizeof(int) * std::max(n, 1));
        d_gids = std::make_shared<DeviceBuffer>(sizeof(int) * std::max(n, 1));
        d_offs = std::make_shared<DeviceBuffer>(sizeof(int64_t) * (n + 1));
        if (n) {
            check(vbh_memcpy(d_bs->ptr, block_sizes.data(), sizeof(int) * n, 0, nullptr), "upload");
            check(vbh_memcpy(d_gids->ptr, owned_global_indices.data(), sizeof(int) * n, 0, nullptr), "upload");
        }
        check(vbh_memcpy(d_offs->ptr, offs.data(), sizeof(int64_t) * (n + 1), 0, nullptr), "upload");
        check(vbh_stream_sync(nullptr), "sync");
    }
};
using GraphP = std::shared_ptr<Graph>;

// ------------------------------------------------------------------ mirrored storage
enum class Fresh { Both, Device, Host };

template <typename T>
struct Mirrored {
    std::shared_ptr<DeviceBuffer> dev;
    std::vector<T> host;
    Fresh fresh = Fresh::Both;
    size_t count = 0;

    void allocate(size_t n) {
        count = n;
        dev = std::make_shared<DeviceBuffer>(sizeof(T) * n);
        host.assign(n, T(0));
        fresh = Fresh::Both;
    }
    void* dptr() {
        to_device();
        return dev->ptr;
    }
    void to_device() {
        if (fresh == Fresh::Host) {
            if (count) check(vbh_memcpy(dev->ptr, host.data(), sizeof(T) * count, 0, nullptr), "upload");
            fresh = Fresh::Both;
        }
    }
    void to_host() {
        if (fresh == Fresh::Device) {
            if (count) check(vbh_memcpy(host.data(), dev->ptr, sizeof(T) * count, 1, nullptr), "download");
            fresh = Fresh::Both;
        }
    }
    void device_written() { fresh = Fresh::Device; }
    void host_written() { fresh = Fresh::Host; }
};

// ------------------------------------------------------------------ vectors
template <typename T>
struct MultiVec {
    GraphP graph;
    int num_vectors = 1, ld = 1, local_rows = 0, ghost_rows = 0;
    bool is_vector = false;  // DistVector: contiguous, ld 1
    Mirrored<T> buf;

    static int compute_ld(int k) {  // dist_multivector.hpp:63-67
        const int per_line = 64 / int(sizeof(T));
        return ((k + per_line - 1) / per_line) * per_line;
    }
    virtual ~MultiVec() = default;
    MultiVec(GraphP g, int k, bool vec) : graph(std::move(g)), num_vectors(k), is_vector(vec) {
        if (k < 0) throw std::invalid_argument("number of vectors must be non-negative");
        ld = vec ? 1 : compute_ld(k);
        local_rows = graph->owned_scalar_rows();
        ghost_rows = graph->local_scalar_cols() - local_rows;
        buf.allocate(size_t(local_rows + ghost_rows) * ld);
    }
    int rows() const { return local_rows + ghost_rows; }
    void* d() { return buf.dptr(); }

    void require_compatible(const MultiVec& o, const char* what) const {
        if (o.local_rows != local_rows || o.num_vectors != num_vectors)
            throw std::runtime_error(std::string(what) + ": size mismatch");
    }
    std::shared_ptr<MultiVec> make_like() const;
    std::shared_ptr<MultiVec> duplicate() const {
        auto c = make_like();
        c->buf.host = buf.host;
        c->buf.fresh = buf.fresh == Fresh::Device ? Fresh::Device : Fresh::Host;
        if (buf.fresh != Fresh::Host && buf.count)
            check(vbh_memcpy(c->buf.dev->ptr, buf.dev->ptr, sizeof(T) * buf.count, 2, nullptr), "copy");
        if (buf.fresh == Fresh::Both) c->buf.fresh = Fresh::Both;
        return c;
    }
    void copy_from(MultiVec& o) {
        if (o.local_rows != local_rows || o.num_vectors != num_vectors)
            throw std::runtime_error(is_vector ? "Vector owned-size mismatch in copy_from"
                                               : "DistMultiVector::copy_from: size mismatch");
        const size_t n = (o.buf.count == buf.count) ? buf.count : size_t(local_rows) * ld;
        check(vbh_memcpy(d(), o.d(), sizeof(T) * n, 2, nullptr), "copy");
        buf.device_written();
    }
    void set_constant(T val) {
        double v[2];
        scalar2(val, v);
        // DistVector fills owned entries (dist_vector.hpp:73-76); DistMultiVector
        // fills owned and ghost rows (dist_multivector.hpp:170-177).
        const int64_t r = is_vector ? local_rows : rows();
        check(vbh_vec_fill(dtype_of<T>(), r, num_vectors, v, d(), ld, nullptr), "set_constant");
        buf.device_written();
    }
    void scale(T alpha) {
        double a[2];
        scalar2(alpha, a);
        const int64_t r = is_vector ? local_rows : rows();  // DistMultiVector scales the whole buffer
        check(vbh_vec_scale(dtype_of<T>(), r, num_vectors, a, d(), ld, nullptr), "scale");
        buf.device_written();
    }
    void axpby(T alpha, MultiVec& x, T beta) {
        require_compatible(x, "axpby");
        double a[2], b[2];
        scalar2(alpha, a);
        scalar2(beta, b);
        check(vbh_vec_axpby(dtype_of<T>(), local_rows, num_vectors, a, x.d(), x.ld, b, d(), ld, nullptr), "axpby");
        buf.device_written();
    }
    void axpy(T alpha, MultiVec& x) { axpby(alpha, x, T(1)); }
    void pointwise(MultiVec& x, bool x_vector) {
        const int64_t r = x_vector ? rows() : local_rows;
        if (x_vector && x.rows() < r) throw std::runtime_error("pointwise_mult: vector too short");
        if (!x_vector) require_compatible(x, "pointwise_mult");
        check(vbh_vec_pointwise(dtype_of<T>(), r, num_vectors, x.d(), x.ld, x_vector ? 1 : 0, d(), ld, nullptr),
              "pointwise_mult");
        buf.device_written();
    }
    std::vector<T> bdot(MultiVec& x) {
        if (x.num_vectors != num_vectors) throw std::runtime_error("Dimension mismatch in bdot");
        std::vector<double> out(2 * std::max(num_vectors, 1));
        check(vbh_vec_bdot(dtype_of<T>(), local_rows, num_vectors, d(), ld, x.d(), x.ld, out.data(), nullptr), "bdot");
        std::vector<T> r(num_vectors);
        for (int v = 0; v < num_vectors; ++v) {
            if constexpr (std::is_same_v<T, double>) r[v] = out[2 * v];
            else r[v] = T(out[2 * v], out[2 * v + 1]);
        }
        return r;
    }
    // Counter-based draw (random_vectors.hpp:62-161): Rademacher for real
    // scalars, random phase for complex; owned rows only, then (normalize)
    // every row times 1/sqrt(N_global) (dist_multivector.hpp:335-372).
    void set_random(uint64_t seed, bool normalize) {
        graph->ensure_device_meta();
        const int kind = std::is_same_v<T, double> ? 0 : 1;
        const double n_global = double(graph->owned_scalar_rows());
        check(vbh_vec_set_random(dtype_of<T>(), graph->n_owned(), (const int*)graph->d_bs->ptr,
                                 (const int*)graph->d_gids->ptr, (const int64_t*)graph->d_offs->ptr, num_vectors, seed,
                                 kind, 1.0, d(), ld, nullptr),
              "set_random");
        buf.device_written();
        if (normalize && n_global > 0) {
            const double f = 1.0 / std::sqrt(n_global);
            if (is_vector) {
                scale(T(f));
            } else {
                double a[2] = {f, 0.0};
                check(vbh_vec_scale(dtype_of<T>(), rows(), num_vectors, a, d(), ld, nullptr), "scale");
            }
        }
    }
    void zero_ghost_tail() {
        if (ghost_rows == 0) return;
        const double z[2] = {0.0, 0.0};
        check(vbh_vec_fill(dtype_of<T>(), ghost_rows, num_vectors, z, (char*)d() + sizeof(T) * size_t(local_rows) * ld,
                           ld, nullptr),
              "zero ghosts");
    }
    py::buffer_info buffer() {
        buf.to_host();
        buf.host_written();  // the exported view may be written
        if (is_vector)
            return py::buffer_info(buf.host.data(), sizeof(T), py::format_descriptor<T>::format(), 1,
                                   {size_t(rows())}, {sizeof(T)});
        return py::buffer_info(buf.host.data(), sizeof(T), py::format_descriptor<T>::format(), 2,
                               {size_t(rows()), size_t(num_vectors)}, {sizeof(T) * size_t(ld), sizeof(T)});
    }
};

// The two Python-visible vector classes (DistVector_*, DistMultiVector_*).
template <typename T>
struct DistVec : MultiVec<T> {
    explicit DistVec(GraphP g) : MultiVec<T>(std::move(g), 1, true) {}
};
template <typename T>
struct DistMVec : MultiVec<T> {
    DistMVec(GraphP g, int k) : MultiVec<T>(std::move(g), k, false) {}
};
template <typename T>
std::shared_ptr<MultiVec<T>> MultiVec<T>::make_like() const {
    if (is_vector) return std::make_shared<DistVec<T>>(graph);
    return std::make_shared<DistMVec<T>>(graph, num_vectors);
}

enum class AssemblyMode { INSERT = 0, ADD = 1 };
enum class RedistOp { Copy = 0, Sum = 1 };

// ------------------------------------------------------------------ matrix
template <typename T>
struct Mat {
    GraphP graph;
    vbh_matrix_t h = nullptr;  // owns the device values
    Fresh fresh = Fresh::Both;
    std::vector<T> host;
    // value offset of every slot (and the total), computed on first host use:
    // products and sums that stay on the device never build it
    std::vector<int64_t> voff_;
    bool voff_ok = false;
    int64_t nnz_ = 0;
    uint32_t page_size_ = 0;

    explicit Mat(GraphP g) : graph(std::move(g)) { create_from_graph(); }
    // The host mirror of a device-made matrix is allocated on first use
    // (to_host): products and sums of large matrices never touch it.
    Mat(GraphP g, vbh_matrix_t handle) : graph(std::move(g)), h(handle) {
        nnz_ = info().nnz;
        fresh = Fresh::Device;
    }
    ~Mat() {
        if (h) vbh_matrix_destroy(h);
    }
    Mat(const Mat&) = delete;

    void create_from_graph() {
        const int n = graph->n_owned();
        std::vector<int> rbs(graph->block_sizes.begin(), graph->block_sizes.begin() + n);
        check(vbh_matrix_create(n, int(graph->block_sizes.size()), rbs.data(), graph->block_sizes.data(),
                                graph->adj_ptr.data(), graph->adj_ind.data(), dtype_of<T>(), &h),
              "matrix create");
        compute_voff();
        host.assign(nnz(), T(0));
        fresh = Fresh::Both;
    }
    void compute_voff() {
        const int n = graph->n_owned();
        const std::vector<int>& rp = graph->adj_ptr;
        const std::vector<int>& ci = graph->adj_ind;
        const std::vector<int>& bs = graph->block_sizes;
        std::vector<int64_t>& voff = voff_;
        voff.assign(ci.size() + 1, 0);
        // per-row value counts, their prefix, then every slot's offset; the
        // row passes run on up to 16 host threads for large patterns
        std::vector<int64_t> rowv(size_t(n) + 1, 0);
        auto rows = [&](auto&& fn) {
            const int nt = int(std::min<int64_t>(std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency())),
                                                 std::max<int64_t>(1, int64_t(ci.size()) >> 18)));
            if (nt <= 1 || n < 2 * nt) {
                fn(0, n);
                return;
            }
            std::vector<std::thread> th;
            for (int t = 0; t < nt; ++t) {
                const int r0 = int(int64_t(n) * t / nt), r1 = int(int64_t(n) * (t + 1) / nt);
                th.emplace_back([&fn, r0, r1] { fn(r0, r1); });
            }
            for (auto& x : th) x.join();
        };
        rows([&](int r0, int r1) {
            for (int i = r0; i < r1; ++i) {
                int64_t v = 0;
                for (int k = rp[i]; k < rp[i + 1]; ++k) v += int64_t(bs[i]) * bs[ci[k]];
                rowv[i + 1] = v;
            }
        });
        for (int i = 0; i < n; ++i) rowv[i + 1] += rowv[i];
        rows([&](int r0, int r1) {
            for (int i = r0; i < r1; ++i) {
                int64_t v = rowv[i];
                for (int k = rp[i]; k < rp[i + 1]; ++k) {
                    v += int64_t(bs[i]) * bs[ci[k]];
                    voff[k + 1] = v;
                }
            }
        });
        nnz_ = voff.back();
        voff_ok = true;
    }
    const std::vector<int64_t>& voff() {
        if (!voff_ok) compute_voff();
        return voff_;
    }
    int64_t nnz() const { return nnz_; }
    int64_t nnzb() const { return int64_t(graph->adj_ind.size()); }
    vbh_matrix_t dev() {
        if (fresh == Fresh::Host) {
            check(vbh_matrix_set_values(h, host.data(), 0, nullptr), "upload values");
            fresh = Fresh::Both;
        }
        return h;
    }
    // Invariant: host holds nnz() values unless fresh == Device.
    void to_host() {
        if (fresh == Fresh::Device) {
            if (int64_t(host.size()) != nnz()) host.assign(nnz(), T(0));
            check(vbh_matrix_get_values(h, host.data(), 0, nullptr), "download values");
            fresh = Fresh::Both;
        }
    }
    // values written on the device (by a kernel or a device copy): the host
    // copy is stale, and so is any FP32 copy the library derived from them
    void device_written() {
        fresh = Fresh::Device;
        if (h) check(vbh_matrix_values_changed(h), "values changed");
    }
    void host_written() { fresh = Fresh::Host; }

    vbh_matrix_info info() {
        vbh_matrix_info i{};
        check(vbh_matrix_info_get(h, &i), "info");
        return i;
    }
    std::string kind() {
        const int k = info().kind;
        return k == 0 ? "csr" : (k == 1 ? "bsr" : "vbcsr");
    }

    int find_slot(int lrow, int lcol) const {
        const auto b = graph->adj_ind.begin() + graph->adj_ptr[lrow];
        const auto e = graph->adj_ind.begin() + graph->adj_ptr[lrow + 1];
        auto it = std::lower_bound(b, e, lcol);
        if (it == e || *it != lcol) return -1;
        return int(it - graph->adj_ind.begin());
    }

    // add_block: block_csr.hpp:613-717 (local part), update_local_block :2960-3015
    void add_block(int grow, int gcol, py::array_t<T, py::array::c_style | py::array::forcecast> data,
                   AssemblyMode amode) {
        const int mode = amode == AssemblyMode::INSERT ? 0 : 1;
        if (data.ndim() != 2) throw std::runtime_error("add_block: array must be 2D");
        const int lrow = graph->local_index(grow);
        const int lcol = graph->local_index(gcol);
        int slot = -1;
        if (lrow >= 0 && lrow < graph->n_owned() && lcol >= 0) slot = find_slot(lrow, lcol);
        if (slot < 0) {
            std::cerr << "Warning: Block (row=" << grow << ", col=" << gcol
                      << ") not found in local graph. Ignoring." << std::endl;
            return;
        }
        const int r = graph->block_sizes[lrow], c = graph->block_sizes[lcol];
        if (data.shape(0) != r || data.shape(1) != c)
            throw std::runtime_error("update_local_block dimension mismatch: row " + std::to_string(lrow) +
                                     " expects " + std::to_string(r) + " (got " + std::to_string(data.shape(0)) +
                                     "), col " + std::to_string(lcol) + " expects " + std::to_string(c) + " (got " +
                                     std::to_string(data.shape(1)) + ")");
        to_host();
        T* dst = host.data() + voff()[slot];
        const T* src = data.data();
        if (mode == 0) std::memcpy(dst, src, sizeof(T) * r * c);  // INSERT
        else
            for (int e = 0; e < r * c; ++e) dst[e] += src[e];
        host_written();
    }

    template <typename V>
    void bind_check(const V& x, const char* what) const {
        if (x.graph.get() != graph.get() && !x.graph->same_owned(*graph))
            throw std::runtime_error(std::string(what) +
                                     ": graph mismatch - target graph has a different owned block structure "
                                     "(partition or block sizes)");
        if (x.rows() != graph->local_scalar_cols())
            throw std::runtime_error(std::string(what) + ": vector does not cover the matrix's local columns");
    }

    void mult(MultiVec<T>& x, MultiVec<T>& y, int precision = 0) {
        bind_check(x, "mult");
        bind_check(y, "mult");
        if (x.num_vectors != y.num_vectors) throw std::runtime_error("mult_dense: X and Y column counts differ");
        vbh_matrix_t m = dev();
        void* xp = x.d();
        void* yp = y.d();
        {
            py::gil_scoped_release nogil;
            if (x.is_vector) check(vbh_spmv(m, xp, yp, precision, nullptr), "mult");
            else check(vbh_spmm(m, x.num_vectors, xp, x.ld, yp, y.ld, precision, nullptr), "mult_dense");
        }
        y.buf.device_written();
        y.zero_ghost_tail();
    }
    void mult_adjoint(MultiVec<T>& x, MultiVec<T>& y) {
        bind_check(x, "mult_adjoint");
        bind_check(y, "mult_adjoint");
        vbh_matrix_t m = dev();
        void* xp = x.d();
        void* yp = y.d();
        {
            py::gil_scoped_release nogil;
            if (x.is_vector) check(vbh_spmv_adjoint(m, xp, yp, 0, nullptr), "mult_adjoint");
            else check(vbh_spmm_adjoint(m, x.num_vectors, xp, x.ld, yp, y.ld, 0, nullptr), "mult_dense_adjoint");
        }
        y.buf.device_written();
    }

    void scale(T alpha) {
        double a[2];
        scalar2(alpha, a);
        check(vbh_matrix_scale(dev(), a, nullptr), "scale");
        device_written();
    }
    void shift(T alpha) {
        double a[2];
        scalar2(alpha, a);
        const int n = graph->n_owned();
        std::vector<int> diag(n, -1);
        for (int i = 0; i < n; ++i) diag[i] = graph->local_index(graph->owned_global_indices[i]);
        check(vbh_matrix_shift(dev(), a, diag.data(), nullptr), "shift");
        device_written();
    }
    void fill(T value) {
        double v[2];
        scalar2(value, v);
        check(vbh_vec_fill(dtype_of<T>(), nnz(), 1, v, vbh_matrix_values(dev()), 1, nullptr), "fill");
        device_written();
    }
    // block_csr.hpp:1328-1360: H_ii += v_i on the diagonal blocks.
    void add_diagonal(MultiVec<T>& v) {
        v.buf.to_host();
        to_host();
        const int n = graph->n_owned();
        for (int i = 0; i < n; ++i) {
            const int slot = find_slot(i, i);
            if (slot < 0) continue;
            const int r = graph->block_sizes[i];
            for (int t = 0; t < r; ++t) host[voff()[slot] + int64_t(t) * r + t] += v.buf.host[graph->block_offsets[i] + t];
        }
        host_written();
    }
    void conjugate() {
        if constexpr (std::is_same_v<T, cplx>) {
            to_host();
            for (auto& x : host) x = std::conj(x);
            host_written();
        }
    }
    void copy_from(Mat& o) {
        if (o.graph->adj_ptr != graph->adj_ptr || o.graph->adj_ind != graph->adj_ind ||
            o.graph->block_sizes != graph->block_sizes)
            throw std::runtime_error("Incompatible graph structure in copy_from");
        check(vbh_memcpy(vbh_matrix_values(h), vbh_matrix_values(o.dev()), sizeof(T) * nnz(), 2, nullptr), "copy");
        device_written();
    }
    std::shared_ptr<Mat> duplicate(bool independent_graph) {
        GraphP g = independent_graph ? std::make_shared<Graph>(*graph) : graph;
        vbh_matrix_t c = nullptr;
        check(vbh_matrix_duplicate(dev(), &c, nullptr), "duplicate");
        auto out = std::make_shared<Mat>(g, c);
        return out;
    }
    // Graph of a new pattern over the same owned rows (and the same column
    // space): used by spmm / filter / union axpby results.
    GraphP graph_for(vbh_matrix_t m) const { return graph_for_pattern(m, *graph); }
    void adopt(vbh_matrix_t m) {
        GraphP g = graph_for(m);
        if (h) vbh_matrix_destroy(h);
        h = m;
        graph = g;
        voff_ok = false;
        std::vector<int64_t>().swap(voff_);
        nnz_ = info().nnz;
        std::vector<T>().swap(host);
        fresh = Fresh::Device;
    }
    void axpby(T alpha, Mat& X, T beta) {
        if (this == &X) {  // block_csr.hpp:1251-1254
            scale(alpha + beta);
            return;
        }
        if (!X.graph->same_owned(*graph)) throw std::runtime_error("Matrix row count mismatch in axpby");
        double a[2], b[2];
        scalar2(alpha, a);
        scalar2(beta, b);
        vbh_matrix_t out = nullptr;
        check(vbh_matrix_axpby(dev(), a, X.dev(), b, &out, nullptr), "axpby");
        if (out) adopt(out);
        else device_written();
    }
    // alpha*X + beta*self as a new matrix, self unchanged: a union of two
    // patterns is built once, without first duplicating self
    std::shared_ptr<Mat> combined(T alpha, Mat& X, T beta) {
        const bool same_pattern = this == &X || (X.graph->adj_ptr == graph->adj_ptr &&
                                                 X.graph->adj_ind == graph->adj_ind &&
                                                 X.graph->block_sizes == graph->block_sizes);
        if (same_pattern) {
            auto c = duplicate(true);
            c->axpby(alpha, this == &X ? *c : X, beta);
            return c;
        }
        if (!X.graph->same_owned(*graph)) throw std::runtime_error("Matrix row count mismatch in axpby");
        double a[2], b[2];
        scalar2(alpha, a);
        scalar2(beta, b);
        vbh_matrix_t out = nullptr;
        check(vbh_matrix_axpby(dev(), a, X.dev(), b, &out, nullptr), "axpby");
        if (!out) throw std::runtime_error("axpby: expected a union result");
        return std::make_shared<Mat>(graph_for(out), out);
    }
    void filter_blocks(double thr) {
        vbh_matrix_t out = nullptr;
        check(vbh_matrix_filter(dev(), thr, &out, nullptr), "filter_blocks");
        if (out) adopt(out);
    }
    std::shared_ptr<Mat> spmm(Mat& B, double thr, bool transA, bool transB, int numeric = 0) {
        if (transA) return transpose()->spmm(B, thr, false, transB, numeric);
        if (transB) return spmm(*B.transpose(), thr, false, false, numeric);
        if (graph->ghost_global_indices.size() || B.graph->ghost_global_indices.size())
            throw std::runtime_error("spmm: ghost columns need the distributed exchange (one rank: none allowed)");
        vbh_matrix_t out = nullptr;
        vbh_matrix_t a = dev(), b = B.dev();
        {
            py::gil_scoped_release nogil;
            check(vbh_spgemm(a, b, thr, numeric, &out, last_stats, nullptr), "spmm");
        }
        auto res = std::make_shared<Mat>(graph_for_pattern(out, *graph), out);
        return res;
    }
    // MatrixMarket coordinate export of the local matrix (block_csr.hpp:
    // 2058-2131): header, "rows cols entries", then every stored scalar in
    // slot order, row-major within a block, 1-based, values in scientific
    // notation with 16 digits (complex: real and imaginary part).  One rank.
    void save_matrix_market(const std::string& filename) {
        std::ofstream file(filename);
        if (!file.is_open()) throw std::runtime_error("Could not open file for writing: " + filename);
        to_host();
        const std::vector<int64_t>& vo = voff();
        constexpr bool cx = std::is_same_v<T, cplx>;
        file << "%%MatrixMarket matrix coordinate " << (cx ? "complex" : "real") << " general\n";
        const int n = graph->n_owned();
        size_t nnz_e = 0;
        for (int i = 0; i < n; ++i)
            for (int k = graph->adj_ptr[i]; k < graph->adj_ptr[i + 1]; ++k)
                nnz_e += size_t(graph->block_sizes[i]) * graph->block_sizes[graph->adj_ind[k]];
        file << graph->owned_scalar_rows() << " " << graph->local_scalar_cols() << " " << nnz_e << "\n";
        file << std::scientific << std::setprecision(16);
        for (int i = 0; i < n; ++i) {
            const int r = graph->block_sizes[i];
            const int row1 = graph->block_offsets[i] + 1;
            for (int k = graph->adj_ptr[i]; k < graph->adj_ptr[i + 1]; ++k) {
                const int j = graph->adj_ind[k];
                const int c = graph->block_sizes[j];
                const int col1 = graph->block_offsets[j] + 1;
                for (int a = 0; a < r; ++a)
                    for (int b = 0; b < c; ++b) {
                        const T v = host[vo[k] + int64_t(a) * c + b];
                        file << (row1 + a) << " " << (col1 + b) << " ";
                        if constexpr (cx) file << v.real() << " " << v.imag();
                        else file << v;
                        file << "\n";
                    }
            }
        }
    }
    // spmm_hermitian (block_csr.hpp:1597-1601): the upper block triangle of
    // op(A) B, then its Hermitian completion; spmm_hermitian_upper alone
    // (:1605-1616) and complete_hermitian (:1913-1920).
    std::shared_ptr<Mat> spmm_hermitian(Mat& B, double thr, bool transA, int numeric = 0) {
        if (transA) return transpose()->spmm_hermitian(B, thr, false, numeric);
        return herm_call(B, thr, numeric, vbh_spgemm_hermitian, "spmm_hermitian");
    }
    std::shared_ptr<Mat> spmm_hermitian_upper(Mat& B, double thr, bool transA, int numeric = 0) {
        if (transA) return transpose()->spmm_hermitian_upper(B, thr, false, numeric);
        return herm_call(B, thr, numeric, vbh_spgemm_upper, "spmm_hermitian_upper");
    }
    std::shared_ptr<Mat> complete_hermitian() {
        if (graph->ghost_global_indices.size())
            throw std::runtime_error("complete_hermitian: ghost columns need the distributed exchange");
        vbh_matrix_t out = nullptr;
        vbh_matrix_t a = dev();
        {
            py::gil_scoped_release nogil;
            check(vbh_matrix_complete_hermitian(a, &out, nullptr), "complete_hermitian");
        }
        return std::make_shared<Mat>(graph_for_pattern(out, *graph), out);
    }
    template <typename F>
    std::shared_ptr<Mat> herm_call(Mat& B, double thr, int numeric, F fn, const char* what) {
        if (graph->ghost_global_indices.size() || B.graph->ghost_global_indices.size())
            throw std::runtime_error(std::string(what) + ": ghost columns need the distributed exchange");
        vbh_matrix_t out = nullptr;
        vbh_matrix_t a = dev(), b = B.dev();
        {
            py::gil_scoped_release nogil;
            check(fn(a, b, thr, numeric, &out, last_stats, nullptr), what);
        }
        return std::make_shared<Mat>(graph_for_pattern(out, *graph), out);
    }
    static GraphP graph_for_pattern(vbh_matrix_t m, const Graph& rows_of) {
        vbh_matrix_info i{};
        check(vbh_matrix_info_get(m, &i), "info");
        // rows_of's distribution with m's adjacency (rows_of's is not copied)
        auto g = std::make_shared<Graph>();
        g->rank = rows_of.rank;
        g->size = rows_of.size;
        g->owned_global_indices = rows_of.owned_global_indices;
        g->ghost_global_indices = rows_of.ghost_global_indices;
        g->block_sizes = rows_of.block_sizes;
        g->g2l = rows_of.g2l;
        g->adj_ptr.assign(i.nbr + 1, 0);
        g->adj_ind.assign(i.nnzb, 0);
        check(vbh_matrix_get_pattern(m, g->adj_ptr.data(), g->adj_ind.data()), "pattern");
        g->finish();
        return g;
    }
    // Conjugate transpose on the device (vbh_matrix_transpose): square local
    // structure without ghosts.
    std::shared_ptr<Mat> transpose() {
        if (graph->ghost_global_indices.size())
            throw std::runtime_error("transpose: ghost columns need the distributed exchange");
        vbh_matrix_t out = nullptr;
        vbh_matrix_t a = dev();
        {
            py::gil_scoped_release nogil;
            check(vbh_matrix_transpose(a, std::is_same_v<T, cplx> ? 1 : 0, &out, nullptr), "transpose");
        }
        return std::make_shared<Mat>(graph_for_pattern(out, *graph), out);
    }
    py::array_t<T> get_values() {
        to_host();
        py::array_t<T> out(nnz());
        std::memcpy(out.mutable_data(), host.data(), sizeof(T) * nnz());
        return out;
    }
    py::array_t<T> get_block(int lrow, int lcol) {
        const int slot = (lrow >= 0 && lrow < graph->n_owned()) ? find_slot(lrow, lcol) : -1;
        if (slot < 0) return py::array_t<T>();
        to_host();
        const int r = graph->block_sizes[lrow], c = graph->block_sizes[lcol];
        py::array_t<T> out({r, c});
        std::memcpy(out.mutable_data(), host.data() + voff()[slot], sizeof(T) * r * c);
        return out;
    }
    py::array_t<T> to_dense() {
        to_host();
        const std::vector<int64_t>& vo = voff();
        const int n = graph->n_owned();
        const int rows = graph->owned_scalar_rows(), cols = graph->local_scalar_cols();
        py::array_t<T> out({rows, cols});
        T* o = out.mutable_data();
        std::fill(o, o + int64_t(rows) * cols, T(0));
        for (int i = 0; i < n; ++i)
            for (int k = graph->adj_ptr[i]; k < graph->adj_ptr[i + 1]; ++k) {
                const int j = graph->adj_ind[k];
                const int r = graph->block_sizes[i], c = graph->block_sizes[j];
                for (int a = 0; a < r; ++a)
                    for (int b = 0; b < c; ++b)
                        o[int64_t(graph->block_offsets[i] + a) * cols + graph->block_offsets[j] + b] =
                            host[vo[k] + int64_t(a) * c + b];
            }
        return out;
    }
    void from_dense(py::array_t<T, py::array::c_style | py::array::forcecast> a) {
        const int rows = graph->owned_scalar_rows(), cols = graph->local_scalar_cols();
        if (a.ndim() != 2 || a.shape(0) != rows || a.shape(1) != cols)
            throw std::runtime_error("from_dense: expected shape (" + std::to_string(rows) + ", " +
                                     std::to_string(cols) + ")");
        to_host();
        const std::vector<int64_t>& vo = voff();
        const T* src = a.data();
        for (int i = 0; i < graph->n_owned(); ++i)
            for (int k = graph->adj_ptr[i]; k < graph->adj_ptr[i + 1]; ++k) {
                const int j = graph->adj_ind[k];
                const int r = graph->block_sizes[i], c = graph->block_sizes[j];
                for (int aa = 0; aa < r; ++aa)
                    for (int b = 0; b < c; ++b)
                        host[vo[k] + int64_t(aa) * c + b] =
                            src[int64_t(graph->block_offsets[i] + aa) * cols + graph->block_offsets[j] + b];
            }
        host_written();
    }
    int64_t last_stats[4] = {0, 0, 0, 0};
};

template <typename T>
using MatP = std::shared_ptr<Mat<T>>;
template <typename T>
using VecP = std::shared_ptr<MultiVec<T>>;

// KPM moments of one matrix (extension of the drop-in for the native recursion).
template <typename T>
py::array_t<double> kpm_moments(Mat<T>& H, MultiVec<T>& V0, int M, int mode, double center, double half, int precision) {
    if (V0.local_rows != H.graph->owned_scalar_rows()) throw std::runtime_error("kpm_moments: V0 size mismatch");
    py::array_t<double> out(std::max(M, 0));
    vbh_matrix_t m = H.dev();
    const int K = V0.num_vectors;
    std::vector<double> acc(std::max(M, 0), 0.0), part(std::max(M, 0));
    for (int v0 = 0; v0 < K; v0 += 64) {
        const int kk = std::min(64, K - v0);
        const char* p = (const char*)V0.d() + sizeof(T) * v0;
        {
            py::gil_scoped_release nogil;
            check(vbh_kpm_moments(m, kk, p, V0.ld, M, mode, center, half, precision, part.data(), nullptr),
                  "kpm_moments");
        }
        for (int n = 0; n < M; ++n) acc[n] += part[n];
    }
    std::memcpy(out.mutable_data(), acc.data(), sizeof(double) * acc.size());
    return out;
}

template <typename T>
py::array_t<double> kpm_moments_batched(std::vector<MatP<T>> Hs, std::vector<VecP<T>> V0s, int M, int mode,
                                        double center, double half, int precision, bool graph) {
    const int n = int(Hs.size());
    if (n == 0 || int(V0s.size()) != n) throw std::invalid_argument("kpm_moments_batched: need equal, non-empty lists");
    const int K = V0s[0]->num_vectors;
    std::vector<vbh_matrix_t> ms(n);
    std::vector<const void*> vs(n);
    std::vector<int> lds(n);
    for (int h = 0; h < n; ++h) {
        if (V0s[h]->num_vectors != K) throw std::invalid_argument("kpm_moments_batched: equal column counts needed");
        if (V0s[h]->local_rows != Hs[h]->graph->owned_scalar_rows())
            throw std::invalid_argument("kpm_moments_batched: V0 size mismatch");
        ms[h] = Hs[h]->dev();
        vs[h] = V0s[h]->d();
        lds[h] = V0s[h]->ld;
    }
    py::array_t<double> out({n, std::max(M, 0)});
    {
        py::gil_scoped_release nogil;
        check(vbh_kpm_moments_batched(n, ms.data(), K, vs.data(), lds.data(), M, mode, center, half, precision,
                                      graph ? 1 : 0, out.mutable_data(), nullptr),
              "kpm_moments_batched");
    }
    return out;
}

template <typename T>
VecP<T> cheb_apply(Mat<T>& H, MultiVec<T>& X, py::array_t<cplx, py::array::c_style | py::array::forcecast> coef,
                   double center, double half, int precision) {
    std::shared_ptr<MultiVec<T>> Y = X.make_like();
    vbh_matrix_t m = H.dev();
    void* xp = X.d();
    void* yp = Y->d();
    const int nt = int(coef.size());
    const double* cp = reinterpret_cast<const double*>(coef.data());
    {
        py::gil_scoped_release nogil;
        check(vbh_cheb_apply(m, X.num_vectors, xp, X.ld, yp, Y->ld, nt, cp, center, half, precision, nullptr),
              "cheb_apply");
    }
    Y->buf.device_written();
    return Y;
}

template <typename T, typename D>
void bind_vec(py::module& m, const std::string& name, bool is_vector) {
    using V = MultiVec<T>;
    py::class_<D, V, std::shared_ptr<D>> c(m, name.c_str(), py::buffer_protocol());
    if constexpr (std::is_same_v<D, DistVec<T>>) c.def(py::init([](GraphP g) { return std::make_shared<D>(g); }));
    else c.def(py::init([](GraphP g, int k) { return std::make_shared<D>(g, k); }));
    c.def("sync_ghosts", [](V&) {})
        .def("reduce_ghosts", [](V&) {})
        .def("set_constant", &V::set_constant)
        .def("scale", &V::scale)
        .def("axpy", &V::axpy)
        .def("axpby", &V::axpby)
        .def("duplicate", &V::duplicate)
        .def("copy_from", &V::copy_from)
        .def("set_random", &V::set_random, py::arg("seed"), py::arg("normalize") = true)
        .def("reset_comm_stats", [](V&) {})
        .def_property_readonly("comm_seconds", [](V&) { return 0.0; })
        .def_property_readonly("comm_calls", [](V&) { return 0L; })
        .def_property_readonly("device_ptr", [](V& v) { return reinterpret_cast<uintptr_t>(v.d()); })
        .def_property_readonly("ld", [](V& v) { return v.ld; })
        .def_property_readonly("is_complex", [](V&) { return std::is_same_v<T, cplx>; })
        .def("mark_device_written", [](V& v) { v.buf.device_written(); })
        .def_buffer([](D& v) { return v.buffer(); });
    if (is_vector) {
        c.def("pointwise_mult", [](V& self, V& o) { self.pointwise(o, false); })
            .def("dot", [](V& self, V& o) { return self.bdot(o)[0]; })
            .def_property_readonly("local_size", [](V& v) { return v.local_rows; })
            .def_property_readonly("ghost_size", [](V& v) { return v.ghost_rows; })
            .def_property_readonly("full_size", [](V& v) { return v.rows(); });
    } else {
        c.def("pointwise_mult", [](V& self, V& o) { self.pointwise(o, false); })
            .def("pointwise_mult_vec", [](V& self, V& o) { self.pointwise(o, true); })
            .def("bdot", &V::bdot)
            .def_property_readonly("local_rows", [](V& v) { return v.local_rows; })
            .def_property_readonly("ghost_rows", [](V& v) { return v.ghost_rows; })
            .def_property_readonly("num_vectors", [](V& v) { return v.num_vectors; });
    }
}

template <typename T>
void bind_mat(py::module& m, const std::string& name) {
    using M = Mat<T>;
    using V = MultiVec<T>;
    py::class_<M, std::shared_ptr<M>>(m, name.c_str())
        .def(py::init([](GraphP g) { return std::make_shared<M>(g); }))
        .def_property_readonly("graph", [](M& self) { return self.graph; })
        .def("add_block", &M::add_block, py::arg("g_row"), py::arg("g_col"), py::arg("data"),
             py::arg("mode") = AssemblyMode::ADD)
        .def("assemble", [](M& self) { self.dev(); })
        .def("mult", [](M& self, V& x, V& y) { self.mult(x, y, 0); })
        .def("mult_dense", [](M& self, V& x, V& y) { self.mult(x, y, 0); })
        .def("mult_adjoint", &M::mult_adjoint)
        .def("mult_dense_adjoint", &M::mult_adjoint)
        // precision 0: FP64 values; 2: the FP32 value copy against these FP64 vectors
        // (precision 1 needs FP32 vector buffers: spmm_raw)
        .def("mult_precision",
             [](M& self, V& x, V& y, int precision) {
                 if (precision != 0 && precision != 2)
                     throw std::invalid_argument("mult_precision: precision must be 0 or 2 for these vectors");
                 self.mult(x, y, precision);
             })
        .def("scale", &M::scale)
        .def("conjugate", &M::conjugate)
        .def("shift", &M::shift)
        .def("add_diagonal", &M::add_diagonal)
        .def("axpy", [](M& self, T alpha, M& X) { self.axpby(alpha, X, T(1)); })
        .def("axpby", &M::axpby)
        .def("combined", &M::combined, py::arg("alpha"), py::arg("X"), py::arg("beta"))
        .def("copy_from", &M::copy_from)
        .def("fill", &M::fill)
        .def("duplicate", &M::duplicate, py::arg("independent_graph") = true)
        .def("spmm", &M::spmm, py::arg("B"), py::arg("threshold"), py::arg("transA") = false,
             py::arg("transB") = false, py::arg("numeric") = 0)
        .def("spmm_hermitian", &M::spmm_hermitian, py::arg("B"), py::arg("threshold"), py::arg("transA") = false,
             py::arg("numeric") = 0)
        .def("spmm_hermitian_upper", &M::spmm_hermitian_upper, py::arg("B"), py::arg("threshold"),
             py::arg("transA") = false, py::arg("numeric") = 0)
        .def("complete_hermitian", &M::complete_hermitian)
        .def("save_matrix_market", &M::save_matrix_market)
        .def("spmm_self",
             [](M& self, double thr, bool transA) { return self.spmm(self, thr, transA, false, 0); },
             py::arg("threshold"), py::arg("transA") = false)
        .def("add",
             [](M& self, M& B, T alpha, T beta) {
                 auto c = self.duplicate(true);
                 c->axpby(alpha, B, beta);
                 return c;
             },
             py::arg("B"), py::arg("alpha") = T(1), py::arg("beta") = T(1))
        .def("transpose", &M::transpose)
        .def("get_block", &M::get_block, py::arg("row"), py::arg("col"))
        .def("get_values", &M::get_values)
        .def("set_values",
             [](M& self, py::array_t<T, py::array::c_style | py::array::forcecast> v) {
                 if (v.size() != self.nnz())
                     throw std::invalid_argument("set_values: expected " + std::to_string(self.nnz()) + " values");
                 self.to_host();
                 std::memcpy(self.host.data(), v.data(), sizeof(T) * self.nnz());
                 self.host_written();
             })
        .def("filter_blocks", &M::filter_blocks)
        .def("to_dense", &M::to_dense)
        .def("from_dense", &M::from_dense, py::arg("array"))
        .def("get_block_density",
             [](M& self) {
                 const double n = double(self.graph->n_owned());
                 return n > 0 ? double(self.nnzb()) / (n * double(self.graph->block_sizes.size())) : 0.0;
             })
        .def_property_readonly("matrix_kind", &M::kind)
        .def_property_readonly("local_nnz", [](M& self) { return self.nnz(); })
        .def_property_readonly("local_block_nnz", [](M& self) { return self.nnzb(); })
        .def_property_readonly("global_nnz", [](M& self) { return self.nnz(); })
        .def_property_readonly("configured_page_size", [](M& self) { return self.page_size_; })
        .def_property("page_size", [](M& self) { return self.page_size_; },
                      [](M& self, uint32_t p) { self.page_size_ = p; })
        .def("set_page_size", [](M& self, uint32_t p) { self.page_size_ = p; }, py::arg("page_size"))
        .def_property_readonly("vendor_backend_name", [](M&) { return std::string("hip-gfx950"); })
        .def_property_readonly("vendor_launch_count", [](M&) { return uint64_t(0); })
        .def("reset_vendor_launch_count", [](M&) {})
        .def_property_readonly("shape_class_count", [](M& self) { return self.info().shape_classes; })
        .def_property_readonly("owned_scalar_rows", [](M& self) { return self.graph->owned_scalar_rows(); })
        .def_property_readonly("global_scalar_rows", [](M& self) { return self.graph->owned_scalar_rows(); })
        .def_property_readonly("global_scalar_cols", [](M& self) { return self.graph->owned_scalar_rows(); })
        .def_property_readonly("shape",
                               [](M& self) {
                                   const int n = self.graph->owned_scalar_rows();
                                   return py::make_tuple(n, n);
                               })
        .def_property_readonly("has_contiguous_layout", [](M&) { return true; })
        .def("pack_contiguous", [](M&) {})
        .def_property_readonly("row_ptr", [](M& self) { return py::array_t<int>(self.graph->adj_ptr.size(), self.graph->adj_ptr.data()); })
        .def_property_readonly("col_ind", [](M& self) { return py::array_t<int>(self.graph->adj_ind.size(), self.graph->adj_ind.data()); })
        .def_property_readonly("spmm_stats",
                               [](M& self) {
                                   return py::make_tuple(self.last_stats[0], self.last_stats[1], self.last_stats[2],
                                                         self.last_stats[3]);
                               })
        .def_property_readonly("values_ptr", [](M& self) { return reinterpret_cast<uintptr_t>(vbh_matrix_values(self.dev())); })
        .def("mark_values_written", [](M& self) { self.device_written(); })
        // the vbh_matrix_t of include/vbcsr_hip.h (values current on the device)
        .def_property_readonly("handle", [](M& self) { return reinterpret_cast<uintptr_t>(self.dev()); })
        .def("spmm_raw",
             [](M& self, int K, uintptr_t X, int ldx, uintptr_t Y, int ldy, int precision, bool adjoint) {
                 vbh_matrix_t m = self.dev();
                 py::gil_scoped_release nogil;
                 if (adjoint) check(vbh_spmm_adjoint(m, K, (const void*)X, ldx, (void*)Y, ldy, precision, nullptr), "spmm_raw");
                 else check(vbh_spmm(m, K, (const void*)X, ldx, (void*)Y, ldy, precision, nullptr), "spmm_raw");
             })
        .def("trace", [](M& self) {
            double t = 0.0;
            check(vbh_matrix_trace(self.dev(), &t, nullptr), "trace");
            return t;
        })
        .def("kpm_moments", &kpm_moments<T>, py::arg("V0"), py::arg("num_moments"), py::arg("mode") = 2,
             py::arg("center") = 0.0, py::arg("half_width") = 1.0, py::arg("precision") = 0)
        .def("cheb_apply", &cheb_apply<T>, py::arg("X"), py::arg("coeffs"), py::arg("center") = 0.0,
             py::arg("half_width") = 1.0, py::arg("precision") = 0);
}

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "vbcsr drop-in core on MI355X (gfx950): device-resident BlockSpMat / DistVector / DistMultiVector";
    m.def("threading_diagnostics", [](bool) {
        py::dict d;
        d["openmp_enabled"] = false;
        d["omp_get_max_threads"] = 1;
        d["mkl_enabled"] = false;
        d["mkl_get_max_threads_before"] = py::none();
        d["mkl_get_max_threads_after"] = py::none();
        int n = 0;
        d["hip_devices"] = vbh_device_count(&n) == VBH_OK ? n : 0;
        return d;
    }, py::arg("configure_vendor_sparse") = false);
    m.def("finalize_mpi", []() {});
    m.def("device_count", []() {
        int n = 0;
        check(vbh_device_count(&n), "device_count");
        return n;
    });
    m.def("library_version", &vbh_version);
    m.def("kpm_last_step_time", []() {
        double ms = 0.0;
        int64_t n = 0;
        check(vbh_kpm_last_step_time(&ms, &n), "kpm_last_step_time");
        return py::make_tuple(ms, n);
    });
    m.def("set_autotune", [](bool on) { check(vbh_set_autotune(on ? 1 : 0), "set_autotune"); });
    m.def("set_batch_plan", [](int t, int p) { check(vbh_set_batch_plan(t, p), "set_batch_plan"); });
    m.def("kpm_last_plan", []() {
        int t = 0, p = 0, tuned = 0;
        check(vbh_kpm_last_plan(&t, &p, &tuned), "kpm_last_plan");
        return py::make_tuple(t, p, bool(tuned));
    });
    m.def("synchronize", []() { check(vbh_stream_sync(nullptr), "synchronize"); });

    py::enum_<AssemblyMode>(m, "AssemblyMode")
        .value("INSERT", AssemblyMode::INSERT)
        .value("ADD", AssemblyMode::ADD)
        .export_values();
    py::enum_<RedistOp>(m, "RedistOp").value("Copy", RedistOp::Copy).value("Sum", RedistOp::Sum).export_values();
    py::class_<MultiVec<double>, std::shared_ptr<MultiVec<double>>>(m, "_DeviceVectorDouble");
    py::class_<MultiVec<cplx>, std::shared_ptr<MultiVec<cplx>>>(m, "_DeviceVectorComplex");

    py::class_<Graph, GraphP>(m, "DistGraph")
        .def(py::init([](py::object) { return std::make_shared<Graph>(); }), py::arg("comm") = py::none())
        .def("construct_serial", &Graph::construct_serial)
        .def("construct_distributed",
             [](Graph& g, const std::vector<int>& owned, const std::vector<int>& sizes,
                const std::vector<std::vector<int>>& adj) { g.build(owned, sizes, adj, {}, {}); })
        .def("construct_distributed_flat",
             [](Graph& g, py::array_t<int, py::array::c_style | py::array::forcecast> owned,
                py::array_t<int, py::array::c_style | py::array::forcecast> sizes,
                py::array_t<int64_t, py::array::c_style | py::array::forcecast> adj_ptr,
                py::array_t<int, py::array::c_style | py::array::forcecast> adj_ind) {
                 const int64_t n = owned.shape(0);
                 if (sizes.shape(0) != n || adj_ptr.shape(0) != n + 1)
                     throw std::runtime_error(
                         "construct_distributed_flat: block_sizes must match owned_indices and adj_ptr must have "
                         "one more entry than owned_indices");
                 std::vector<std::vector<int>> adj(n);
                 for (int64_t i = 0; i < n; ++i) adj[i].assign(adj_ind.data() + adj_ptr.data()[i], adj_ind.data() + adj_ptr.data()[i + 1]);
                 g.build(std::vector<int>(owned.data(), owned.data() + n), std::vector<int>(sizes.data(), sizes.data() + n), adj, {}, {});
             },
             py::arg("owned_indices"), py::arg("block_sizes"), py::arg("adj_ptr"), py::arg("adj_ind"))
        .def("construct_local", &Graph::build, py::arg("owned_indices"), py::arg("block_sizes"), py::arg("adjacency"),
             py::arg("ghost_indices"), py::arg("ghost_sizes"),
             "One rank's local share: owned blocks plus explicitly sized ghost columns.")
        .def_readonly("owned_global_indices", &Graph::owned_global_indices)
        .def_readonly("ghost_global_indices", &Graph::ghost_global_indices)
        .def_readonly("block_sizes", &Graph::block_sizes)
        .def_readonly("adj_ptr", &Graph::adj_ptr)
        .def_property_readonly("adj_ind", [](Graph& g) { return py::array_t<int>(g.adj_ind.size(), g.adj_ind.data()); })
        .def_readonly("block_offsets", &Graph::block_offsets)
        .def_property_readonly("send_counts", [](Graph&) { return std::vector<int>{0}; })
        .def_property_readonly("recv_counts", [](Graph&) { return std::vector<int>{0}; })
        .def_property_readonly("send_counts_scalar", [](Graph&) { return std::vector<int>{0}; })
        .def_property_readonly("recv_counts_scalar", [](Graph&) { return std::vector<int>{0}; })
        .def_property_readonly("send_ranks", [](Graph&) { return std::vector<int>{}; })
        .def_property_readonly("recv_ranks", [](Graph&) { return std::vector<int>{}; })
        .def_property_readonly("owned_scalar_rows", &Graph::owned_scalar_rows)
        .def_property_readonly("local_scalar_cols", &Graph::local_scalar_cols)
        .def_property_readonly("global_scalar_rows", &Graph::owned_scalar_rows)
        .def("get_local_index", &Graph::local_index)
        .def("get_global_index", &Graph::global_index)
        .def_readonly("rank", &Graph::rank)
        .def_readonly("size", &Graph::size);

    bind_vec<double, DistVec<double>>(m, "DistVector_Double", true);
    bind_vec<cplx, DistVec<cplx>>(m, "DistVector_Complex", true);
    bind_vec<double, DistMVec<double>>(m, "DistMultiVector_Double", false);
    bind_vec<cplx, DistMVec<cplx>>(m, "DistMultiVector_Complex", false);
    bind_mat<double>(m, "BlockSpMat_Double");
    bind_mat<cplx>(m, "BlockSpMat_Complex");
    m.def("kpm_moments_batched", &kpm_moments_batched<double>, py::arg("mats"), py::arg("V0s"),
          py::arg("num_moments"), py::arg("mode") = 2, py::arg("center") = 0.0, py::arg("half_width") = 1.0,
          py::arg("precision") = 0, py::arg("graph") = true);
    m.def("kpm_moments_batched", &kpm_moments_batched<cplx>, py::arg("mats"), py::arg("V0s"),
          py::arg("num_moments"), py::arg("mode") = 2, py::arg("center") = 0.0, py::arg("half_width") = 1.0,
          py::arg("precision") = 0, py::arg("graph") = true);
}
