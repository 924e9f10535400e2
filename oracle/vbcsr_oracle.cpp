// vbcsr CPU oracle -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// A plain scalar C++ restatement of the reference vbcsr semantics for the
// variable-block sparse hot path.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load it (through oracle/__init__.py), and only as
// the checker or the CPU baseline; the HIP product never links or calls it.
//
// Layout shared by every entry point (one rank's local share):
//   nbr block rows with sizes rbs[nbr]; nbc block columns with sizes cbs[nbc]
//   (owned columns first, then ghost columns); CSR row_ptr[nbr+1] /
//   col_ind[nnzb] in slot order; block values packed in slot order, each block
//   row-major (element (a,b) of an r x c block at blk[a*c+b]) -- the
//   reference's get_values() export (vbcsr/core/block_csr.hpp:1202-1248,
//   block layout block_csr.hpp:16-22).
//
// Pinning of every function (closed forms, golden values, the compiled
// reference in oracle/_ref) is listed in tests/test_oracle.py; no function
// here is "parity unpinned".
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

using cplx = std::complex<double>;

inline double sqnorm(double v) { return v * v; }
inline double sqnorm(const cplx& v) { return std::norm(v); }
inline double conj_(double v) { return v; }
inline cplx conj_(const cplx& v) { return std::conj(v); }

struct Pattern {
    int nbr, nbc;
    const int* rbs;
    const int* cbs;
    const int* row_ptr;
    const int* col_ind;
    std::vector<int64_t> roff, coff, voff;  // scalar row/col offsets, value offsets
    Pattern(int nbr_, int nbc_, const int* rbs_, const int* cbs_, const int* rp, const int* ci)
        : nbr(nbr_), nbc(nbc_), rbs(rbs_), cbs(cbs_), row_ptr(rp), col_ind(ci) {
        roff.assign(nbr + 1, 0);
        coff.assign(nbc + 1, 0);
        for (int i = 0; i < nbr; ++i) roff[i + 1] = roff[i] + rbs[i];
        for (int j = 0; j < nbc; ++j) coff[j + 1] = coff[j] + cbs[j];
        const int nnzb = row_ptr[nbr];
        voff.assign(nnzb + 1, 0);
        for (int i = 0; i < nbr; ++i)
            for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
                voff[k + 1] = voff[k] + int64_t(rbs[i]) * cbs[col_ind[k]];
    }
};

// y = A x: y zeroed, then per row, per slot (ascending), y_row += A_blk x_col.
// vbcsr/core/detail/kernels/vbcsr_apply.hpp:81-103,193-244 (row direct),
// per-block arithmetic rowmajor_kernels.hpp:80-90 (generic rm_gemv).
template <typename T>
void spmv(const Pattern& p, const T* vals, const T* x, T* y) {
    for (int64_t s = 0; s < p.roff[p.nbr]; ++s) y[s] = T(0);
    for (int i = 0; i < p.nbr; ++i) {
        const int r = p.rbs[i];
        T* yr = y + p.roff[i];
        for (int k = p.row_ptr[i]; k < p.row_ptr[i + 1]; ++k) {
            const int j = p.col_ind[k];
            const int c = p.cbs[j];
            const T* blk = vals + p.voff[k];
            const T* xc = x + p.coff[j];
            for (int a = 0; a < r; ++a) {
                T acc = T(0);
                for (int b = 0; b < c; ++b) acc += blk[a * c + b] * xc[b];
                yr[a] += acc;
            }
        }
    }
}

// Y = A X over K columns (row-major, leading dims ldx/ldy).
// vbcsr_apply.hpp:277-340 (row direct, each row's Y segment zeroed first),
// per-block arithmetic rowmajor_kernels.hpp:44-78 (generic rm_gemm).
template <typename T>
void spmm(const Pattern& p, const T* vals, int K, const T* X, int ldx, T* Y, int ldy) {
    for (int i = 0; i < p.nbr; ++i) {
        const int r = p.rbs[i];
        T* yr = Y + p.roff[i] * ldy;
        for (int a = 0; a < r; ++a)
            for (int v = 0; v < K; ++v) yr[int64_t(a) * ldy + v] = T(0);
        for (int k = p.row_ptr[i]; k < p.row_ptr[i + 1]; ++k) {
            const int j = p.col_ind[k];
            const int c = p.cbs[j];
            const T* blk = vals + p.voff[k];
            const T* xc = X + p.coff[j] * ldx;
            for (int a = 0; a < r; ++a)
                for (int b = 0; b < c; ++b) {
                    const T av = blk[a * c + b];
                    for (int v = 0; v < K; ++v)
                        yr[int64_t(a) * ldy + v] += av * xc[int64_t(b) * ldx + v];
                }
        }
    }
}

// Y = A^H X (columns of the local column space): Y zeroed over all columns,
// then every block adds A_blk^H x_row into its column segment.
// vbcsr_apply.hpp:57-74 (zero whole buffer, column direct), per-block
// rowmajor_kernels.hpp:92-101 (rm_gemv_adjoint: y[b] += conj(a[a*c+b]) x[a]).
template <typename T>
void spmm_adjoint(const Pattern& p, const T* vals, int K, const T* X, int ldx, T* Y, int ldy) {
    for (int64_t s = 0; s < p.coff[p.nbc]; ++s)
        for (int v = 0; v < K; ++v) Y[s * ldy + v] = T(0);
    for (int i = 0; i < p.nbr; ++i) {
        const int r = p.rbs[i];
        const T* xr = X + p.roff[i] * ldx;
        for (int k = p.row_ptr[i]; k < p.row_ptr[i + 1]; ++k) {
            const int j = p.col_ind[k];
            const int c = p.cbs[j];
            const T* blk = vals + p.voff[k];
            T* yc = Y + p.coff[j] * ldy;
            for (int a = 0; a < r; ++a)
                for (int b = 0; b < c; ++b) {
                    const T av = conj_(blk[a * c + b]);
                    for (int v = 0; v < K; ++v)
                        yc[int64_t(b) * ldy + v] += av * xr[int64_t(a) * ldx + v];
                }
        }
    }
}

// Frobenius norm per block, accumulated in double.  block_csr.hpp:2486-2502.
template <typename T>
void block_norms(const Pattern& p, const T* vals, double* out) {
    const int nnzb = p.row_ptr[p.nbr];
    for (int k = 0; k < nnzb; ++k) {
        double s = 0.0;
        for (int64_t e = p.voff[k]; e < p.voff[k + 1]; ++e) s += sqnorm(vals[e]);
        out[k] = std::sqrt(s);
    }
}

template <typename T>
struct Result {
    std::vector<int> row_ptr, col_ind;
    std::vector<T> vals;
};

// Thresholded block SpGEMM C = A B (serial, one rank).
//  symbolic: est_ij = sum_k |A_ik|_F |B_kj|_F accumulated in (a_slot, b_slot)
//            order, keep j iff est_ij > thr (strict), row sorted by column
//            -- vbcsr/core/detail/ops/spmm/common.hpp:507-690 (keep rule :664,
//            sort :668).
//  numeric:  C blocks start at zero; for each A slot (i,k) in order,
//            row_eps = thr / max(1, nnzb_row_A(i)); for each B slot (k,j) in
//            order skip when |A_ik||B_kj| < row_eps, otherwise, if j is in C's
//            row, C_ij += A_ik B_kj -- ops/spmm/vbcsr.hpp:342-420 (row_eps
//            :371-372, skip :385-387, hashed destination :505-527).
//  filter:   only when thr > 0, drop blocks with |C_ij|_F < thr, keep >= thr
//            -- block_csr.hpp:1399-1528 (rule :1411), called at vbcsr.hpp:95.
// upper: keep only block columns j >= i (the upper block triangle of a
// square result, spmm_hermitian_upper -- common.hpp:502-513, skip :625).
// A: nbr x nbk block rows/cols, B: nbk x nbc.
template <typename T>
Result<T>* spgemm(const Pattern& A, const T* av, const Pattern& B, const T* bv, double thr, bool upper = false) {
    std::vector<double> an(A.row_ptr[A.nbr]), bn(B.row_ptr[B.nbr]);
    block_norms(A, av, an.data());
    block_norms(B, bv, bn.data());
    auto* res = new Result<T>();
    res->row_ptr.assign(A.nbr + 1, 0);
    std::vector<std::vector<int>> cols(A.nbr);
    for (int i = 0; i < A.nbr; ++i) {
        std::map<int, double> est;  // key order irrelevant; per-key sum order is slot order
        for (int ka = A.row_ptr[i]; ka < A.row_ptr[i + 1]; ++ka) {
            const int k = A.col_ind[ka];
            for (int kb = B.row_ptr[k]; kb < B.row_ptr[k + 1]; ++kb) {
                const int j = B.col_ind[kb];
                if (upper && j < i) continue;
                auto it = est.find(j);
                if (it == est.end()) est.emplace(j, an[ka] * bn[kb]);
                else it->second += an[ka] * bn[kb];
            }
        }
        for (auto& kv : est)
            if (kv.second > thr) cols[i].push_back(kv.first);
        res->row_ptr[i + 1] = res->row_ptr[i] + int(cols[i].size());
    }
    for (int i = 0; i < A.nbr; ++i) res->col_ind.insert(res->col_ind.end(), cols[i].begin(), cols[i].end());
    std::vector<int> crbs(A.rbs, A.rbs + A.nbr);
    Pattern C(A.nbr, B.nbc, A.rbs, B.cbs, res->row_ptr.data(), res->col_ind.data());
    res->vals.assign(C.voff.back(), T(0));
    for (int i = 0; i < A.nbr; ++i) {
        const int r = A.rbs[i];
        const int row_count = A.row_ptr[i + 1] - A.row_ptr[i];
        const double row_eps = thr / std::max(1, row_count);
        for (int ka = A.row_ptr[i]; ka < A.row_ptr[i + 1]; ++ka) {
            const int k = A.col_ind[ka];
            const int m = A.cbs[k];
            const T* ablk = av + A.voff[ka];
            for (int kb = B.row_ptr[k]; kb < B.row_ptr[k + 1]; ++kb) {
                if (an[ka] * bn[kb] < row_eps) continue;
                const int j = B.col_ind[kb];
                const int* first = res->col_ind.data() + res->row_ptr[i];
                const int* last = res->col_ind.data() + res->row_ptr[i + 1];
                const int* pos = std::lower_bound(first, last, j);
                if (pos == last || *pos != j) continue;
                const int slot = int(pos - res->col_ind.data());
                const int c = B.cbs[j];
                const T* bblk = bv + B.voff[kb];
                T* cblk = res->vals.data() + C.voff[slot];
                for (int a = 0; a < r; ++a)
                    for (int q = 0; q < m; ++q) {
                        const T aq = ablk[a * m + q];
                        for (int b = 0; b < c; ++b) cblk[a * c + b] += aq * bblk[q * c + b];
                    }
            }
        }
    }
    if (thr > 0.0) {
        std::vector<double> cn(C.voff.size() - 1);
        block_norms(C, res->vals.data(), cn.data());
        auto* f = new Result<T>();
        f->row_ptr.assign(A.nbr + 1, 0);
        for (int i = 0; i < A.nbr; ++i) {
            for (int s = res->row_ptr[i]; s < res->row_ptr[i + 1]; ++s) {
                if (cn[s] < thr) continue;
                f->col_ind.push_back(res->col_ind[s]);
                f->vals.insert(f->vals.end(), res->vals.begin() + C.voff[s], res->vals.begin() + C.voff[s + 1]);
            }
            f->row_ptr[i + 1] = int(f->col_ind.size());
        }
        delete res;
        res = f;
    }
    return res;
}

// filter_blocks(thr): block_csr.hpp:1399-1528 -- no-op for thr <= 0,
// otherwise keep blocks with |blk|_F >= thr in slot order.
template <typename T>
Result<T>* filter(const Pattern& p, const T* vals, double thr) {
    std::vector<double> n(p.row_ptr[p.nbr]);
    block_norms(p, vals, n.data());
    auto* f = new Result<T>();
    f->row_ptr.assign(p.nbr + 1, 0);
    for (int i = 0; i < p.nbr; ++i) {
        for (int s = p.row_ptr[i]; s < p.row_ptr[i + 1]; ++s) {
            if (thr > 0.0 && n[s] < thr) continue;
            f->col_ind.push_back(p.col_ind[s]);
            f->vals.insert(f->vals.end(), vals + p.voff[s], vals + p.voff[s + 1]);
        }
        f->row_ptr[i + 1] = int(f->col_ind.size());
    }
    return f;
}

// self <- alpha X + beta self over the union pattern (sorted columns per
// row; blocks present in only one operand take only that operand's term;
// union blocks present in neither cannot occur).  ops/axpby.hpp:417-527.
template <typename T>
Result<T>* axpby_union(const Pattern& Y, const T* yv, const Pattern& X, const T* xv, T alpha, T beta) {
    auto* f = new Result<T>();
    f->row_ptr.assign(Y.nbr + 1, 0);
    for (int i = 0; i < Y.nbr; ++i) {
        int ky = Y.row_ptr[i], kx = X.row_ptr[i];
        const int ey = Y.row_ptr[i + 1], ex = X.row_ptr[i + 1];
        while (ky < ey || kx < ex) {
            const int cy = ky < ey ? Y.col_ind[ky] : INT32_MAX;
            const int cx = kx < ex ? X.col_ind[kx] : INT32_MAX;
            const int col = std::min(cy, cx);
            const int64_t sz = int64_t(Y.rbs[i]) * Y.cbs[col];
            f->col_ind.push_back(col);
            const size_t base = f->vals.size();
            f->vals.resize(base + sz, T(0));
            if (cy == col) {
                for (int64_t e = 0; e < sz; ++e) f->vals[base + e] = beta * yv[Y.voff[ky] + e];
                ++ky;
            }
            if (cx == col) {
                for (int64_t e = 0; e < sz; ++e) f->vals[base + e] += alpha * xv[X.voff[kx] + e];
                ++kx;
            }
        }
        f->row_ptr[i + 1] = int(f->col_ind.size());
    }
    return f;
}

// shift(alpha): alpha added to the leading min(r,c) diagonal of the block
// whose column is the row's own block, when stored.  block_csr.hpp:1289-1325.
// scale(alpha): every stored value times alpha.  block_csr.hpp:1053-1068.
template <typename T>
void shift(const Pattern& p, T* vals, T alpha, const int* col_of_row) {
    for (int i = 0; i < p.nbr; ++i) {
        for (int k = p.row_ptr[i]; k < p.row_ptr[i + 1]; ++k) {
            const int j = p.col_ind[k];
            if (j != col_of_row[i]) continue;
            const int r = p.rbs[i], c = p.cbs[j];
            for (int d = 0; d < std::min(r, c); ++d) vals[p.voff[k] + int64_t(d) * c + d] += alpha;
            break;
        }
    }
}

// splitmix64 / entry_hash: vbcsr/core/random_vectors.hpp:64-79 (Steele, Lea,
// Flood 2014 finaliser; the four multipliers are the published constants).
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
inline uint64_t entry_hash(uint64_t seed, uint64_t block, uint64_t row, uint64_t col) {
    uint64_t h = splitmix64(seed + 0x9e3779b97f4a7c15ULL);
    h = splitmix64(h ^ (block * 0xbf58476d1ce4e5b9ULL));
    h = splitmix64(h ^ (row * 0x94d049bb133111ebULL));
    return splitmix64(h ^ (col * 0xd6e8feb86659fd93ULL));
}

// KPM moments, summed over the K columns (not yet divided by K*N):
//  mode 1 (one moment per matvec): v1 = H~ v0, v_{n+1} = 2 H~ v_n - v_{n-1},
//         mu_n = sum_v Re <v0, v_n> -- examples/ex08_graphene_dos.py:95-114.
//  mode 2 (two moments per matvec): mu_0 = <t0,t0>, mu_1 = <t0,t1>,
//         mu_{2n} = 2<t_n,t_n> - mu_0, mu_{2n-1} = 2<t_n,t_{n-1}> - mu_1
//         -- examples/inp_nanocrystal_dos/kpm.py:46-73.
//  H~ = (H - center)/half_width applied as in kpm.py:36-39.
template <typename T>
void kpm_moments(const Pattern& p, const T* vals, int K, const T* V0, int M, int mode,
                 double center, double half_width, double* mu) {
    const int64_t N = p.roff[p.nbr];
    const int64_t Nc = p.coff[p.nbc];
    std::vector<T> a(Nc * K, T(0)), b(Nc * K, T(0)), c(Nc * K, T(0));
    std::copy(V0, V0 + N * K, a.begin());
    const std::vector<T> v0(a);
    auto apply = [&](const std::vector<T>& src, std::vector<T>& dst) {
        spmm(p, vals, K, src.data(), K, dst.data(), K);
        for (int64_t e = 0; e < N * K; ++e) dst[e] = (dst[e] - center * src[e]) / half_width;
    };
    auto dot = [&](const std::vector<T>& u, const std::vector<T>& w) {
        double s = 0.0;
        for (int v = 0; v < K; ++v) {
            T acc = T(0);
            for (int64_t r = 0; r < N; ++r) acc += conj_(u[r * K + v]) * w[r * K + v];
            s += std::real(acc);
        }
        return s;
    };
    for (int n = 0; n < M; ++n) mu[n] = 0.0;
    mu[0] = dot(a, a);
    if (M == 1) return;
    apply(a, b);
    mu[1] = dot(a, b);
    if (mode == 1) {
        // a = v_{n-1}, b = v_n
        for (int n = 2; n < M; ++n) {
            apply(b, c);
            for (int64_t e = 0; e < N * K; ++e) c[e] = T(2.0) * c[e] - a[e];
            mu[n] = dot(v0, c);
            std::swap(a, b);
            std::swap(b, c);
        }
        return;
    }
    // mode 2: a = t_{n-1}, b = t_n
    for (int n = 1;; ++n) {
        const int odd = 2 * n - 1, even = 2 * n;
        if (odd > 1 && odd < M) mu[odd] = 2.0 * dot(b, a) - mu[1];
        if (even < M) mu[even] = 2.0 * dot(b, b) - mu[0];
        if (2 * n + 1 >= M) break;
        apply(b, c);
        for (int64_t e = 0; e < N * K; ++e) c[e] = T(2.0) * c[e] - a[e];
        std::swap(a, b);
        std::swap(b, c);
    }
}

// Chebyshev series apply y = sum_n coef_n T_n(H~) x with the three-term
// recursion of examples/ex09_graphene_dynamics.py:115-131 and
// vbcsr/utils.py:71-126 (kpm_apply; H~ = (H - center)/half_width).
template <typename T>
void cheb_apply(const Pattern& p, const T* vals, int K, const T* X, int nterms, const cplx* coef,
                double center, double half_width, T* Y) {
    const int64_t N = p.roff[p.nbr];
    const int64_t Nc = p.coff[p.nbc];
    std::vector<T> a(Nc * K, T(0)), b(Nc * K, T(0)), c(Nc * K, T(0));
    std::copy(X, X + N * K, a.begin());
    auto cf = [&](int n) -> T {
        if constexpr (std::is_same_v<T, cplx>) return coef[n];
        else return T(coef[n].real());
    };
    for (int64_t e = 0; e < N * K; ++e) Y[e] = cf(0) * a[e];
    if (nterms == 1) return;
    spmm(p, vals, K, a.data(), K, b.data(), K);
    for (int64_t e = 0; e < N * K; ++e) b[e] = (b[e] - center * a[e]) / half_width;
    for (int64_t e = 0; e < N * K; ++e) Y[e] += cf(1) * b[e];
    for (int n = 2; n < nterms; ++n) {
        spmm(p, vals, K, b.data(), K, c.data(), K);
        for (int64_t e = 0; e < N * K; ++e) c[e] = T(2.0) * (c[e] - center * b[e]) / half_width - a[e];
        for (int64_t e = 0; e < N * K; ++e) Y[e] += cf(n) * c[e];
        std::swap(a, b);
        std::swap(b, c);
    }
}

}  // namespace

#define PAT Pattern P(nbr, nbc, rbs, cbs, rp, ci)
#define PATARGS int nbr, int nbc, const int *rbs, const int *cbs, const int *rp, const int *ci

ORC_API uint64_t orc_entry_hash(uint64_t seed, uint64_t block, uint64_t row, uint64_t col) {
    return entry_hash(seed, block, row, col);
}

// Rademacher columns: vbcsr/core/random_vectors.hpp:99-143 (+1 iff bit 63 of
// the hash is set), keyed (seed, global block, row in block, column);
// normalised by 1/sqrt(N_global) -- dist_multivector.hpp:335-372.
ORC_API void orc_random_rademacher(int nblocks, const int* bs, const int* gids, int K, uint64_t seed,
                                   int normalize, int64_t n_global, double* out, int ld) {
    const double f = normalize ? 1.0 / std::sqrt(double(n_global)) : 1.0;
    int64_t row = 0;
    for (int b = 0; b < nblocks; ++b)
        for (int a = 0; a < bs[b]; ++a, ++row)
            for (int v = 0; v < K; ++v)
                out[row * ld + v] = ((entry_hash(seed, uint64_t(gids[b]), uint64_t(a), uint64_t(v)) >> 63) ? 1.0 : -1.0) * f;
}

#define ORC_TYPED(SFX, T)                                                                          \
    ORC_API void orc_spmv_##SFX(PATARGS, const void* v, const void* x, void* y) {                  \
        PAT;                                                                                       \
        spmv(P, (const T*)v, (const T*)x, (T*)y);                                                  \
    }                                                                                              \
    ORC_API void orc_spmm_##SFX(PATARGS, const void* v, int K, const void* X, int ldx, void* Y,   \
                                int ldy) {                                                         \
        PAT;                                                                                       \
        spmm(P, (const T*)v, K, (const T*)X, ldx, (T*)Y, ldy);                                     \
    }                                                                                              \
    ORC_API void orc_spmm_adjoint_##SFX(PATARGS, const void* v, int K, const void* X, int ldx,     \
                                        void* Y, int ldy) {                                        \
        PAT;                                                                                       \
        spmm_adjoint(P, (const T*)v, K, (const T*)X, ldx, (T*)Y, ldy);                             \
    }                                                                                              \
    ORC_API void orc_block_norms_##SFX(PATARGS, const void* v, double* out) {                      \
        PAT;                                                                                       \
        block_norms(P, (const T*)v, out);                                                          \
    }                                                                                              \
    ORC_API void* orc_spgemm_##SFX(int nbr, int nbk, int nbc, const int* rbs, const int* kbs,      \
                                   const int* cbs, const int* arp, const int* aci, const void* av, \
                                   const int* brp, const int* bci, const void* bv, double thr,     \
                                   int upper) {                                                    \
        Pattern A(nbr, nbk, rbs, kbs, arp, aci);                                                   \
        Pattern B(nbk, nbc, kbs, cbs, brp, bci);                                                   \
        return spgemm(A, (const T*)av, B, (const T*)bv, thr, upper != 0);                          \
    }                                                                                              \
    ORC_API void* orc_filter_##SFX(PATARGS, const void* v, double thr) {                           \
        PAT;                                                                                       \
        return filter(P, (const T*)v, thr);                                                        \
    }                                                                                              \
    ORC_API void* orc_axpby_##SFX(PATARGS, const void* yv, const int* xrp, const int* xci,        \
                                  const void* xv, const void* alpha, const void* beta) {           \
        PAT;                                                                                       \
        Pattern X(nbr, nbc, rbs, cbs, xrp, xci);                                                   \
        return axpby_union(P, (const T*)yv, X, (const T*)xv, *(const T*)alpha, *(const T*)beta);  \
    }                                                                                              \
    ORC_API void orc_shift_##SFX(PATARGS, void* v, const void* alpha, const int* col_of_row) {     \
        PAT;                                                                                       \
        shift(P, (T*)v, *(const T*)alpha, col_of_row);                                             \
    }                                                                                              \
    ORC_API long orc_result_nnzb_##SFX(void* r) { return long(((Result<T>*)r)->col_ind.size()); } \
    ORC_API long orc_result_nvals_##SFX(void* r) { return long(((Result<T>*)r)->vals.size()); }   \
    ORC_API void orc_result_copy_##SFX(void* r, int* rp, int* ci, void* v) {                       \
        auto* R = (Result<T>*)r;                                                                   \
        std::copy(R->row_ptr.begin(), R->row_ptr.end(), rp);                                       \
        std::copy(R->col_ind.begin(), R->col_ind.end(), ci);                                       \
        std::copy(R->vals.begin(), R->vals.end(), (T*)v);                                          \
    }                                                                                              \
    ORC_API void orc_result_free_##SFX(void* r) { delete (Result<T>*)r; }                          \
    ORC_API void orc_kpm_moments_##SFX(PATARGS, const void* v, int K, const void* V0, int M,       \
                                       int mode, double center, double half_width, double* mu) {   \
        PAT;                                                                                       \
        kpm_moments(P, (const T*)v, K, (const T*)V0, M, mode, center, half_width, mu);            \
    }                                                                                              \
    ORC_API void orc_cheb_apply_##SFX(PATARGS, const void* v, int K, const void* X, int nterms,    \
                                      const void* coef, double center, double half_width,          \
                                      void* Y) {                                                   \
        PAT;                                                                                       \
        cheb_apply(P, (const T*)v, K, (const T*)X, nterms, (const cplx*)coef, center, half_width, \
                   (T*)Y);                                                                         \
    }

ORC_TYPED(d, double)
ORC_TYPED(z, cplx)
