/* vbcsr_hip.h -- C ABI of the MI355X (gfx950) variable-block sparse hot path.
 *
 * Plain C: opaque handles, pointers and sizes, int status codes (0 = success),
 * a thread-local message behind vbh_last_error(), and an explicit HIP stream
 * (passed as void*, NULL = the default stream) on every call that launches
 * device work.  No C++ exception crosses this boundary.
 *
 * A matrix is one rank's local share of a vbcsr block matrix: nbr owned block
 * rows, nbc local block columns (owned first, then ghosts), block-row CSR in
 * slot order with columns ascending per row, and block values packed in slot
 * order with every block row-major -- the reference's get_values() export
 * (vbcsr/core/block_csr.hpp:1202-1248; element (a,b) of an r x c block at
 * blk[a*c+b], block_csr.hpp:16-22).  Vectors are device arrays over the local
 * scalar rows; multivectors are row-major with a leading dimension ld >= K
 * whose pad lanes hold zeros (vbcsr/core/dist_multivector.hpp:24-47).
 *
 * Each entry point names the reference interface it replaces.  The pybind
 * host module vbcsr_amd/csrc/pybind_core.cpp binds them behind the unchanged
 * VBCSR / DistVector / DistMultiVector Python API.
 */
#ifndef VBCSR_HIP_H
#define VBCSR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VBH_API __attribute__((visibility("default")))

/* Value types. */
enum {
    VBH_F64 = 0,  /* double */
    VBH_C128 = 1, /* complex<double>, interleaved (re, im) */
    VBH_F32 = 2,  /* float */
    VBH_C64 = 3   /* complex<float>, interleaved (re, im) */
};

/* Status codes. */
enum {
    VBH_OK = 0,
    VBH_ERR_ARG = 1,     /* invalid argument / shape mismatch */
    VBH_ERR_HIP = 2,     /* HIP runtime failure */
    VBH_ERR_NOMEM = 3,   /* device allocation failed */
    VBH_ERR_STATE = 4,   /* handle in the wrong state */
    VBH_ERR_NOTSUP = 5   /* unsupported combination */
};

typedef struct vbh_matrix* vbh_matrix_t;

typedef struct {
    int32_t nbr, nbc;       /* block rows, local block columns */
    int64_t nnzb;           /* stored blocks */
    int64_t nnz;            /* stored scalars (sum r*c) */
    int64_t n_rows, n_cols; /* scalar rows / columns of the local share */
    int32_t dtype;          /* VBH_F64 or VBH_C128 */
    int32_t kind;           /* 0 csr (all 1x1), 1 bsr (uniform), 2 vbcsr */
    int32_t max_block;      /* largest block dimension */
    int32_t shape_classes;  /* distinct (r, c) block shapes */
} vbh_matrix_info;

/* ---- runtime ---------------------------------------------------------- */
VBH_API const char* vbh_last_error(void);
VBH_API int vbh_version(void);
VBH_API int vbh_device_count(int* count);
VBH_API int vbh_set_device(int device);
/* Device memory of the library's allocator.  Buffers of 64 MiB and more are
 * kept for reuse when freed (per device, up to 16 GiB in total; released when
 * an allocation fails); vbh_free synchronises the device, as hipFree does. */
VBH_API int vbh_malloc(void** ptr, size_t bytes);
VBH_API int vbh_free(void* ptr);
VBH_API int vbh_memcpy(void* dst, const void* src, size_t bytes, int kind /*0 h2d 1 d2h 2 d2d*/,
                       void* stream);
VBH_API int vbh_memset(void* dst, int value, size_t bytes, void* stream);
VBH_API int vbh_stream_sync(void* stream);
/* Threading.  The library keeps per-(device, stream) state: scratch buffers
 * and cached launch objects (the batched-KPM graph).  Every call holds the
 * state of its stream for its whole duration, so calls on different streams
 * may run concurrently from different threads, and calls on one stream are
 * serialised.  Structures a matrix builds on first use (FP32 value copy,
 * row-class lists, adjoint plan) are built under the matrix's own lock and
 * are complete on the device before any kernel reads them.  Writing a
 * matrix's values while another call reads them is the caller's race, as in
 * any library.  vbh_stream_release frees the state of a stream the caller is
 * about to destroy (no reference counterpart: the reference holds the GIL
 * and runs no streams). */
VBH_API int vbh_stream_release(void* stream);

/* ---- matrix lifetime and values --------------------------------------- */
/* Build the device structure from host CSR arrays (DistGraph local view,
 * vbcsr/core/dist_graph.hpp:19-77; kind detection block_csr.hpp:2770-2798).
 * rbs[nbr] row block sizes, cbs[nbc] column block sizes, row_ptr[nbr+1],
 * col_ind[nnzb] local column ids ascending per row.  Values start at zero. */
VBH_API int vbh_matrix_create(int nbr, int nbc, const int* rbs, const int* cbs, const int* row_ptr,
                              const int* col_ind, int dtype, vbh_matrix_t* out);
VBH_API int vbh_matrix_destroy(vbh_matrix_t m);
VBH_API int vbh_matrix_info_get(vbh_matrix_t m, vbh_matrix_info* info);
/* Host copy of the pattern (row_ptr[nbr+1], col_ind[nnzb]).  block_csr.hpp:234-241. */
VBH_API int vbh_matrix_get_pattern(vbh_matrix_t m, int* row_ptr, int* col_ind);
/* Values in slot order, row-major blocks.  set: add_block(INSERT)+assemble
 * (block_csr.hpp:613,861); get: get_values (block_csr.hpp:1202). location: 0
 * host pointer, 1 device pointer. */
VBH_API int vbh_matrix_set_values(vbh_matrix_t m, const void* vals, int location, void* stream);
VBH_API int vbh_matrix_get_values(vbh_matrix_t m, void* vals, int location, void* stream);
/* Device pointer to the values (for in-place kernels such as the adjoint).
 * After writing through it, call vbh_matrix_values_changed. */
VBH_API void* vbh_matrix_values(vbh_matrix_t m);
/* The values were written through vbh_matrix_values (or by any device copy
 * the library did not make): copies derived from them (the FP32 copy of
 * precisions 1 and 2) are rebuilt on next use.  The counterpart of the
 * reference's in-place value writes (add_block INSERT/ADD after assembly,
 * block_csr.hpp:613-700), which have no derived copies to refresh. */
VBH_API int vbh_matrix_values_changed(vbh_matrix_t m);
/* Deep copy (block_csr.hpp:579 duplicate). */
VBH_API int vbh_matrix_duplicate(vbh_matrix_t m, vbh_matrix_t* out, void* stream);
/* Transpose (conj != 0: conjugate transpose) as a new matrix: nbc x nbr
 * blocks, block (j, i) = A_ij^T (^H), rows sorted (BlockSpMat::transpose,
 * block_csr.hpp:1530-1543 with detail/ops/transpose.hpp:178-).  The pattern
 * comes from the handle's host mirror; the values are transposed on the
 * device. */
VBH_API int vbh_matrix_transpose(vbh_matrix_t m, int conj, vbh_matrix_t* out, void* stream);

/* ---- applies ------------------------------------------------------------
 * precision: 0 = the matrix's own type (FP64 / complex128 vectors),
 *            1 = FP32 (float / complex<float> vectors; the matrix values
 *                are read from an FP32 copy kept on the device),
 *            2 = FP32 values against the matrix's own vector type (FP64 /
 *                complex128 vectors, FP32 value copy): the value stream
 *                halves, the vectors keep full precision.  1 <= K <= 64;
 *                not for the adjoint or per-column fused dots.
 * Every precision accumulates in FP64 registers.
 * x has n_cols entries (owned + ghost columns), y has n_rows; y is fully
 * overwritten.  Multivectors: K columns, leading dims ldx, ldy >= K; only
 * columns < K of each row are read or written. */
/* BlockSpMat::mult, block_csr.hpp:985 (vbcsr_apply.hpp:25-42, bsr_apply.hpp:331-408). */
VBH_API int vbh_spmv(vbh_matrix_t m, const void* x, void* y, int precision, void* stream);
/* BlockSpMat::mult_dense, block_csr.hpp:1011 (vbcsr_apply.hpp:277-340, bsr_apply.hpp:410-485). */
VBH_API int vbh_spmm(vbh_matrix_t m, int K, const void* X, int ldx, void* Y, int ldy, int precision,
                     void* stream);
/* BlockSpMat::mult_adjoint / mult_dense_adjoint, block_csr.hpp:1024,1040
 * (vbcsr_apply.hpp:57-74,342-424).  Y has n_cols rows and is fully overwritten. */
VBH_API int vbh_spmv_adjoint(vbh_matrix_t m, const void* x, void* y, int precision, void* stream);
VBH_API int vbh_spmm_adjoint(vbh_matrix_t m, int K, const void* X, int ldx, void* Y, int ldy,
                             int precision, void* stream);
/* Fused apply: Y = alpha (A X) + beta X[owned rows] + gamma Z, with optional
 * per-column FP64 dot products d1[v] = Re sum conj(W)Y, d2[v] = Re sum
 * conj(X_own)Y added into dots (device pointer, 2K doubles: d1 then d2; NULL
 * to skip).  The composition examples/ex08_graphene_dos.py:109-113 and
 * vbcsr/utils.py:107-117 build from mult_dense + axpby + bdot in one pass.
 * alpha/beta/gamma are real.  W NULL means W = Y. */
VBH_API int vbh_spmm_fused(vbh_matrix_t m, int K, const void* X, int ldx, void* Y, int ldy,
                           double alpha, double beta, double gamma, const void* Z, int ldz,
                           const void* W, int ldw, double* dots, int precision, void* stream);

/* ---- KPM / Chebyshev ------------------------------------------------------
 * Moments mu[n] = sum_v Re <v0_v, T_n(H~) v0_v>, n = 0..M-1, H~ =
 * (H - center)/half_width; V0 is a device multivector (n_rows x K, ld ldv) of
 * the matrix's type (square block structure, no ghosts).  mode 1: one moment
 * per matvec (examples/ex08_graphene_dos.py:95-114); mode 2: two moments per
 * matvec (examples/inp_nanocrystal_dos/kpm.py:46-73).  precision 1 runs the
 * recursion in FP32 (FP32 vectors and values), precision 2 with FP32 values
 * and FP64 vectors; the moment reductions are FP64 in every precision.
 * moments: host, M doubles. */
VBH_API int vbh_kpm_moments(vbh_matrix_t m, int K, const void* V0, int ldv, int M, int mode,
                            double center, double half_width, int precision, double* moments,
                            void* stream);
/* vbh_kpm_moments without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as vbh_kpm_moments (the
 * same kernels run); the work buffers (recursion vectors, partial dots,
 * moments) live in a static device array of 128 MiB per GPU instead of
 * library-allocated scratch.  A call needing more returns VBH_ERR_ARG with
 * the size (about (3 or 4) * n_rows * K * sizeof(scalar) + 8 * steps *
 * work-groups * 2 bytes).  The matrix's lazily built structures must exist:
 * run one vbh_kpm_moments (or apply) at the same precision on it first, else
 * VBH_ERR_STATE.  Calls in one process are serialised by a host lock (each
 * synchronises its stream before returning), so calls never overlap on the
 * array; the array keeps the last call's contents. */
VBH_API int vbh_kpm_moments_noalloc(vbh_matrix_t m, int K, const void* V0, int ldv, int M, int mode,
                                    double center, double half_width, int precision, double* moments,
                                    void* stream);
/* Batched KPM over many independent matrices in one launch per recursion
 * step, the steps captured in one HIP graph (graph != 0).  V0s[h] device
 * multivectors (n_rows_h x K, ld ldv[h]).  moments: host, count*M doubles. */
VBH_API int vbh_kpm_moments_batched(int count, const vbh_matrix_t* ms, int K, const void* const* V0s,
                                    const int* ldv, int M, int mode, double center, double half_width,
                                    int precision, int graph, double* moments, void* stream);
/* HIP-event duration (ms) of the step kernels of the last vbh_kpm_moments or
 * vbh_kpm_moments_batched call on this process, and the number of kernel
 * launches (graph nodes) it covers.  Profiling aid; no reference counterpart. */
VBH_API int vbh_kpm_last_step_time(double* ms, int64_t* launches);
/* Tile autotuning of vbh_kpm_moments_batched (on by default).  When on, the
 * first call on a batch (same matrix handles, K, precision, mode) times one
 * recursion step for each candidate plan -- the class threshold (block rows
 * of t..16 rows in the class, t in {4, 8, 17 = none}), the kernel of the
 * class rows (the FP64 matrix cores, or for FP32 values and K in 16, 32, 64
 * the block-row wave kernel on the vector ALUs), and the kernel of the other
 * rows: the lane family with its slot partitions, or the row-wave kernel (one
 * scalar row per wavefront, real types) with its rows per work-group -- and
 * keeps the fastest for every later call on that batch.  Plans change the
 * order of floating-point sums only.  No reference counterpart (the reference
 * has no device tiles). */
VBH_API int vbh_set_autotune(int on);
/* Pin the plan of later vbh_kpm_moments_batched calls (threshold 1..17: the
 * class rows on the matrix cores, -16..-1: class rows of -threshold..16 rows
 * on the block-row wave kernel, where that kernel serves the call; parts: a
 * power of two in 1..64 = lane-family slot partitions, clamped to the lane
 * shape, or its negative = row-wave kernel with -parts rows per work-group,
 * where that kernel serves the call), bypassing the cache and the autotuner;
 * threshold 0 releases it.  For reproducible runs and for tests of every
 * plan.  No reference counterpart. */
VBH_API int vbh_set_batch_plan(int threshold, int parts);
/* Plan of the last vbh_kpm_moments_batched call: class threshold (negative:
 * class rows on the block-row wave kernel), slot partitions (negative:
 * row-wave rows per work-group), and whether it came from timing (1) or the
 * default rule (0). */
VBH_API int vbh_kpm_last_plan(int* threshold, int* parts, int* tuned);
/* Chebyshev series Y = sum_n coef_n T_n(H~) X (complex coefficients as
 * interleaved doubles, nterms of them; real matrices use the real parts).
 * examples/ex09_graphene_dynamics.py:102-133, vbcsr/utils.py:71-126. */
VBH_API int vbh_cheb_apply(vbh_matrix_t m, int K, const void* X, int ldx, void* Y, int ldy,
                           int nterms, const double* coef, double center, double half_width,
                           int precision, void* stream);

/* ---- adjoint of the Chebyshev propagation ---------------------------------
 * Building blocks of the reverse-mode gradient of y = sum_n c_n T_n(H~) x with
 * respect to x and to the stored block values (no reference counterpart: the
 * reference propagates forward only, examples/ex09_graphene_dynamics.py:102-133,
 * vbcsr/utils.py:71-126).  Multivector layouts are row-major with the given
 * leading dimensions; "column block n" of an n_rows x (nterms*K) array is
 * columns [n*K, (n+1)*K).  precision 0 or 1 as for the applies. */
/* Chebyshev basis: column block n of T receives t_n = T_n(H~) X, n < nterms. */
VBH_API int vbh_cheb_basis(vbh_matrix_t m, int K, const void* X, int ldx, int nterms, void* T, int ldt,
                           double center, double half_width, int precision, void* stream);
/* Clenshaw recurrence with H~^H: column block n of B receives
 * b_n = conj(c_n) L + 2 H~^H b_{n+1} - b_{n+2} (n >= 1) and
 * b_0 = conj(c_0) L + H~^H b_1 - b_2, the gradient with respect to X for the
 * cotangent L of y (torch convention: dL/dRe + i dL/dIm).  coef: nterms complex
 * coefficients as interleaved doubles. */
VBH_API int vbh_cheb_adjoint(vbh_matrix_t m, int K, const void* L, int ldl, int nterms, const double* coef,
                             void* B, int ldb, double center, double half_width, int precision,
                             void* stream);
/* Block-sampled outer-product sum (SDDMM): for every stored block (i, j),
 * G_blk[a][b] += alpha * sum_{q<L} U[row_i+a][q] * conj(V[col_j+b][q]).
 * U: n_rows x L (ld ldu), V: n_cols x L (ld ldv), of the precision's vector
 * type; G: device array of nnz values in the matrix's own type and slot order
 * (real matrices accumulate Re).  alpha: complex as 2 doubles. */
VBH_API int vbh_sddmm(vbh_matrix_t m, int L, const void* U, int ldu, const void* V, int ldv,
                      const double* alpha, void* G, int precision, void* stream);

/* ---- vector kernels --------------------------------------------------------
 * rows x K row-major blocks with leading dims; dtype VBH_*.  alpha/beta are
 * (re, im) pairs.  DistMultiVector::axpby/scale/bdot,
 * dist_multivector.hpp:179-206,283-314; DistVector dist_vector.hpp:82-123. */
VBH_API int vbh_vec_axpby(int dtype, int64_t rows, int K, const double* alpha, const void* x, int ldx,
                          const double* beta, void* y, int ldy, void* stream);
VBH_API int vbh_vec_scale(int dtype, int64_t rows, int K, const double* alpha, void* y, int ldy,
                          void* stream);
/* y[r, v] = value for v < K (pad lanes untouched): set_constant,
 * dist_multivector.hpp:170-177. */
VBH_API int vbh_vec_fill(int dtype, int64_t rows, int K, const double* value, void* y, int ldy, void* stream);
VBH_API int vbh_vec_pointwise(int dtype, int64_t rows, int K, const void* x, int ldx, int x_is_vector,
                              void* y, int ldy, void* stream);
/* out[v] = sum_rows conj(x[r,v]) y[r,v]; out is host memory, 2K doubles
 * (re, im) per column. */
VBH_API int vbh_vec_bdot(int dtype, int64_t rows, int K, const void* x, int ldx, const void* y, int ldy,
                         double* out, void* stream);
/* Counter-based Rademacher / random-phase columns keyed by (seed, global
 * block, row in block, column): random_vectors.hpp:62-143,
 * dist_multivector.hpp:335-372.  gids[nblocks] global block ids, bs block
 * sizes; scale multiplies every entry (1/sqrt(N) when normalising). */
VBH_API int vbh_vec_set_random(int dtype, int nblocks, const int* bs_dev, const int* gids_dev,
                               const int64_t* offs_dev, int K, uint64_t seed, int kind, double scale,
                               void* y, int ldy, void* stream);

/* ---- matrix algebra --------------------------------------------------------- */
/* block_csr.hpp:1053 scale; alpha (re, im). */
VBH_API int vbh_matrix_scale(vbh_matrix_t m, const double* alpha, void* stream);
/* block_csr.hpp:1289 shift: alpha added to the diagonal of each row's
 * diagonal block; diag_col[nbr] is the local column of row i's own block
 * (-1 if none). */
VBH_API int vbh_matrix_shift(vbh_matrix_t m, const double* alpha, const int* diag_col_host,
                             void* stream);
/* Frobenius norm of every block, FP64, device output (block_csr.hpp:2486). */
VBH_API int vbh_matrix_block_norms(vbh_matrix_t m, double* norms_dev, void* stream);
/* filter_blocks (block_csr.hpp:1399-1528): when thr > 0 drop blocks with
 * norm < thr.  *out receives a new matrix, or NULL when nothing was dropped. */
VBH_API int vbh_matrix_filter(vbh_matrix_t m, double thr, vbh_matrix_t* out, void* stream);
/* self <- alpha X + beta self (ops/axpby.hpp:254-528): same pattern in place
 * (*out = NULL), or the union pattern into a new matrix *out with self left
 * unchanged. */
VBH_API int vbh_matrix_axpby(vbh_matrix_t self, const double* alpha, vbh_matrix_t X,
                             const double* beta, vbh_matrix_t* out, void* stream);
/* Thresholded block SpGEMM C = A B (BlockSpMat::spmm, block_csr.hpp:1530;
 * symbolic ops/spmm/common.hpp:507-735, numeric ops/spmm/vbcsr.hpp:206-471,
 * filter block_csr.hpp:1399).  A: nbr x nbk blocks, B: nbk x nbc; B's block
 * rows must be A's local columns.  numeric: 0 = native FP64/complex128,
 * 1 = FP32 block products (from FP32 copies of A's and B's values, kept on
 * the device as for precision-1 applies), 2 = BF16 block products with FP32
 * accumulation (real matrices only): the BF16 matrix cores for products
 * whose C fills >= 10 % of its block pattern (one dense GEMM, every
 * candidate pair multiplied) and for blocks over 8 x 8, else BF16-rounded
 * FP32 copies on the vector ALUs.  The threshold filter uses C's block norms
 * from the numeric kernels.  Result in *out.  stats (optional, 4 int64):
 * symbolic pairs, numeric pairs (for the dense BF16 GEMM: the symbolic
 * pairs), C blocks before filter, after. */
VBH_API int vbh_spgemm(vbh_matrix_t A, vbh_matrix_t B, double thr, int numeric, vbh_matrix_t* out,
                       int64_t* stats, void* stream);
/* Upper block triangle of the thresholded SpGEMM (block columns >= the
 * row; BlockSpMat::spmm_hermitian_upper, block_csr.hpp:1605-1616, symbolic
 * upper_only common.hpp:502-513,625): the same rules as vbh_spgemm restricted
 * to j >= i.  Square results without ghost columns. */
VBH_API int vbh_spgemm_upper(vbh_matrix_t A, vbh_matrix_t B, double thr, int numeric, vbh_matrix_t* out,
                             int64_t* stats, void* stream);
/* Hermitian completion of an upper-triangle matrix
 * (BlockSpMat::complete_hermitian, block_csr.hpp:1913-1920; mirror
 * transpose.hpp:139-200; diagonal average block_csr.hpp:1948-1968): blocks on
 * and above the diagonal kept, C_ji = U_ij^H for every strict-upper block,
 * diagonal blocks D <- (D + D^H)/2; blocks below the diagonal are ignored. */
VBH_API int vbh_matrix_complete_hermitian(vbh_matrix_t U, vbh_matrix_t* out, void* stream);
/* C = A B for a product known to be Hermitian: vbh_spgemm_upper then
 * vbh_matrix_complete_hermitian (BlockSpMat::spmm_hermitian,
 * block_csr.hpp:1597-1601).  The result is exactly Hermitian. */
VBH_API int vbh_spgemm_hermitian(vbh_matrix_t A, vbh_matrix_t B, double thr, int numeric, vbh_matrix_t* out,
                                 int64_t* stats, void* stream);
/* FP64 trace and Frobenius norm of (P^2 - P) for idempotency checks:
 * out[0] = tr P, out[1] = ||P2 - P||_F with P2 = P*P given. */
VBH_API int vbh_matrix_trace(vbh_matrix_t m, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* VBCSR_HIP_H */
