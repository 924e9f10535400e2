// Lane-per-row SpMM family for 2 <= K <= 64 and the matrix-core class
// (included by the per-type lanes_*.hip translation units).
//
// Lane family: a wavefront owns R = 64/(S*P) consecutive scalar rows; lane l
// works on row l % R, on the column slice (l / R) % S of KT columns of the K
// right-hand sides, and on slot partition l / (R*S) of the row's blocks (the
// partitions merge by xor shuffles).  Per block the lane reads its row of the
// block once (one 16-byte load for 4x4 FP32 blocks) and, for every block
// column b, the KT contiguous elements of X row (col + b) with vector loads,
// then issues KT FMAs per loaded A element: all K columns share one A load.
// For uniform 4x4 blocks the 4 rows of a block row are a lane quad that
// shares the X rows by DPP broadcast.  Rows are independent, so block rows of
// any size tile the wave; lanes of short block rows finish their loop early.
//
// Matrix-core class (mfma_rows): block rows of kMfmaMinRows..16 rows of real
// matrices, one block row per wavefront on FP64 16x16x4 MFMA tiles (see the
// section below).
//
// Accumulation is in FP64 registers for every storage precision.  The
// epilogue is the fused one of row_loop: Y = alpha*AX + beta*X_own + gamma*Z
// with optional FP64 dots <W|Y>, <X_own|Y> summed over the columns (the
// Chebyshev moments need only the column sums), one partial per work-group.
#pragma once

#include "apply.hpp"

namespace vbh {

template <typename T, int VN>
struct __attribute__((aligned(VN * sizeof(T)))) VecN {
    T v[VN];
};

// Columns [0, nvalid) of a KT-wide slice starting at src, for 1 <= nvalid
// and VN dividing nvalid (the host picks VN dividing K and KT).  Branch-free:
// vector q reads vector min(q, last), always in bounds, so the compiler
// issues every load of the slice before the first wait.  Columns at or past
// nvalid hold a copy of a valid vector (finite); callers never store them or
// count them in a dot product.
template <typename T, int KT, int VN>
__device__ __forceinline__ void load_slice(const T* __restrict__ src, int nvalid, T (&out)[KT]) {
    const int last = nvalid / VN - 1;
#pragma unroll
    for (int q = 0; q < KT / VN; ++q) {
        const VecN<T, VN> w = reinterpret_cast<const VecN<T, VN>*>(src)[q < last ? q : last];
#pragma unroll
        for (int e = 0; e < VN; ++e) out[q * VN + e] = w.v[e];
    }
}

template <typename T, int KT, int VN>
__device__ __forceinline__ void store_slice(T* __restrict__ dst, int nvalid, const T (&in)[KT]) {
#pragma unroll
    for (int q = 0; q < KT / VN; ++q) {
        if ((q + 1) * VN <= nvalid) {
            VecN<T, VN> w;
#pragma unroll
            for (int e = 0; e < VN; ++e) w.v[e] = in[q * VN + e];
            reinterpret_cast<VecN<T, VN>*>(dst)[q] = w;
        }
    }
}

// Epilogue of one row slice (row yr of Y, row xr of X for the X_own terms),
// one VN-vector at a time so that only one vector of each operand is live.
// Vector indices are clamped (in-bounds loads), stores are guarded.
template <typename T, typename V_, int KT, int VN, bool DOT>
__device__ __forceinline__ void emit_slice(const ApplyArgs<T, V_>& g, int64_t yr, int64_t xr, int v0, int nvalid,
                                           const typename Wide<T>::type (&acc)[KT], double& d1, double& d2) {
    using A = typename Wide<T>::type;
    using V = VecN<T, VN>;
    const int last = nvalid / VN - 1;
    const V* xp = reinterpret_cast<const V*>(g.X + xr * g.ldx + v0);
    const V* zp = reinterpret_cast<const V*>(g.Z + yr * g.ldz + v0);
    const V* wp = reinterpret_cast<const V*>(g.W + yr * g.ldw + v0);
    V* yp = reinterpret_cast<V*>(g.Y + yr * g.ldy + v0);
    const bool need_x = DOT || g.beta != 0.0;
    const bool need_z = g.gamma != 0.0;
    const bool need_w = DOT && g.W != nullptr;
#pragma unroll
    for (int q = 0; q < KT / VN; ++q) {
        const int qc = q < last ? q : last;
        V xo{}, zo{}, wo{}, out;
        if (need_x) xo = xp[qc];
        if (need_z) zo = zp[qc];
        if (need_w) wo = wp[qc];
#pragma unroll
        for (int e = 0; e < VN; ++e) {
            A y = rscale(g.alpha, acc[q * VN + e]);
            if (g.beta != 0.0) y = y + rscale(g.beta, widen(xo.v[e]));
            if (need_z) y = y + rscale(g.gamma, widen(zo.v[e]));
            out.v[e] = narrow_to<T>(y);
            if constexpr (DOT) {
                // dots of the stored (rounded) value: the state the recursion carries on
                const A ys = widen(out.v[e]);
                const double t1 = redot(need_w ? widen(wo.v[e]) : ys, ys);
                const double t2 = redot(widen(xo.v[e]), ys);
                d1 += q <= last ? t1 : 0.0;
                d2 += q <= last ? t2 : 0.0;
            }
        }
        if (q <= last) yp[q] = out;
    }
}

// Value of lane (l & ~3) + B of v's quad (DPP quad permutation, any 4/8/16-byte type).
template <int B, typename V>
__device__ __forceinline__ V quad_bcast(V v) {
    static_assert(sizeof(V) % 4 == 0, "32-bit granular types only");
    constexpr int ctrl = B * 0x55;  // quad_perm [B, B, B, B]
    union {
        V v;
        int w[sizeof(V) / 4];
    } in{v}, out;
#pragma unroll
    for (int i = 0; i < int(sizeof(V) / 4); ++i) out.w[i] = __builtin_amdgcn_mov_dpp(in.w[i], ctrl, 0xf, 0xf, false);
    return out.v;
}

// Words (32-bit) of element w of a quad-contiguous X chunk set: chunk q of
// the quad's 4*KT elements (16 bytes each) lives in lane q % 4 of the quad,
// register q / 4.  Element f (0 <= f < 4*KT) of the quad's rows, broadcast
// from its owner lane to the whole quad.
template <typename T, int KT, int F, int NCL>
__device__ __forceinline__ T quad_elem(const vbh_w4 (&xc)[NCL]) {
    constexpr int WPE = int(sizeof(T) / 4);  // words per element
    constexpr int word = F * WPE;
    constexpr int q = word / 4, w = word % 4;
    static_assert(q / 4 < NCL, "chunk out of range");
    union {
        T v;
        int x[WPE];
    } out;
#pragma unroll
    for (int i = 0; i < WPE; ++i) out.x[i] = quad_bcast<q % 4>(xc[q / 4][w + i]);
    return out.v;
}

// acc[j] += sum_b ad[b] * x_b[j] for columns J..KT-1 of a quad-contiguous chunk set.
template <typename T, typename A, int KT, int NCL, int J>
__device__ __forceinline__ void quad_fma_cols(const A (&ad)[4], const vbh_w4 (&xc)[NCL], A (&acc)[KT]) {
    if constexpr (J < KT) {
        A t = Sc<A>::fma(ad[0], widen(quad_elem<T, KT, 0 * KT + J>(xc)), acc[J]);
        t = Sc<A>::fma(ad[1], widen(quad_elem<T, KT, 1 * KT + J>(xc)), t);
        t = Sc<A>::fma(ad[2], widen(quad_elem<T, KT, 2 * KT + J>(xc)), t);
        acc[J] = Sc<A>::fma(ad[3], widen(quad_elem<T, KT, 3 * KT + J>(xc)), t);
        quad_fma_cols<T, A, KT, NCL, J + 1>(ad, xc, acc);
    }
}

// Slot loop of one lane for uniform 4x4 blocks.  The 4 rows of a block row
// sit in one quad of lanes.  Column indices of 4*CB slots are loaded
// lane-parallel up front (lane a holds slots a, a+4, ...) and complete
// together, and U slots are issued per batch -- their values, then their X
// chunks, in one basic block -- so a row of 10 blocks costs about 1 + 10/U
// memory latencies (measured: U = 2 beats 1 and 4 for 10 FP32 columns).
// Against a loop that waited for each batch's column index and values before
// addressing its X rows: the 200 x 200 graphene KPM step 26.4 -> 24.5 us
// (FP32) and 46.1 -> 38.3 us (FP64); on the 800 x 800 lattice (HBM-resident)
// FP32 422 -> 431 us, and FP64 760 -> 906 us with 2-slot batches or 720 us
// with 4-slot batches (QB = 4, chosen by the launcher).  Two ways to read a block's 4
// X rows (QX, a template parameter: a branch on it inside the batch would
// make the compiler wait for the value loads before the X loads):
//  * quad-contiguous (QX: one lane slice of all K columns, rows stored
//    back to back, ldx == K, X 16-byte aligned -- the KPM recursion
//    buffers): the 4 rows are 4*K contiguous elements, read as 16-byte
//    chunks, chunk q by lane q % 4 of the quad, so each load instruction
//    covers 64 contiguous bytes per quad (the texture addresser's cost goes
//    with the cache lines an instruction touches);
//  * per row (otherwise): lane a reads X row 4*col + a.
// Either way each element reaches the whole quad by DPP broadcast, so X
// traffic is not replicated across the block's rows.  Slots past the
// partition's count are clamped to its last slot and enter with zero A
// values (one select per A value instead of one per product).
template <typename T, typename V, int KT, int VN, bool QX, int QB>
__device__ __forceinline__ void quad_slots(const ApplyArgs<T, V>& g, int kb, int cnt, int stride, int a, int v0,
                                           int nvalid, typename Wide<T>::type (&acc)[KT]) {
    using A = typename Wide<T>::type;
    constexpr int CB = 4;                                   // slots per lane per pass
    constexpr int per = int(KT * sizeof(T) / 4);            // VGPRs of one X slice
    // slots per batch: QB when given, else by the slice width
    constexpr int U = QB > 0 ? QB : per <= 8 ? 4 : per <= 24 ? 2 : 1;
    constexpr int NC = int(4 * KT * sizeof(T) / 16);        // 16-byte chunks of a quad's 4 X rows
    constexpr int NCL = (NC + 3) / 4;                       // chunks per lane
    static_assert(4 * KT * sizeof(T) % 16 == 0, "quad rows must fill whole 16-byte chunks");
    const DevPattern& p = g.p;
    // logical slot j of this lane's partition is physical slot kb + j*stride
    for (int j0 = 0; j0 < cnt; j0 += 4 * CB) {
        // ns >= 1; slot indices past the partition are clamped to its last
        // slot, so every load below is unconditional and in bounds
        const int ns = cnt - j0 < 4 * CB ? cnt - j0 : 4 * CB;
        const int k0 = kb + j0 * stride;
        int colv[CB];
#pragma unroll
        for (int t = 0; t < CB; ++t) colv[t] = p.col_ind[k0 + (a + 4 * t < ns ? a + 4 * t : ns - 1) * stride];
        // the four column loads complete here, together: without this the
        // compiler sinks each into the batch that first uses it, and every
        // batch then waits for its column index before it can even address
        // its X rows (two dependent memory latencies per batch)
        static_assert(CB == 4, "column preload of four registers");
        asm volatile("" : "+v"(colv[0]), "+v"(colv[1]), "+v"(colv[2]), "+v"(colv[3]));
#pragma unroll
        for (int s0 = 0; s0 < 4 * CB; s0 += U) {
            if (s0 >= ns) break;
            V av[U][4];  // raw values, widened where they are used, after the X loads are issued
            int col[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int s = s0 + u;
                switch (s & 3) {
                    case 0: col[u] = quad_bcast<0>(colv[s >> 2]); break;
                    case 1: col[u] = quad_bcast<1>(colv[s >> 2]); break;
                    case 2: col[u] = quad_bcast<2>(colv[s >> 2]); break;
                    default: col[u] = quad_bcast<3>(colv[s >> 2]); break;
                }
                const int sc = s < ns ? s : ns - 1;
                const V* ar = static_cast<const V*>(__builtin_assume_aligned(
                    g.vals + int64_t(k0 + sc * stride) * 16 + a * 4, (4 * sizeof(V)) % 16 ? sizeof(V) : 16));
#pragma unroll
                for (int b = 0; b < 4; ++b) av[u][b] = ar[b];
            }
            // value loads issue before the X loads (the first products wait
            // for the oldest loads only)
            asm volatile("" ::: "memory");
            auto widened = [&](int u, A (&ad)[4]) {
                const bool on = s0 + u < ns;
#pragma unroll
                for (int b = 0; b < 4; ++b) ad[b] = widen(pick(on, av[u][b], Sc<V>::zero()));
            };
            // (the X layout is a template parameter: a branch on it here would
            // make the compiler wait for the value loads before the X loads)
            if constexpr (QX) {
                vbh_w4 xc[U][NCL];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const vbh_w4* src = reinterpret_cast<const vbh_w4*>(g.X + int64_t(col[u]) * 4 * KT);
#pragma unroll
                    for (int t = 0; t < NCL; ++t) xc[u][t] = src[a + 4 * t < NC ? a + 4 * t : NC - 1];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    A ad[4];
                    widened(u, ad);
                    quad_fma_cols<T, A, KT, NCL, 0>(ad, xc[u], acc);
                }
            } else {
                T xr[U][KT];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    load_slice<T, KT, VN>(g.X + (int64_t(col[u]) * 4 + a) * g.ldx + v0, nvalid, xr[u]);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    A ad[4];
                    widened(u, ad);
#pragma unroll
                    for (int j = 0; j < KT; ++j) {
                        const T x0 = quad_bcast<0>(xr[u][j]), x1 = quad_bcast<1>(xr[u][j]);
                        const T x2 = quad_bcast<2>(xr[u][j]), x3 = quad_bcast<3>(xr[u][j]);
                        A t = Sc<A>::fma(ad[0], widen(x0), acc[j]);
                        t = Sc<A>::fma(ad[1], widen(x1), t);
                        t = Sc<A>::fma(ad[2], widen(x2), t);
                        acc[j] = Sc<A>::fma(ad[3], widen(x3), t);
                    }
                }
            }
        }
    }
}

// Rows [sb, se) of the matrix, wave-interleaved in segments of R = 64/S
// rows; lane l works on row l % R of the segment and column slice l / R, so
// the rows of a block row occupy consecutive lanes.  BS > 0: uniform BS x BS
// blocks (BS == 4: quad-shared X rows, quad_slots).
template <typename T, typename V, int BS, int S, int KT, int VN, bool DOT, int QB = 0>
__device__ __forceinline__ void lane_rows(const ApplyArgs<T, V>& g, int64_t sb, int64_t se, int lane, int wave,
                                          double& d1, double& d2) {
    using A = typename Wide<T>::type;
    const DevPattern& p = g.p;
    // lane = (part * S + slice) * R + row: P partitions of each row's slots
    const int P = g.parts;
    const int R = 64 / (S * P);
    const int grp = lane / R;
    const int part = grp / S;
    const int v0 = (grp - part * S) * KT;
    const int nvalid = g.K - v0;
    for (int64_t base = sb + int64_t(wave) * R; base < se; base += int64_t(kWaves) * R) {
        const int64_t idx = base + lane % R;
        const bool live = idx < se && nvalid > 0;
        int64_t srow = idx;
        if constexpr (BS == 0) {
            if (g.row_list && live) srow = p.small_rows[idx];
        }
        int kb = 0, ke = 0, r = BS, a = 0;
        int64_t vrow = 0, xown = 0;
        if (live) {
            if constexpr (BS > 0) {
                const int i = int(srow / BS);
                kb = p.row_ptr[i];
                ke = p.row_ptr[i + 1];
                a = int(srow - int64_t(i) * BS);
                vrow = int64_t(kb) * (BS * BS) + a * BS;
                xown = srow;
            } else {
                const int i = p.row_of[srow];
                kb = p.row_ptr[i];
                ke = p.row_ptr[i + 1];
                r = p.rbs[i];
                a = int(srow - p.roff[i]);
                vrow = p.voff[kb];
                xown = int64_t(p.coff[i]) + a;
            }
        }
        A acc[KT];
#pragma unroll
        for (int j = 0; j < KT; ++j) acc[j] = Sc<A>::zero();
        if constexpr (BS == 4) {
            const int kp = kb + part;
            // quad-contiguous X rows: one slice of all K columns, rows back to
            // back, 16-byte aligned (wave-uniform)
            const bool qx = S == 1 && nvalid == KT && g.ldx == KT && (reinterpret_cast<uintptr_t>(g.X) & 15) == 0;
            const int cnt = ke > kp ? (ke - kp + P - 1) / P : 0;
            if (qx) quad_slots<T, V, KT, VN, true, QB>(g, kp, cnt, P, a, v0, nvalid, acc);
            else quad_slots<T, V, KT, VN, false, QB>(g, kp, cnt, P, a, v0, nvalid, acc);
        } else if constexpr (BS > 0) {
            for (int k = kb + part; k < ke; k += P) {
                const int col = p.col_ind[k];
                const V* ar = static_cast<const V*>(
                    __builtin_assume_aligned(g.vals + vrow + int64_t(k - kb) * (BS * BS), (BS * sizeof(V)) % 16 ? sizeof(V) : 16));
                const T* xb = g.X + int64_t(col) * BS * g.ldx + v0;
                V av[BS];
#pragma unroll
                for (int b = 0; b < BS; ++b) av[b] = ar[b];
#pragma unroll
                for (int b = 0; b < BS; ++b) {
                    T xv[KT];
                    load_slice<T, KT, VN>(xb + int64_t(b) * g.ldx, nvalid, xv);
                    const A ab = widen(av[b]);
#pragma unroll
                    for (int j = 0; j < KT; ++j) acc[j] = Sc<A>::fma(ab, widen(xv[j]), acc[j]);
                }
            }
        } else {
            // variable blocks: the lane's row of each block is read in batches of
            // BU columns with clamped (in-bounds) indices, all loads of a batch
            // issued before its FMAs; the next slot's metadata is fetched ahead
            constexpr int per = int(KT * sizeof(T) / 4);
#ifdef VBH_GEN_BATCH
            constexpr int BU = VBH_GEN_BATCH;
#else
            constexpr int BU = per <= 8 ? 4 : 2;
#endif
            const int kp = kb + part;
            int2 mn = p.meta[kp < ke ? kp : 0];
            for (int k = kp; k < ke; k += P) {
                const int2 mm = mn;
                mn = p.meta[k + P < ke ? k + P : k];
                const int c = mm.y & 255;
                const V* ar = g.vals + vrow + int64_t(r) * (mm.y >> 8) + a * c;
                const T* xb = g.X + int64_t(mm.x) * g.ldx + v0;
                for (int b0 = 0; b0 < c; b0 += BU) {
                    V av[BU];
                    T xv[BU][KT];
#pragma unroll
                    for (int u = 0; u < BU; ++u) {
                        const int b = b0 + u < c ? b0 + u : c - 1;
                        av[u] = ar[b];
                        load_slice<T, KT, VN>(xb + int64_t(b) * g.ldx, nvalid, xv[u]);
                    }
#pragma unroll
                    for (int u = 0; u < BU; ++u) {
                        const A ab = widen(av[u]);
                        const bool on = b0 + u < c;
#pragma unroll
                        for (int j = 0; j < KT; ++j) {
                            const A t = Sc<A>::fma(ab, widen(xv[u][j]), acc[j]);
                            acc[j] = pick(on, t, acc[j]);
                        }
                    }
                }
            }
        }
        // merge the partitions (lanes 64/P apart) into part 0
        for (int m = 64 / P; m < 64; m <<= 1) {
#pragma unroll
            for (int j = 0; j < KT; ++j) acc[j] = acc[j] + shfl_xor(acc[j], m);
        }
        if (live && part == 0) emit_slice<T, V, KT, VN, DOT>(g, srow, xown, v0, nvalid, acc, d1, d2);
    }
}

template <typename T, typename V, int BS, int S, int KT, int VN, bool DOT, int QB = 0>
__global__ __launch_bounds__(kThreads)
void k_apply_lanes(ApplyArgs<T, V> g, int64_t n_rows) {
    const int RPG = kWaves * (64 / (S * g.parts));
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int chunk = xcd_chunk(g.G);
    const int64_t sb = int64_t(chunk) * RPG;
    const int64_t se = sb + RPG < n_rows ? sb + RPG : n_rows;
    double d1 = 0.0, d2 = 0.0;
    lane_rows<T, V, BS, S, KT, VN, DOT, QB>(g, sb, se, lane, wave, d1, d2);
    if constexpr (DOT) write_partials(1, lane, wave, d1, d2, g.partials + int64_t(chunk) * 2);
}

// Apply arguments of batch item lo for one recursion step (buffers rotate
// t_{n-1}, t_n, t_{n+1} through buf[3]).  The item table is read through the
// constant address space: it is read-only during the launch, and the
// compiler then takes the pointers loaded from it to be global (global_load
// rather than flat_load, whose completion also holds the LDS/scalar wait
// counter).
template <typename T, typename V>
__device__ __forceinline__ ApplyArgs<T, V> batch_args(const BatchItem<T, V>* items, int lo, int K, int ld, int step,
                                                      double alpha, double beta, double gamma0, double gamma,
                                                      int mode) {
    using CItem = const __attribute__((address_space(4))) BatchItem<T, V>;
    CItem& it = ((CItem*)items)[lo];
    ApplyArgs<T, V> a{};
    a.p.nbr = it.p.nbr;
    a.p.nbc = it.p.nbc;
    a.p.bs = it.p.bs;
    a.p.rbs = it.p.rbs;
    a.p.cbs = it.p.cbs;
    a.p.row_ptr = it.p.row_ptr;
    a.p.col_ind = it.p.col_ind;
    a.p.roff = it.p.roff;
    a.p.coff = it.p.coff;
    a.p.voff = it.p.voff;
    a.p.meta = it.p.meta;
    a.p.row_of = it.p.row_of;
    a.p.small_rows = it.p.small_rows;
    a.p.big_desc = it.p.big_desc;
    a.p.n_small = it.p.n_small;
    a.p.n_big = it.p.n_big;
    a.vals = it.vals;
    a.K = K;
    a.X = it.buf[step % 3];
    a.ldx = ld;
    a.Y = it.buf[(step + 1) % 3];
    a.ldy = ld;
    a.alpha = alpha;
    a.beta = beta;
    a.gamma = step == 0 ? gamma0 : gamma;
    a.Z = it.buf[(step + 2) % 3];
    a.ldz = ld;
    a.W = mode == 1 ? it.v0 : nullptr;
    a.ldw = ld;
    return a;
}

// Largest item index whose base (g_base or g_base2) is <= gg, uniform.
template <typename T, typename V, bool SECOND>
__device__ __forceinline__ int batch_item(const BatchItem<T, V>* items, int nitems, int gg) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((SECOND ? items[mid].g_base2 : items[mid].g_base) <= gg) lo = mid;
        else hi = mid - 1;
    }
    return uniform(lo);
}

// One KPM recursion step over every matrix of a batch (see kpm.hip): work-group
// gg finds its matrix by binary search over g_base; groups past the last
// matrix's rows write zero partials.
template <typename T, typename V, int S, int KT, int VN>
__global__ __launch_bounds__(kThreads)
void k_batched_lanes(const BatchItem<T, V>* __restrict__ items, int nitems,
                                                            int Gtot, int K, int ld, int step, double alpha,
                                                            double beta, double gamma0, double gamma, int mode,
                                                            double* __restrict__ partials, int parts) {
    const int RPG = kWaves * (64 / (S * parts));
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int gg = xcd_chunk(Gtot);
    const int lo = batch_item<T, V, false>(items, nitems, gg);
    using CItem = const __attribute__((address_space(4))) BatchItem<T, V>;
    CItem& it = ((CItem*)items)[lo];
    ApplyArgs<T, V> a = batch_args(items, lo, K, ld, step, alpha, beta, gamma0, gamma, mode);
    a.ldw = ld;
    a.row_list = it.row_list;
    a.parts = parts;
    const int64_t sb = int64_t(gg - it.g_base) * RPG;
    const int64_t se = sb + RPG < it.lane_rows ? sb + RPG : it.lane_rows;
    double d1 = 0.0, d2 = 0.0;
    lane_rows<T, V, 0, S, KT, VN, true>(a, sb, se, lane, wave, d1, d2);
    write_partials(1, lane, wave, d1, d2, partials + int64_t(gg) * 2);
}

// ------------------------------------------------------------ matrix-core class
// Block rows of kMfmaMinRows..16 rows (real types): one wavefront per block row computes
// the r x K output as ceil(K/16) tiles of 16 x 16 with v_mfma_f64_16x16x4_f64
// (A operand: lane l holds row l%16, block column kk + l/16; B operand: lane l
// holds X row kk + l/16, column l%16; D: register e of lane l holds row
// 4e + l/16 of column l%16 -- the FP64 layout, unlike the FP32 16x16x4 one
// with rows 4(l/16) + e).  Each X element of a block is loaded once per block
// -- not once per row as in the lane family -- and the FP32 values of
// precision 1/2 are widened exactly, so the products accumulate in FP64 as
// everywhere else.  Rows past r and block columns past c enter as zeros
// (clamped loads + selects, no guarded loads).
typedef double vbh_d4 __attribute__((ext_vector_type(4)));

// Operand layout.  The k index of the MFMAs runs over a block's columns in
// the order b(s, lk) = 8*(s/2) + 2*lk + s%2 (k-step s, lane group lk), so a
// lane's A values of k-steps 2m and 2m+1 are adjacent, A[li][8m + 2lk + {0,1}]
// (one 2-wide load per pair of k-steps); a block of c columns needs k-steps
// s with 8*(s/2) + s%2 < c.  The output columns are interleaved: tile t,
// column j is X column NT*j + t, so lane j's NT columns of an X row are
// adjacent (one NT-wide load per k-step, NT-wide stores in the epilogue).
// VX: K % NT == 0, the X rows are read with NT-wide loads; otherwise column
// by column.
__device__ __forceinline__ int mfma_col(int s, int lk) { return 8 * (s >> 1) + 2 * lk + (s & 1); }

template <typename T, int N>
__device__ __forceinline__ void load_n(const T* src, T (&out)[N]) {
    typedef T vec __attribute__((ext_vector_type(N)));
    vec v;
    __builtin_memcpy(&v, src, sizeof(vec));  // element-aligned: the target reads it unaligned
#pragma unroll
    for (int e = 0; e < N; ++e) out[e] = v[e];
}
// The same as a non-temporal (evict-first) load, for data read once per
// launch: the block values, whose lines should not displace gathered X rows.
template <typename T, int N>
__device__ __forceinline__ void load_n_stream(const T* src, T (&out)[N]) {
    typedef T vec __attribute__((ext_vector_type(N)));
    typedef T uvec __attribute__((ext_vector_type(N), aligned(sizeof(T))));
    const vec v = __builtin_nontemporal_load(reinterpret_cast<const uvec*>(src));
#pragma unroll
    for (int e = 0; e < N; ++e) out[e] = v[e];
}
template <typename T, int N>
__device__ __forceinline__ void store_n(T* dst, const T (&in)[N]) {
    typedef T vec __attribute__((ext_vector_type(N)));
    vec v;
#pragma unroll
    for (int e = 0; e < N; ++e) v[e] = in[e];
    __builtin_memcpy(dst, &v, sizeof(vec));
}

// ------------------------------------------------------------ row-wave family
// Rows of the lane class of a batched step, one scalar row per wavefront at a
// time.  The K columns split into LQ = K / CPL groups of CPL contiguous
// columns (16 bytes), and the 64 lanes into LP = 64 / LQ partitions: lane
// (partition p, group q) = p * LQ + q.  The row's blocks are flattened into
// entries, one per stored column of the row (X row, value offset of the
// block's first row at that column, block width), written to a per-wave LDS
// list for up to kRowWaveEnt entries at a time; each step then handles LP
// entries, one per partition, so a load instruction reads LP whole X rows
// (each partition's lanes read one row, contiguous) -- the lane family reads
// one slice per lane, and its load instructions touch one cache line per
// lane.  Partitions merge by xor shuffles; the epilogue is the fused one.
constexpr int kRowWaveEnt = 512;  // LDS entries per wave

template <typename T, typename V, int CPL>
__device__ __forceinline__ void rowwave_rows(const ApplyArgs<T, V>& g, int64_t sb, int64_t se, int lane, int wave,
                                             int2* __restrict__ ent, double& d1, double& d2) {
    using A = typename Wide<T>::type;
    const DevPattern& p = g.p;
    const int LQ = g.K / CPL;
    const int lgq = __builtin_ctz(LQ);
    const int LP = 64 >> lgq;
    const int q = lane & (LQ - 1), part = lane >> lgq;
    for (int64_t idx = sb + wave; idx < se; idx += kWaves) {
        const int64_t srow = g.row_list ? int64_t(p.small_rows[idx]) : idx;
        const int i = p.row_of[srow];
        const int kb = __builtin_amdgcn_readfirstlane(p.row_ptr[i]);
        const int ke = __builtin_amdgcn_readfirstlane(p.row_ptr[i + 1]);
        const int r = p.rbs[i];
        const int a = int(srow - p.roff[i]);
        const int64_t vrow = p.voff[kb];
        const int64_t xown = int64_t(p.coff[i]) + a;
        A acc[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) acc[j] = Sc<A>::zero();
        for (int t0 = kb; t0 < ke; t0 += 64) {
            const int nb = ke - t0 < 64 ? ke - t0 : 64;
            const int2 mm = p.meta[t0 + (lane < nb ? lane : nb - 1)];
            const int xo = mm.x, c = mm.y & 255, panel = mm.y >> 8;
            for (int u0 = 0; u0 < nb;) {
                // the run of slots from u0 whose entries fit the list
                const int panel0 = __builtin_amdgcn_readlane(panel, u0);
                const bool fits = lane >= u0 && lane < nb && panel + c - panel0 <= kRowWaveEnt;
                const int u1 = u0 + __popcll(__ballot(fits));
                const int W = __builtin_amdgcn_readlane(panel + c, u1 - 1) - panel0;
                if (lane >= u0 && lane < u1) {
                    for (int b = 0; b < c; ++b)
                        ent[panel - panel0 + b] = make_int2(int(uint32_t(xo + b) | (uint32_t(c) << 24)), r * panel + b);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // two steps per pass, all their loads issued first; entries
                // past W repeat the last one with a zero value
                for (int e0 = 0; e0 < W; e0 += 2 * LP) {
                    A av[2];
                    T xv[2][CPL];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int e = e0 + h * LP + part;
                        const int2 en = ent[e < W ? e : W - 1];
                        const int xrow = en.x & 0xffffff, cw = int(uint32_t(en.x) >> 24);
                        const V vv = g.vals[vrow + en.y + a * cw];
                        av[h] = e < W ? widen(vv) : Sc<A>::zero();
                        load_n<T, CPL>(g.X + int64_t(xrow) * g.ldx + CPL * q, xv[h]);
                    }
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int j = 0; j < CPL; ++j) acc[j] = Sc<A>::fma(av[h], widen(xv[h][j]), acc[j]);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                u0 = u1;
            }
        }
        for (int m = LQ; m < 64; m <<= 1) {
#pragma unroll
            for (int j = 0; j < CPL; ++j) acc[j] = acc[j] + shfl_xor(acc[j], m);
        }
        if (part == 0) emit_slice<T, V, CPL, CPL, true>(g, srow, xown, CPL * q, CPL, acc, d1, d2);
    }
}

// Batched step over the lane-class rows with the row-wave kernel: rw rows
// per work-group.
template <typename T, typename V, int CPL>
__global__ __launch_bounds__(kThreads) void k_batched_rowwave(const BatchItem<T, V>* __restrict__ items, int nitems,
                                                              int Gtot, int K, int ld, int rw, int step, double alpha,
                                                              double beta, double gamma0, double gamma, int mode,
                                                              double* __restrict__ partials) {
    __shared__ int2 ents[kWaves][kRowWaveEnt];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int gg = xcd_chunk(Gtot);
    const int lo = batch_item<T, V, false>(items, nitems, gg);
    using CItem = const __attribute__((address_space(4))) BatchItem<T, V>;
    CItem& it = ((CItem*)items)[lo];
    ApplyArgs<T, V> a = batch_args(items, lo, K, ld, step, alpha, beta, gamma0, gamma, mode);
    a.ldw = ld;
    a.row_list = it.row_list;
    const int64_t sb = int64_t(gg - it.g_base) * rw;
    const int64_t se = sb + rw < it.lane_rows ? sb + rw : it.lane_rows;
    double d1 = 0.0, d2 = 0.0;
    rowwave_rows<T, V, CPL>(a, sb, se, lane, wave, ents[wave], d1, d2);
    write_partials(1, lane, wave, d1, d2, partials + int64_t(gg) * 2);
}

// ------------------------------------------------------------ block-row wave family
// The matrix-core class rows of a batched step on the vector ALUs, one block
// row per wavefront at a time.  The wave's four 16-lane rows are four
// partitions; lane q of a partition holds columns CPL*q .. CPL*q+CPL-1 of the
// K = 16*CPL right-hand sides for every row of the block row.  The block
// row's stored scalar columns ("entries": block by block, column by column)
// go round-robin to the partitions.  For entry e a partition loads its X row
// (the 16 lanes read the whole row, contiguous) and column e of the block row
// from the matrix's column-major FP32 copy, where the r values of a column
// are contiguous (lane a < r takes row a); row a's value then reaches the 16
// lanes by a DPP row broadcast folded into the FMA (v_fmac_f64 with
// row_newbcast:a).  Every product is a useful one: no tile padding, and each
// X row is read once per stored column, not once per row of the block.  The
// row count is a compile-time constant of the loop (one instance per
// r = 1..16).  At the end of a row the partitions reduce-scatter their sums
// (partition p keeps rows a = p mod 4) and write them through the fused
// epilogue.

// acc[j] += (a of lane A_ of this 16-lane row) * x[j], j < CPL, as one asm
// block: v_fmac_f64 with a DPP row broadcast of its first source.  The
// compiler does not see inside the block, so its hazard recognizer cannot
// pad for it: a VALU write of any VGPR that a DPP instruction reads needs 2
// wait states before it, which the leading s_nop 1 provides whatever the
// scheduler puts before the block (the FMAs inside write only their own
// accumulators, which no later instruction of the block reads).
template <int A_, int CPL>
__device__ __forceinline__ void fmac_rowbcast(double (&acc)[CPL], double a, const double (&x)[CPL]) {
    if constexpr (CPL == 1) {
        asm("s_nop 1\n\tv_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
            : "+v"(acc[0])
            : "v"(a), "v"(x[0]), "n"(A_));
    } else if constexpr (CPL == 2) {
        asm("s_nop 1\n\tv_fmac_f64_dpp %0, %2, %3 row_newbcast:%5 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %1, %2, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf"
            : "+v"(acc[0]), "+v"(acc[1])
            : "v"(a), "v"(x[0]), "v"(x[1]), "n"(A_));
    } else {
        static_assert(CPL == 4, "block-row wave kernel: CPL in 1, 2, 4");
        asm("s_nop 1\n\tv_fmac_f64_dpp %0, %4, %5 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %1, %4, %6 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %2, %4, %7 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %3, %4, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf"
            : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3])
            : "v"(a), "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "n"(A_));
    }
}
template <int CPL, int R4, int... As>
__device__ __forceinline__ void brw_fmas(double (&acc)[R4][CPL], double a, const double (&x)[CPL],
                                         std::integer_sequence<int, As...>) {
    (fmac_rowbcast<As, CPL>(acc[As], a, x), ...);
}

template <typename T, int R, int CPL>
__device__ __forceinline__ void brw_row(const ApplyArgs<T, float>& g, const float* __restrict__ vt,
                                        const int* __restrict__ ecol, int64_t eb, int ne, int64_t vrow, int64_t row0,
                                        int64_t x0r, int lane, double& d1, double& d2) {
    constexpr int R4 = (R + 3) & ~3;  // rows padded to the reduce-scatter groups (rows >= R stay zero)
    const int q = lane & 15, part = lane >> 4;
    const float* vcol = vt + vrow + (q < R ? q : R - 1);  // + R * e: this lane's row of column e
    double acc[R4][CPL];
#pragma unroll
    for (int a = 0; a < R4; ++a)
#pragma unroll
        for (int j = 0; j < CPL; ++j) acc[a][j] = 0.0;
    // Software pipeline over passes of U entries per partition (4U per wave;
    // more for short rows, whose passes have few FMAs to cover the loads):
    // while the FMAs of pass j run, the X rows and values of pass j+1 are in
    // flight and the entry columns of pass j+2 are being read.  Entries past
    // ne repeat the last one (in bounds) with a zero value.
    // (CPL >= 2: 4 entries per partition up to 9 rows and 3 above, within
    // the 128 VGPRs of 4 waves per SIMD; against 4 up to 5 rows and 2 above,
    // the C4 step (blocks of 1, 4, 9, 13 rows, K = 32) 497 -> 447-457 us)
    constexpr int U = CPL >= 2 ? (R <= 2 ? 8 : R <= 9 ? 4 : 3) : (R <= 2 ? 8 : R <= 5 ? 4 : 2);
    constexpr int W = 4 * U;  // entries per pass
    auto ent = [&](int e) { return e < ne ? e : ne - 1; };
    int xr[U];             // entry columns of the next pass
    float an[U];           // the next pass's values and X rows, as loaded
    T xsn[U][CPL];
#pragma unroll
    for (int h = 0; h < U; ++h) {
        const int x0 = ecol[eb + ent(4 * h + part)];
        xr[h] = ecol[eb + ent(W + 4 * h + part)];
        an[h] = vcol[int64_t(R) * ent(4 * h + part)];
        load_n<T, CPL>(g.X + int64_t(x0) * g.ldx + CPL * q, xsn[h]);
    }
    // DPP after a write of EXEC needs 5 wait states: the loop runs with the
    // whole wave active (the row is wave-uniform), and this pads its entry
    asm volatile("s_nop 4");
    for (int e0 = 0; e0 < ne; e0 += W) {
        double av[U], xv[U][CPL];
#pragma unroll
        for (int h = 0; h < U; ++h) {
            av[h] = e0 + 4 * h + part < ne ? double(an[h]) : 0.0;
#pragma unroll
            for (int j = 0; j < CPL; ++j) xv[h][j] = double(xsn[h][j]);
        }
        int xn[U];
#pragma unroll
        for (int h = 0; h < U; ++h) {
            xn[h] = ecol[eb + ent(e0 + 2 * W + 4 * h + part)];
            an[h] = vcol[int64_t(R) * ent(e0 + W + 4 * h + part)];
            load_n<T, CPL>(g.X + int64_t(xr[h]) * g.ldx + CPL * q, xsn[h]);
        }
#pragma unroll
        for (int h = 0; h < U; ++h) brw_fmas<CPL, R4>(acc, av[h], xv[h], std::make_integer_sequence<int, R>{});
#pragma unroll
        for (int h = 0; h < U; ++h) xr[h] = xn[h];
    }
    // reduce-scatter over the partitions: lanes 0-31 keep rows a with bit 1
    // clear, lanes 32-63 those with it set; then within each half the even
    // and odd partitions split bit 0.  Partition p ends with rows 4b + p.
    const bool hi = part >= 2, odd = part & 1;
    double k1[R4 / 4][2][CPL];
#pragma unroll
    for (int b = 0; b < R4 / 4; ++b)
#pragma unroll
        for (int s0 = 0; s0 < 2; ++s0)
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
                const double keep = hi ? acc[4 * b + 2 + s0][j] : acc[4 * b + s0][j];
                const double send = hi ? acc[4 * b + s0][j] : acc[4 * b + 2 + s0][j];
                k1[b][s0][j] = keep + shfl_xor(send, 32);
            }
#pragma unroll
    for (int b = 0; b < R4 / 4; ++b) {
        double k2[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const double keep = odd ? k1[b][1][j] : k1[b][0][j];
            const double send = odd ? k1[b][0][j] : k1[b][1][j];
            k2[j] = keep + shfl_xor(send, 16);
        }
        const int a = 4 * b + part;
        if (a < R) emit_slice<T, float, CPL, CPL, true>(g, row0 + a, x0r + a, CPL * q, CPL, k2, d1, d2);
    }
}

// Batched step over the matrix-core class rows with the block-row wave
// kernel (precisions 1 and 2: FP32 values): work-group gg takes
// kMfmaRowsPerGroup entries of its matrix's class-row list, one per wave.
// (Splitting a row's K columns over two waves halves the accumulators and
// fits 7 waves per SIMD instead of 4, but measured 25 % slower on C4.)
template <typename T, int CPL>
__global__ __launch_bounds__(kThreads) void k_batched_brw(const BatchItem<T, float>* __restrict__ items, int nitems,
                                                          int Gtot, int K, int ld, int step, double alpha,
                                                          double beta, double gamma0, double gamma, int mode,
                                                          double* __restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int gg = xcd_chunk(Gtot);
    const int lo = batch_item<T, float, true>(items, nitems, gg);
    using CItem = const __attribute__((address_space(4))) BatchItem<T, float>;
    CItem& it = ((CItem*)items)[lo];
    ApplyArgs<T, float> a = batch_args(items, lo, K, ld, step, alpha, beta, gamma0, gamma, mode);
    const int64_t q = int64_t(gg - it.g_base2) * kMfmaRowsPerGroup + __builtin_amdgcn_readfirstlane(wave);
    double d1 = 0.0, d2 = 0.0;
    if (it.row_list && q < a.p.n_big) {
        const int4 h0 = a.p.big_desc[2 * q], h1 = a.p.big_desc[2 * q + 1];
        const int r = __builtin_amdgcn_readfirstlane(h0.z);
        const int64_t row0 = __builtin_amdgcn_readfirstlane(h0.w), x0r = __builtin_amdgcn_readfirstlane(h1.x);
        const int64_t vrow = int64_t((uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(h1.z))) << 32) |
                                     uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(h1.y))));
        const int i = __builtin_amdgcn_readfirstlane(h1.w);
        const int64_t eb = it.eoff[i];
        const int ne = int(it.eoff[i + 1] - eb);
        const float* vt = it.vt;
        const int* ecol = it.ecol;
        if (ne > 0) {
            // class rows have 1 <= r <= 16 (the row-class lists hold no others)
            switch (r) {
#define VBH_BRW_CASE(RR) \
    case RR: brw_row<T, RR, CPL>(a, vt, ecol, eb, ne, vrow, row0, x0r, lane, d1, d2); break;
                VBH_BRW_CASE(1) VBH_BRW_CASE(2) VBH_BRW_CASE(3) VBH_BRW_CASE(4) VBH_BRW_CASE(5) VBH_BRW_CASE(6)
                VBH_BRW_CASE(7) VBH_BRW_CASE(8) VBH_BRW_CASE(9) VBH_BRW_CASE(10) VBH_BRW_CASE(11)
                VBH_BRW_CASE(12) VBH_BRW_CASE(13) VBH_BRW_CASE(14) VBH_BRW_CASE(15) VBH_BRW_CASE(16)
#undef VBH_BRW_CASE
                default: break;
            }
        } else {
            // a block row without stored blocks: Y = beta X + gamma Z
            const double zero[CPL] = {};
            for (int ar = (lane >> 4); ar < r; ar += 4)
                emit_slice<T, float, CPL, CPL, true>(a, row0 + ar, x0r + ar, CPL * (lane & 15), CPL, zero, d1, d2);
        }
    }
    write_partials(1, lane, wave, d1, d2, partials + int64_t(gg) * 2);
}

// Operands of slot u of the current 64-slot chunk, loaded straight into the
// MFMA layout: a[m][e] = A[li][8m + 2lk + e], x[s][t] = X[xo + b(s, lk)][NT*li + t].
// Indices are clamped into the block and into K, so every load is
// unconditional and in bounds (a pair that starts on a row's last value reads
// one value past it, still inside the value array and its padding), and the
// number of loads per slot is fixed (the compiler then waits for exactly one
// slot's loads while the next slot's are in flight).  The slot's x offset and
// width come from lanes of mx, mw.
template <typename T, typename V, int NT, bool VX>
__device__ __forceinline__ void mfma_load(const ApplyArgs<T, V>& g, int mx, int mw, int u, int64_t vrow, int r,
                                          int li, int lk, V (&a)[2][2], T (&x)[4][NT], int& c_u) {
    const int xo = lane_read(mx, u), w = lane_read(mw, u);
    const int c = w & 255;
    __builtin_assume(c >= 1);
    c_u = c;
    const V* ab = g.vals + vrow + int64_t(r) * (w >> 8) + (li < r ? li : r - 1) * c;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const int p0 = 8 * m + 2 * lk;
        load_n_stream<V, 2>(ab + (p0 < c ? p0 : c - 1), a[m]);
    }
    const int K = g.K;
    const int c0 = NT * li < K - NT ? NT * li : K - NT;  // VX: lanes past K read the last NT columns
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int b = mfma_col(s, lk);
        const T* xr = g.X + int64_t(xo + (b < c ? b : c - 1)) * g.ldx;
        if constexpr (VX) {
            load_n<T, NT>(xr + c0, x[s]);
        } else {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int col = NT * li + t;
                x[s][t] = xr[col < K ? col : K - 1];
            }
        }
    }
}

// The MFMAs of one slot: the k-steps its width c needs (a wave-uniform test).
template <typename T, typename V, int NT>
__device__ __forceinline__ void mfma_slot(int c, int r, int li, int lk, int K, const V (&a)[2][2],
                                          const T (&x)[4][NT], vbh_d4 (&acc)[NT]) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (8 * (s >> 1) + (s & 1) < c) {
            const bool kon = mfma_col(s, lk) < c;
            const double av = (li < r && kon) ? widen(a[s >> 1][s & 1]) : 0.0;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const double bv = (kon && NT * li + t < K) ? widen(x[s][t]) : 0.0;
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
            }
        }
    }
}

// Epilogue of one matrix-core block row: Y rows row0..row0+r-1 from the
// MFMA accumulators (register e of lane l holds row 4e + l/16 of column
// NT*(l%16) + t of tile t), fused with beta*X_own + gamma*Z and the dots.
template <typename T, typename V, int NT, bool VX, bool DOT>
__device__ __forceinline__ void mfma_epilogue(const ApplyArgs<T, V>& g, const vbh_d4 (&acc)[NT], int r, int64_t row0,
                                              int64_t x0r, int li, int lk, double& d1, double& d2) {
    const int K = g.K;
    // All epilogue loads (clamped rows and columns) are
    // issued before the first use, so the row costs one more latency,
    // not one per element
    const bool need_x = DOT || g.beta != 0.0, need_z = g.gamma != 0.0, need_w = DOT && g.W != nullptr;
    const int c0 = NT * li < K - NT ? NT * li : K - NT;
    T xo[4][NT], zo[4][NT], wo[4][NT];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int a = 4 * e + lk;
        const int ac = a < r ? a : r - 1;
        const T* xp = g.X + (x0r + ac) * g.ldx;
        const T* zp = g.Z + (row0 + ac) * g.ldz;
        const T* wp = g.W + (row0 + ac) * g.ldw;
        if constexpr (VX) {
            if (need_x) load_n<T, NT>(xp + c0, xo[e]);
            if (need_z) load_n<T, NT>(zp + c0, zo[e]);
            if (need_w) load_n<T, NT>(wp + c0, wo[e]);
        } else {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int col = NT * li + t;
                const int cc = col < K ? col : K - 1;
                if (need_x) xo[e][t] = xp[cc];
                if (need_z) zo[e][t] = zp[cc];
                if (need_w) wo[e][t] = wp[cc];
            }
        }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int a = 4 * e + lk;
        T ys[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bool on = a < r && NT * li + t < K;
            const double xv = need_x ? widen(xo[e][t]) : 0.0;
            double y = g.alpha * acc[t][e];
            if (g.beta != 0.0) y += g.beta * xv;
            if (need_z) y += g.gamma * widen(zo[e][t]);
            ys[t] = narrow_to<T>(y);
            if constexpr (DOT) {
                const double yw = widen(ys[t]);
                d1 += on ? (need_w ? widen(wo[e][t]) : yw) * yw : 0.0;
                d2 += on ? xv * yw : 0.0;
            }
        }
        T* yp = g.Y + (row0 + a) * g.ldy;
        if constexpr (VX) {
            if (a < r && NT * li < K) store_n<T, NT>(yp + NT * li, ys);
        } else {
#pragma unroll
            for (int t = 0; t < NT; ++t)
                if (a < r && NT * li + t < K) yp[NT * li + t] = ys[t];
        }
    }}

// Block rows [q0, q1) of the matrix-core list, wave-interleaved (row q0 +
// wave, then + kWaves, ...), one block row at a time per wavefront.  Within
// a row two slots are in flight: while slot u feeds the MFMAs, the operands
// of slots u+1 and u+2 are being loaded (past the chunk's last slot the loads
// repeat it; those products are never issued).  Across rows, the next row's
// descriptor is loaded when a row starts and its first metadata chunk before
// the epilogue, so a row begins with its operand loads.
template <typename T, typename V, int NT, bool VX, bool DOT>
__device__ __forceinline__ void mfma_rows(const ApplyArgs<T, V>& g, int64_t q0, int64_t q1, int lane, int wave,
                                          double& d1, double& d2) {
    const DevPattern& p = g.p;
    const int li = lane & 15, lk = lane >> 4;
    const int K = g.K;
    // the row state is wave-uniform: readfirstlane keeps the slot loop a
    // scalar loop (the compiler cannot see that the wave index is uniform)
    int64_t q = q0 + __builtin_amdgcn_readfirstlane(wave);
    if (q >= q1) return;
    // metadata of the first chunk of a row: one slot per lane (clamped; an
    // empty row reads a neighbouring valid slot and never uses it)
    auto chunk_meta = [&](int kb, int ke, int& mx, int& mw) {
        const int nb = ke - kb < 64 ? ke - kb : 64;
        const int idx = nb > 0 ? kb + (lane < nb ? lane : nb - 1) : (kb > 0 ? kb - 1 : 0);
        const int2 mm = p.meta[idx];
        mx = mm.x;
        mw = mm.y;
    };
    int4 h0 = p.big_desc[2 * q], h1 = p.big_desc[2 * q + 1];
    int mx = 0, mw = 0;
    chunk_meta(__builtin_amdgcn_readfirstlane(h0.x), __builtin_amdgcn_readfirstlane(h0.y), mx, mw);
    for (;;) {
        const int kb = __builtin_amdgcn_readfirstlane(h0.x), ke = __builtin_amdgcn_readfirstlane(h0.y);
        const int r = __builtin_amdgcn_readfirstlane(h0.z);
        const int64_t row0 = __builtin_amdgcn_readfirstlane(h0.w), x0r = __builtin_amdgcn_readfirstlane(h1.x);
        const int64_t vrow = int64_t((uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(h1.z))) << 32) |
                                     uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(h1.y))));
        // the next row's descriptor (this row's again at the end of the range)
        const int64_t qn = q + kWaves;
        const int64_t qc = qn < q1 ? qn : q;
        const int4 n0 = p.big_desc[2 * qc], n1 = p.big_desc[2 * qc + 1];
        vbh_d4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = vbh_d4{0.0, 0.0, 0.0, 0.0};
        for (int t0 = kb; t0 < ke; t0 += 64) {
            const int nb = ke - t0 < 64 ? ke - t0 : 64;
            if (t0 > kb) chunk_meta(t0, ke, mx, mw);  // rows of more than 64 slots
            V a0[2][2], a1[2][2];
            T x0[4][NT], x1[4][NT];
            int c0 = 0, c1 = 0;
            mfma_load<T, V, NT, VX>(g, mx, mw, 0, vrow, r, li, lk, a0, x0, c0);
            mfma_load<T, V, NT, VX>(g, mx, mw, nb > 1 ? 1 : 0, vrow, r, li, lk, a1, x1, c1);
            for (int u = 0; u < nb; u += 2) {
                mfma_slot<T, V, NT>(c0, r, li, lk, K, a0, x0, acc);
                mfma_load<T, V, NT, VX>(g, mx, mw, u + 2 < nb ? u + 2 : nb - 1, vrow, r, li, lk, a0, x0, c0);
                if (u + 1 < nb) mfma_slot<T, V, NT>(c1, r, li, lk, K, a1, x1, acc);
                mfma_load<T, V, NT, VX>(g, mx, mw, u + 3 < nb ? u + 3 : nb - 1, vrow, r, li, lk, a1, x1, c1);
            }
        }
        // the next row's first metadata chunk, in flight during the epilogue
        chunk_meta(__builtin_amdgcn_readfirstlane(n0.x), __builtin_amdgcn_readfirstlane(n0.y), mx, mw);
        mfma_epilogue<T, V, NT, VX, DOT>(g, acc, r, row0, x0r, li, lk, d1, d2);
        if (qn >= q1) break;
        q = qn;
        h0 = n0;
        h1 = n1;
    }
}

template <typename T, typename V, int NT, bool VX, bool DOT>
__global__ __launch_bounds__(kThreads) void k_apply_mfma(ApplyArgs<T, V> g) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int chunk = xcd_chunk(g.G);
    const int64_t q0 = int64_t(chunk) * kMfmaRowsPerGroup;
    const int64_t q1 = q0 + kMfmaRowsPerGroup < g.p.n_big ? q0 + kMfmaRowsPerGroup : g.p.n_big;
    double d1 = 0.0, d2 = 0.0;
    mfma_rows<T, V, NT, VX, DOT>(g, q0, q1, lane, wave, d1, d2);
    if constexpr (DOT) write_partials(1, lane, wave, d1, d2, g.partials + int64_t(chunk) * 2);
}

// The matrix-core class of every matrix of a batch in one launch: work-group
// gg takes kMfmaRowsPerGroup entries of its matrix's big_brows.
template <typename T, typename V, int NT, bool VX>
__global__ __launch_bounds__(kThreads) void k_batched_mfma(const BatchItem<T, V>* __restrict__ items, int nitems,
                                                           int Gtot, int K, int ld, int step, double alpha,
                                                           double beta, double gamma0, double gamma, int mode,
                                                           double* __restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int gg = xcd_chunk(Gtot);
    const int lo = batch_item<T, V, true>(items, nitems, gg);
    using CItem = const __attribute__((address_space(4))) BatchItem<T, V>;
    CItem& it = ((CItem*)items)[lo];
    ApplyArgs<T, V> a = batch_args(items, lo, K, ld, step, alpha, beta, gamma0, gamma, mode);
    const int64_t q0 = int64_t(gg - it.g_base2) * kMfmaRowsPerGroup;
    const int64_t q1 = q0 + kMfmaRowsPerGroup < a.p.n_big ? q0 + kMfmaRowsPerGroup : a.p.n_big;
    double d1 = 0.0, d2 = 0.0;
    if (it.row_list) mfma_rows<T, V, NT, VX, true>(a, q0, q1, lane, wave, d1, d2);
    write_partials(1, lane, wave, d1, d2, partials + int64_t(gg) * 2);
}

// ------------------------------------------------------------ host dispatch
template <typename T>
constexpr int vec_max() {
    return sizeof(T) >= 16 ? 1 : int(16 / sizeof(T));
}

// Calls f(std::integral_constant<int, S>, std::integral_constant<int, KT>) for
// the instantiated lane shapes.
template <typename T, typename F>
static void with_shape(LaneShape sh, F&& f) {
    using std::integral_constant;
    const int code = sh.S * 100 + sh.KT;
    if constexpr (Sc<T>::is_complex) {
        switch (code) {
            case 104: f(integral_constant<int, 1>{}, integral_constant<int, 4>{}); return;
            case 108: f(integral_constant<int, 1>{}, integral_constant<int, 8>{}); return;
            case 208: f(integral_constant<int, 2>{}, integral_constant<int, 8>{}); return;
            case 408: f(integral_constant<int, 4>{}, integral_constant<int, 8>{}); return;
            case 808: f(integral_constant<int, 8>{}, integral_constant<int, 8>{}); return;
        }
    } else {
        switch (code) {
            case 104: f(integral_constant<int, 1>{}, integral_constant<int, 4>{}); return;
            case 108: f(integral_constant<int, 1>{}, integral_constant<int, 8>{}); return;
            case 110: f(integral_constant<int, 1>{}, integral_constant<int, 10>{}); return;
            case 112: f(integral_constant<int, 1>{}, integral_constant<int, 12>{}); return;
            case 116: f(integral_constant<int, 1>{}, integral_constant<int, 16>{}); return;
            case 216: f(integral_constant<int, 2>{}, integral_constant<int, 16>{}); return;
            case 416: f(integral_constant<int, 4>{}, integral_constant<int, 16>{}); return;
        }
    }
    throw Error(VBH_ERR_ARG, "internal: no lane kernel for this shape");
}

// Calls f(std::integral_constant<int, VN>) for vector width vn (which divides KT).
template <typename T, int KT, typename F>
static void with_vector(int vn, F&& f) {
    using std::integral_constant;
    if constexpr (vec_max<T>() >= 4 && KT % 4 == 0) {
        if (vn == 4) return f(integral_constant<int, 4>{});
    }
    if constexpr (vec_max<T>() >= 2 && KT % 2 == 0) {
        if (vn == 2) return f(integral_constant<int, 2>{});
    }
    f(integral_constant<int, 1>{});
}

template <typename T, typename V>
void launch_lanes(const ApplyArgs<T, V>& a, int64_t n_rows, LaneShape sh, int bs, hipStream_t s) {
    const int vn = lane_vector<T>(sh.KT, a.K, {{a.X, a.ldx}, {a.Y, a.ldy}, {a.gamma != 0.0 ? a.Z : nullptr, a.ldz},
                                               {a.partials ? a.W : nullptr, a.ldw}});
    const bool dot = a.partials != nullptr;
    const dim3 grid(a.G), block(kThreads);
    with_shape<T>(sh, [&](auto S_, auto KT_) {
        with_vector<T, decltype(KT_)::value>(vn, [&](auto VN_) {
            constexpr int S = decltype(S_)::value, KT = decltype(KT_)::value, VN = decltype(VN_)::value;
            if (bs == 4) {
                // FP64 quads of launches with one partition per row (large
                // matrices, HBM-resident): 4-slot batches.  Measured on the
                // 800 x 800 graphene KPM step: 720 us against 906 us with the
                // 2-slot batches that the 200 x 200 step (two partitions)
                // prefers (3.9 against 5.2 ms per call with 4-slot batches).
                if constexpr (std::is_same_v<T, double> && std::is_same_v<V, double>) {
                    if (a.parts == 1) {
                        if (dot) hipLaunchKernelGGL((k_apply_lanes<T, V, 4, S, KT, VN, true, 4>), grid, block, 0, s, a, n_rows);
                        else hipLaunchKernelGGL((k_apply_lanes<T, V, 4, S, KT, VN, false, 4>), grid, block, 0, s, a, n_rows);
                        return;
                    }
                }
                if (dot) hipLaunchKernelGGL((k_apply_lanes<T, V, 4, S, KT, VN, true>), grid, block, 0, s, a, n_rows);
                else hipLaunchKernelGGL((k_apply_lanes<T, V, 4, S, KT, VN, false>), grid, block, 0, s, a, n_rows);
            } else {
                if (dot) hipLaunchKernelGGL((k_apply_lanes<T, V, 0, S, KT, VN, true>), grid, block, 0, s, a, n_rows);
                else hipLaunchKernelGGL((k_apply_lanes<T, V, 0, S, KT, VN, false>), grid, block, 0, s, a, n_rows);
            }
        });
    });
    VBH_CHECK_HIP(hipGetLastError());
}

template <typename T, typename V>
void launch_batched_lanes(const BatchItem<T, V>* items, int nitems, int Gtot, int K, int ld, LaneShape sh, int step,
                          double alpha, double beta, double gamma0, double gamma, int mode, double* partials,
                          hipStream_t s) {
    // every recursion buffer starts on a 256-byte boundary (kpm.hip)
    const int vn = lane_vector<T>(sh.KT, K, {{reinterpret_cast<const void*>(256), ld}});
    with_shape<T>(sh, [&](auto S_, auto KT_) {
        with_vector<T, decltype(KT_)::value>(vn, [&](auto VN_) {
            constexpr int S = decltype(S_)::value, KT = decltype(KT_)::value, VN = decltype(VN_)::value;
            hipLaunchKernelGGL((k_batched_lanes<T, V, S, KT, VN>), dim3(Gtot), dim3(kThreads), 0, s, items, nitems, Gtot,
                               K, ld, step, alpha, beta, gamma0, gamma, mode, partials, sh.P);
        });
    });
    VBH_CHECK_HIP(hipGetLastError());
}

template <typename T, typename V>
void launch_batched_rowwave(const BatchItem<T, V>* items, int nitems, int Gtot, int K, int ld, int rw, int step,
                            double alpha, double beta, double gamma0, double gamma, int mode, double* partials,
                            hipStream_t s) {
    if constexpr (Sc<T>::is_complex) {
        throw Error(VBH_ERR_NOTSUP, "internal: the row-wave kernel is real-only");
    } else {
        constexpr int CPL = int(16 / sizeof(T));
        VBH_REQUIRE(rowwave_cpl<T>(K) == CPL && ld % CPL == 0, "internal: row-wave kernel shape");
        hipLaunchKernelGGL((k_batched_rowwave<T, V, CPL>), dim3(Gtot), dim3(kThreads), 0, s, items, nitems, Gtot, K,
                           ld, rw, step, alpha, beta, gamma0, gamma, mode, partials);
        VBH_CHECK_HIP(hipGetLastError());
    }
}

template <typename T, typename V>
void launch_batched_brw(const BatchItem<T, V>* items, int nitems, int Gtot, int K, int ld, int step, double alpha,
                        double beta, double gamma0, double gamma, int mode, double* partials, hipStream_t s) {
    if constexpr (Sc<T>::is_complex || !std::is_same_v<V, float>) {
        throw Error(VBH_ERR_NOTSUP, "internal: the block-row wave kernel takes real vectors and FP32 values");
    } else {
        VBH_REQUIRE(brw_cpl<T>(K) > 0 && ld % brw_cpl<T>(K) == 0, "internal: block-row wave kernel shape");
        auto go = [&](auto CPL_) {
            constexpr int CPL = decltype(CPL_)::value;
            hipLaunchKernelGGL((k_batched_brw<T, CPL>), dim3(Gtot), dim3(kThreads), 0, s, items, nitems, Gtot, K, ld,
                               step, alpha, beta, gamma0, gamma, mode, partials);
        };
        switch (K / 16) {
            case 1: go(std::integral_constant<int, 1>{}); break;
            case 2: go(std::integral_constant<int, 2>{}); break;
            default: go(std::integral_constant<int, 4>{}); break;
        }
        VBH_CHECK_HIP(hipGetLastError());
    }
}

template <typename T, typename V>
void launch_batched_mfma(const BatchItem<T, V>* items, int nitems, int Gtot, int K, int ld, int step, double alpha,
                         double beta, double gamma0, double gamma, int mode, double* partials, hipStream_t s) {
    if constexpr (Sc<T>::is_complex) {
        throw Error(VBH_ERR_NOTSUP, "internal: matrix-core class is real-only");
    } else {
        auto go = [&](auto NT_) {
            constexpr int NT = decltype(NT_)::value;
            if (K % NT == 0)
                hipLaunchKernelGGL((k_batched_mfma<T, V, NT, true>), dim3(Gtot), dim3(kThreads), 0, s, items, nitems,
                                   Gtot, K, ld, step, alpha, beta, gamma0, gamma, mode, partials);
            else
                hipLaunchKernelGGL((k_batched_mfma<T, V, NT, false>), dim3(Gtot), dim3(kThreads), 0, s, items, nitems,
                                   Gtot, K, ld, step, alpha, beta, gamma0, gamma, mode, partials);
        };
        switch ((K + 15) / 16) {
            case 1: go(std::integral_constant<int, 1>{}); break;
            case 2: go(std::integral_constant<int, 2>{}); break;
            case 3: go(std::integral_constant<int, 3>{}); break;
            default: go(std::integral_constant<int, 4>{}); break;
        }
        VBH_CHECK_HIP(hipGetLastError());
    }
}

template <typename T, typename V>
void launch_mfma(const ApplyArgs<T, V>& a, hipStream_t s) {
    if constexpr (Sc<T>::is_complex) {
        throw Error(VBH_ERR_NOTSUP, "internal: matrix-core class is real-only");
    } else {
        const int nt = (a.K + 15) / 16;
        const bool dot = a.partials != nullptr;
        const dim3 grid(a.G), block(kThreads);
        auto go = [&](auto NT_) {
            constexpr int NT = decltype(NT_)::value;
            const bool vx = a.K % NT == 0;
            if (dot && vx) hipLaunchKernelGGL((k_apply_mfma<T, V, NT, true, true>), grid, block, 0, s, a);
            else if (dot) hipLaunchKernelGGL((k_apply_mfma<T, V, NT, false, true>), grid, block, 0, s, a);
            else if (vx) hipLaunchKernelGGL((k_apply_mfma<T, V, NT, true, false>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((k_apply_mfma<T, V, NT, false, false>), grid, block, 0, s, a);
        };
        switch (nt) {
            case 1: go(std::integral_constant<int, 1>{}); break;
            case 2: go(std::integral_constant<int, 2>{}); break;
            case 3: go(std::integral_constant<int, 3>{}); break;
            default: go(std::integral_constant<int, 4>{}); break;
        }
        VBH_CHECK_HIP(hipGetLastError());
    }
}

#define VBH_LANES_INSTANTIATE(T, V)                                                                             \
    template void launch_mfma<T, V>(const ApplyArgs<T, V>&, hipStream_t);                                       \
    template void launch_batched_mfma<T, V>(const BatchItem<T, V>*, int, int, int, int, int, double, double,    \
                                            double, double, int, double*, hipStream_t);                        \
    template void launch_batched_rowwave<T, V>(const BatchItem<T, V>*, int, int, int, int, int, int, double,   \
                                               double, double, double, int, double*, hipStream_t);             \
    template void launch_batched_brw<T, V>(const BatchItem<T, V>*, int, int, int, int, int, double, double,     \
                                           double, double, int, double*, hipStream_t);                         \
    template void launch_lanes<T, V>(const ApplyArgs<T, V>&, int64_t, LaneShape, int, hipStream_t);            \
    template void launch_batched_lanes<T, V>(const BatchItem<T, V>*, int, int, int, int, LaneShape, int, double, \
                                             double, double, double, int, double*, hipStream_t);

}  // namespace vbh
