// Lane-per-row SpMM family for 2 <= K <= 64 (included by the per-type
// lanes_*.hip translation units).
//
// A wavefront owns 64/S consecutive scalar rows; lane l works on row
// base + l/S and on the column slice [(l%S)*KT, (l%S)*KT + KT) of the K
// right-hand sides.  Per block the lane reads its row of the block once (one
// 16-byte load for 4x4 FP32 blocks) and, for every block column b, the KT
// contiguous elements of X row (col + b) with vector loads, then issues KT
// FMAs per loaded A element: all K columns share one A load and every lane
// does useful work, where the column-per-lane map of row_loop spends most
// of its issue slots on index and shuffle work at small K.  Lanes of one
// block row read the same X rows (one cache line, broadcast).  Rows are
// independent, so block rows of any size tile the wave; lanes of short
// block rows simply finish their slot loop early.
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

// Slot loop of one lane for uniform 4x4 blocks.  The 4 rows of a block row
// sit in one quad of lanes; lane a of the quad loads X row (4*col + a) of
// every block (distinct addresses, no redundant traffic) and the quad shares
// the 4 rows by DPP broadcast.  Column indices of 4*CB slots are preloaded
// lane-parallel (lane a holds slots a, a+4, ...), and U slots are issued per
// batch, so a row of 10 blocks costs 1 + 10/U memory latencies; U shrinks
// with the slice width to keep 5+ waves per SIMD resident (measured: U = 2
// beats 1 and 4 for 10 FP32 columns).
template <typename T, typename V, int KT, int VN>
__device__ __forceinline__ void quad_slots(const ApplyArgs<T, V>& g, int kb, int ke, int a, int v0, int nvalid,
                                           typename Wide<T>::type (&acc)[KT]) {
    using A = typename Wide<T>::type;
    constexpr int CB = 4;                                   // slots per lane per pass
    constexpr int per = int(KT * sizeof(T) / 4);            // VGPRs of one X slice
#ifdef VBH_QUAD_BATCH
    constexpr int U = VBH_QUAD_BATCH;
#else
    constexpr int U = per <= 8 ? 4 : per <= 24 ? 2 : 1;     // slots per batch
#endif
    const DevPattern& p = g.p;
    for (int k0 = kb; k0 < ke; k0 += 4 * CB) {
        // ns >= 1; slot indices past the row are clamped to its last slot, so
        // every load below is unconditional and in bounds, and the extra
        // slots' products are dropped by a select
        const int ns = ke - k0 < 4 * CB ? ke - k0 : 4 * CB;
        int colv[CB];
#pragma unroll
        for (int t = 0; t < CB; ++t) colv[t] = p.col_ind[k0 + (a + 4 * t < ns ? a + 4 * t : ns - 1)];
#pragma unroll
        for (int s0 = 0; s0 < 4 * CB; s0 += U) {
            if (s0 >= ns) break;
            V av[U][4];
            T xr[U][KT];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int s = s0 + u;
                int col = 0;
                switch (s & 3) {
                    case 0: col = quad_bcast<0>(colv[s >> 2]); break;
                    case 1: col = quad_bcast<1>(colv[s >> 2]); break;
                    case 2: col = quad_bcast<2>(colv[s >> 2]); break;
                    default: col = quad_bcast<3>(colv[s >> 2]); break;
                }
                const int sc = s < ns ? s : ns - 1;
                const V* ar = static_cast<const V*>(
                    __builtin_assume_aligned(g.vals + int64_t(k0 + sc) * 16 + a * 4, (4 * sizeof(V)) % 16 ? sizeof(V) : 16));
#pragma unroll
                for (int b = 0; b < 4; ++b) av[u][b] = ar[b];
                load_slice<T, KT, VN>(g.X + (int64_t(col) * 4 + a) * g.ldx + v0, nvalid, xr[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool on = s0 + u < ns;
#pragma unroll
                for (int j = 0; j < KT; ++j) {
                    const T x0 = quad_bcast<0>(xr[u][j]), x1 = quad_bcast<1>(xr[u][j]);
                    const T x2 = quad_bcast<2>(xr[u][j]), x3 = quad_bcast<3>(xr[u][j]);
                    A t = Sc<A>::fma(widen(av[u][0]), widen(x0), acc[j]);
                    t = Sc<A>::fma(widen(av[u][1]), widen(x1), t);
                    t = Sc<A>::fma(widen(av[u][2]), widen(x2), t);
                    t = Sc<A>::fma(widen(av[u][3]), widen(x3), t);
                    acc[j] = on ? t : acc[j];
                }
            }
        }
    }
}

// Rows [sb, se) of the matrix, wave-interleaved in segments of R = 64/S
// rows; lane l works on row l % R of the segment and column slice l / R, so
// the rows of a block row occupy consecutive lanes.  BS > 0: uniform BS x BS
// blocks (BS == 4: quad-shared X rows, quad_slots).
template <typename T, typename V, int BS, int S, int KT, int VN, bool DOT>
__device__ __forceinline__ void lane_rows(const ApplyArgs<T, V>& g, int64_t sb, int64_t se, int lane, int wave,
                                          double& d1, double& d2) {
    using A = typename Wide<T>::type;
    constexpr int R = 64 / S;
    const DevPattern& p = g.p;
    const int v0 = (lane / R) * KT;
    const int nvalid = g.K - v0;
    for (int64_t base = sb + int64_t(wave) * R; base < se; base += int64_t(kWaves) * R) {
        const int64_t srow = base + lane % R;
        const bool live = srow < se && nvalid > 0;
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
            quad_slots<T, V, KT, VN>(g, kb, ke, a, v0, nvalid, acc);
        } else if constexpr (BS > 0) {
            for (int k = kb; k < ke; ++k) {
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
            int2 mn = p.meta[kb < ke ? kb : 0];
            for (int k = kb; k < ke; ++k) {
                const int2 mm = mn;
                mn = p.meta[k + 1 < ke ? k + 1 : k];
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
                            acc[j] = on ? t : acc[j];
                        }
                    }
                }
            }
        }
        if (live) emit_slice<T, V, KT, VN, DOT>(g, srow, xown, v0, nvalid, acc, d1, d2);
    }
}

template <typename T, typename V, int BS, int S, int KT, int VN, bool DOT>
__global__ __launch_bounds__(kThreads)
void k_apply_lanes(ApplyArgs<T, V> g, int64_t n_rows) {
    constexpr int RPG = kWaves * (64 / S);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int chunk = xcd_chunk(g.G);
    const int64_t sb = int64_t(chunk) * RPG;
    const int64_t se = sb + RPG < n_rows ? sb + RPG : n_rows;
    double d1 = 0.0, d2 = 0.0;
    lane_rows<T, V, BS, S, KT, VN, DOT>(g, sb, se, lane, wave, d1, d2);
    if constexpr (DOT) write_partials(1, lane, wave, d1, d2, g.partials + int64_t(chunk) * 2);
}

// One KPM recursion step over every matrix of a batch (see kpm.hip): work-group
// gg finds its matrix by binary search over g_base; groups past the last
// matrix's rows write zero partials.
template <typename T, typename V, int S, int KT, int VN>
__global__ __launch_bounds__(kThreads)
void k_batched_lanes(const BatchItem<T, V>* __restrict__ items, int nitems,
                                                            int Gtot, int K, int ld, int step, double alpha,
                                                            double beta, double gamma0, double gamma, int mode,
                                                            double* __restrict__ partials) {
    constexpr int RPG = kWaves * (64 / S);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int gg = xcd_chunk(Gtot);
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].g_base <= gg) lo = mid;
        else hi = mid - 1;
    }
    lo = uniform(lo);  // the same for the whole work-group: scalar loads below
    // The item table is read through the constant address space: it is
    // read-only during the launch, and the compiler then takes the pointers
    // loaded from it to be global (global_load rather than flat_load, whose
    // completion also holds the LDS/scalar wait counter).
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
    const int64_t sb = int64_t(gg - it.g_base) * RPG;
    const int64_t se = sb + RPG < it.n_rows ? sb + RPG : it.n_rows;
    double d1 = 0.0, d2 = 0.0;
    lane_rows<T, V, 0, S, KT, VN, true>(a, sb, se, lane, wave, d1, d2);
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
                               K, ld, step, alpha, beta, gamma0, gamma, mode, partials);
        });
    });
    VBH_CHECK_HIP(hipGetLastError());
}

#define VBH_LANES_INSTANTIATE(T, V)                                                                             \
    template void launch_lanes<T, V>(const ApplyArgs<T, V>&, int64_t, LaneShape, int, hipStream_t);            \
    template void launch_batched_lanes<T, V>(const BatchItem<T, V>*, int, int, int, int, LaneShape, int, double, \
                                             double, double, double, int, double*, hipStream_t);

}  // namespace vbh
