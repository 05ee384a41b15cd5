// Multi-vector SpMM row engine shared by the multi-right-hand-side PCG
// (k_spmm.hip) and the eigensolver's Chebyshev filter (k_eigen.hip).
#pragma once

#include "dev_common.h"

namespace nlhip {

constexpr uint32_t kSpmmChunk = 1024;  // (colind, val) entries staged in LDS at a time
constexpr uint32_t kSpmmPass = 64;     // rows per pass (at least)
constexpr int kSpmmU = 8;              // gathers in flight per lane

// Rows of one pass of the engine for G lanes per row.
template <int G>
constexpr uint32_t spmm_pass_rows() {
    return uint32_t(kBlock / G) > kSpmmPass ? uint32_t(kBlock / G) : kSpmmPass;
}

// Rows of logical block L (the passes of max(64, kBlock / G) consecutive rows
// of its XCD's range that fall to it in turn, see below): for each pass the
// block copies the pass's (colind, val) range into LDS with coalesced
// loads, then each G-lane group sums its rows of the pass from LDS, lane j
// gathering X[col * k + j].  Row sums are left to right with one rounding
// per product and per add.  For each row, lane j calls pre(row, j) (loads
// the epilogue needs that do not depend on the sum) and then
// fin(row, j, sum, pre value), row by row in the same order whatever the
// path.  Every thread of the block must call this (it synchronises the
// block).  zero_x: X is all zeros; the rows' sums are +0 (what the sums
// would give) and the matrix is not read.
//
// ONE: every pass's entries fit in one LDS chunk (the caller checked it,
// spmm_one_chunk): a group finishes each row before the next, holding
// one accumulator and issuing the row's epilogue loads with its gathers.
// Otherwise passes are staged in chunks and the group keeps one accumulator
// per row of the pass until the pass ends.
// V columns per lane (V = 1 or 2): lane j of a group owns columns
// j*V .. j*V + V - 1, gathers them as one copy per entry (16 bytes for two
// FP64 columns; k * V-aligned when V = 2 and k is even) and sums each of
// them as above; pre / fin are called per column, columns in order.  G
// lanes per row cover G * V >= k columns.
template <int G, int V, bool ONE, typename T, typename Pre, typename Fin>
__device__ __forceinline__ void spmm_rows_v(uint32_t m, uint32_t k, uint32_t rows_per_block,
                                            uint32_t L, const uint32_t* __restrict__ rp,
                                            const uint32_t* __restrict__ ci,
                                            const T* __restrict__ val, const T* __restrict__ X,
                                            Pre& pre, Fin& fin, const int32_t* done = nullptr,
                                            bool zero_x = false) {
    static_assert(V == 1 || V == 2, "one or two columns per lane");
    __shared__ uint32_t s_ci[kSpmmChunk];
    __shared__ T s_val[kSpmmChunk];
    constexpr uint32_t groups = kBlock / G;
    constexpr uint32_t pass = spmm_pass_rows<G>();
    const uint32_t g = threadIdx.x / G, jb = (threadIdx.x % G) * V;
    const bool mine = jb < k;  // the lane owns at least one column
    // columns that have stopped gather nothing (their sums are not used)
    bool active_col = false;
#pragma unroll
    for (int v = 0; v < V; ++v)
        active_col |= jb + v < k && !(done && done[jb + v]);
    // the blocks that run on one XCD take the passes of their joint row range
    // in turn, so at any time they work on neighbouring rows and the X rows
    // those gather (a stencil's rows i - w .. i + w) are shared in that
    // XCD's L2 instead of fetched once per block
    uint32_t lo, hi;
    xcd_block_range(L, gridDim.x, lo, hi);
    const uint64_t xb0 = uint64_t(lo) * rows_per_block, xb1 = uint64_t(hi) * rows_per_block;
    const uint32_t r0 = xb0 < m ? uint32_t(xb0) : m;
    const uint32_t r1 = xb1 < m ? uint32_t(xb1) : m;
    const uint32_t first = uint32_t(uint64_t(L - lo) * pass);
    const uint32_t step = (hi - lo) * pass;
    constexpr uint32_t R = pass / groups;  // rows per group in one pass
    // U gathers of entries [p, hi) (LDS offsets), added in order to acc
    auto gather = [&](uint32_t p, uint32_t hi, T (&acc)[V]) {
        T av[kSpmmU], xv[kSpmmU][V];
#pragma unroll
        for (int u = 0; u < kSpmmU; ++u) {
            const uint32_t q = p + u < hi ? p + u : hi - 1u;
            av[u] = s_val[q];
            if (active_col) {
                __builtin_memcpy(xv[u], X + uint64_t(s_ci[q]) * k + jb, sizeof(T) * V);
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) xv[u][v] = T(0);
            }
        }
#pragma unroll
        for (int u = 0; u < kSpmmU; ++u)
            if (p + u < hi) {
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = add_rn(acc[v], mul_rn(av[u], xv[u][v]));
            }
    };
    using PV = decltype(pre(0u, 0u));
    auto finish = [&](uint32_t row, const T (&acc)[V], const PV (&pv)[V]) {
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (jb + v < k) fin(row, jb + v, acc[v], pv[v]);
    };
    auto fetch = [&](uint32_t row, PV (&pv)[V]) {
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (jb + v < k) pv[v] = pre(row, jb + v);
    };
    for (uint64_t a64 = uint64_t(r0) + first; a64 < r1; a64 += step) {
        const uint32_t a = uint32_t(a64);
        const uint32_t pe = (a + pass < r1) ? a + pass : r1;
        const uint32_t J0 = rp[a], J1 = rp[pe];
        if (zero_x) {  // X = 0: every sum is +0 (block-uniform)
            for (uint32_t t = 0; t < R; ++t) {
                const uint32_t row = a + t * groups + g;
                if (row >= pe) break;
                if (!mine) continue;
                PV pv[V]{};
                fetch(row, pv);
                T z[V];
#pragma unroll
                for (int v = 0; v < V; ++v) z[v] = T(0);
                finish(row, z, pv);
            }
            continue;
        }
        if constexpr (ONE) {
            __syncthreads();  // the previous pass has been consumed
            for (uint32_t q = J0 + threadIdx.x; q < J1; q += kBlock) {
                s_ci[q - J0] = ci[q];
                s_val[q - J0] = val[q];
            }
            __syncthreads();
            for (uint32_t t = 0; t < R; ++t) {
                const uint32_t row = a + t * groups + g;
                if (row >= pe) break;
                const uint32_t lo = rp[row] - J0, hi = rp[row + 1] - J0;
                PV pv[V]{};
                if (mine) fetch(row, pv);
                T acc[V];
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = T(0);
                for (uint32_t p = lo; p < hi; p += kSpmmU) gather(p, hi, acc);
                if (mine) finish(row, acc, pv);
            }
        } else {
            uint32_t rb[R], re[R];
            T acc[R][V];
#pragma unroll
            for (uint32_t t = 0; t < R; ++t) {
                const uint32_t row = a + t * groups + g;
                rb[t] = re[t] = 0;
                if (row < pe) {
                    rb[t] = rp[row];
                    re[t] = rp[row + 1];
                }
#pragma unroll
                for (int v = 0; v < V; ++v) acc[t][v] = T(0);
            }
            for (uint32_t c0 = J0; c0 < J1; c0 += kSpmmChunk) {
                const uint32_t c1 = (c0 + kSpmmChunk < J1) ? c0 + kSpmmChunk : J1;
                __syncthreads();  // the previous chunk has been consumed
                for (uint32_t q = c0 + threadIdx.x; q < c1; q += kBlock) {
                    s_ci[q - c0] = ci[q];
                    s_val[q - c0] = val[q];
                }
                __syncthreads();
#pragma unroll
                for (uint32_t t = 0; t < R; ++t) {
                    const uint32_t lo = rb[t] > c0 ? rb[t] : c0, hi = re[t] < c1 ? re[t] : c1;
                    for (uint32_t p = lo; p < hi; p += kSpmmU) gather(p - c0, hi - c0, acc[t]);
                }
            }
#pragma unroll
            for (uint32_t t = 0; t < R; ++t) {
                const uint32_t row = a + t * groups + g;
                if (mine && row < pe) {
                    PV pv[V]{};
                    fetch(row, pv);
                    finish(row, acc[t], pv);
                }
            }
        }
    }
}

// One column per lane (spmm_rows_v with V = 1): lane j of a group owns
// column j.
template <int G, bool ONE, typename T, typename Pre, typename Fin>
__device__ __forceinline__ void spmm_rows(uint32_t m, uint32_t k, uint32_t rows_per_block,
                                          uint32_t L, const uint32_t* __restrict__ rp,
                                          const uint32_t* __restrict__ ci,
                                          const T* __restrict__ val, const T* __restrict__ X,
                                          Pre& pre, Fin& fin, const int32_t* done = nullptr,
                                          bool zero_x = false) {
    spmm_rows_v<G, 1, ONE>(m, k, rows_per_block, L, rp, ci, val, X, pre, fin, done, zero_x);
}

// ---- stream SpMM for a few right-hand sides --------------------------
//
// For k <= 4 right-hand sides the row-group engine above leaves lanes idle
// (4 lanes for k = 3) and waits for each pass's staging before it gathers.
// This engine streams the matrix as the scalar stream engine does (one row
// per lane, 64-row tiles, a wavefront's nonzeros as one contiguous range in
// chunks of 64E entries, the next chunk's (colind, val) loads issued before
// the current chunk is summed) and gathers the K values X[col*K + j] of every
// entry.  The K products of each entry are staged in LDS and lane l sums
// row l's products column by column, left to right with one rounding per
// product and per add: the sums of spmm_rows, bitwise.  Columns with
// act[j] false sum zeros (their sums are not used).  Block L owns tiles [L*sp, L*sp + sp), one contiguous quarter
// per wavefront; `stage` holds kStage values per wavefront.
//
// Epilogue: pv = epi.pre(row) when the row's tile starts, then
// epi.done(row, sums[K], pv), by the row's lane, rows in order.  zero_x: X
// is all zeros; the sums are +0 (what the sums would give) and the matrix
// is not read.
template <int K, int E, bool D16>
struct MultiStreamEngine {
    using Loads = StreamEngine<E, true, D16>;
    static constexpr uint32_t kStage = 64u * uint32_t(E) * uint32_t(K);

    template <typename T, typename Epi>
    __device__ __forceinline__ static void run_rows(uint32_t R0, uint32_t R1,
                                                    const uint32_t* __restrict__ rp,
                                                    const Cols& cols, const T* __restrict__ val,
                                                    const T* __restrict__ X, const bool (&act)[K],
                                                    T* __restrict__ st, int lane, Epi& epi) {
        constexpr uint32_t C = 64u * E;  // entries per chunk
        const uint32_t ul = uint32_t(lane);
        const uint32_t J1 = __builtin_amdgcn_readfirstlane(rp[R1]);
        uint32_t r0 = R0;
        uint32_t nrows = (R1 - r0) < 64u ? (R1 - r0) : 64u;
        uint32_t tile_end = rp[r0 + nrows];
        uint32_t rb = ul < nrows ? rp[r0 + ul] : tile_end;
        typename Epi::PV pv{};
        if (ul < nrows) pv = epi.pre(r0 + ul);
        uint32_t re = __shfl_down(rb, 1, kWave);
        if (ul + 1 >= nrows) re = tile_end;
        uint32_t base = __builtin_amdgcn_readfirstlane(__shfl(rb, 0, kWave));
        T acc[K];
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] = T(0);
        uint32_t c[E];
        T a[E];
        if (base < J1) Loads::load_chunk(cols, val, base, J1, lane, c, a);
        while (base < J1) {
            const uint32_t craw = base + C;
            const uint32_t cend = craw < J1 ? craw : J1;
            // each entry's K values as one copy: the widest loads the
            // alignment allows (k = 3 FP64: 16 + 8 bytes; FP32: 12 bytes),
            // so the gather instructions per entry are 1-2 instead of K
            T p[E][K];
#pragma unroll
            for (int e = 0; e < E; ++e)
                __builtin_memcpy(p[e], X + uint64_t(c[e]) * K, sizeof(T) * K);
            // every gather ahead of the prefetch (loads return in order)
            __builtin_amdgcn_sched_barrier(0);
            uint32_t cn[E];
            T an[E];
            Loads::load_chunk(cols, val, craw, J1, lane, cn, an);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const bool in = base + uint32_t(e) * 64u + ul < J1;
#pragma unroll
                for (int j = 0; j < K; ++j)
                    st[uint32_t(j) * C + uint32_t(e) * 64u + ul] =
                        (in && act[j]) ? mul_rn(a[e], p[e][j]) : T(0);
            }
            __builtin_amdgcn_wave_barrier();
            // consume this chunk; it may complete several tiles
            for (;;) {
                const uint32_t lo = rb > base ? rb : base;
                const uint32_t hi = re < cend ? re : cend;
                for (uint32_t q = lo; q < hi; ++q) {
#pragma unroll
                    for (int j = 0; j < K; ++j) acc[j] = add_rn(acc[j], st[uint32_t(j) * C + q - base]);
                }
                if (tile_end > cend) break;
                if (ul < nrows) epi.done(r0 + ul, acc, pv);
#pragma unroll
                for (int j = 0; j < K; ++j) acc[j] = T(0);
                r0 += 64u;
                if (r0 >= R1) return;
                nrows = (R1 - r0) < 64u ? (R1 - r0) : 64u;
                tile_end = rp[r0 + nrows];
                rb = ul < nrows ? rp[r0 + ul] : tile_end;
                re = __shfl_down(rb, 1, kWave);
                if (ul + 1 >= nrows) re = tile_end;
                if (ul < nrows) pv = epi.pre(r0 + ul);
            }
            __builtin_amdgcn_wave_barrier();
            base = craw;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                c[e] = cn[e];
                a[e] = an[e];
            }
        }
        // tiles left without nonzeros (all their rows are empty)
        for (;;) {
            if (ul < nrows) epi.done(r0 + ul, acc, pv);
            r0 += 64u;
            if (r0 >= R1) return;
            nrows = (R1 - r0) < 64u ? (R1 - r0) : 64u;
            if (ul < nrows) pv = epi.pre(r0 + ul);
        }
    }

    template <typename T, typename Epi>
    __device__ __forceinline__ static void run(uint32_t m, uint32_t sp, uint32_t L,
                                               const uint32_t* __restrict__ rp,
                                               const Cols& cols, const T* __restrict__ val,
                                               const T* __restrict__ X, const bool (&act)[K],
                                               T* stage, Epi& epi, bool zero_x = false) {
        const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
        const uint64_t ntiles = (uint64_t(m) + 63) / 64;
        uint64_t tb = uint64_t(L) * sp;
        if (tb > ntiles) tb = ntiles;
        const uint64_t te = tb + sp < ntiles ? tb + sp : ntiles;
        const uint64_t q = (te - tb + kWavesPerBlock - 1) / kWavesPerBlock;
        uint64_t t0 = tb + uint64_t(w) * q;
        if (t0 > te) t0 = te;
        const uint64_t t1 = t0 + q < te ? t0 + q : te;
        if (t0 >= t1) return;
        const uint32_t R1 = uint32_t(t1 * 64 < m ? t1 * 64 : m);
        if (zero_x) {  // X = 0: the same epilogue calls, in order, with sums of +0
            T z[K];
#pragma unroll
            for (int j = 0; j < K; ++j) z[j] = T(0);
            for (uint32_t r0 = uint32_t(t0 * 64); r0 < R1; r0 += 64u) {
                const uint32_t row = r0 + uint32_t(lane);
                if (row < R1) {
                    const typename Epi::PV pv = epi.pre(row);
                    epi.done(row, z, pv);
                }
            }
            return;
        }
        run_rows(uint32_t(t0 * 64), R1, rp, cols, val, X, act, stage + w * kStage, lane, epi);
    }
};

}  // namespace nlhip
