// Device-side helpers shared by the HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

#include "nlhip_types.h"

namespace nlhip {

constexpr int kWave = 64;
constexpr int kBlock = 256;           // threads per workgroup for all kernels
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kNumXcd = 8;       // MI355X: 8 XCDs, each with its own L2

// Sum over each aligned group of W lanes (W a power of two <= 64).  A xor
// butterfly: every lane of the group ends with the same, fixed-order sum.
template <int W, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

// Block-wide sum; the result is returned to every thread.  `lds` must hold
// kWavesPerBlock elements.  Deterministic: waves summed in index order.
template <typename T>
__device__ __forceinline__ T block_sum_all(T v, T* lds) {
    v = group_sum<kWave>(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    T s = lds[0];
#pragma unroll
    for (int w = 1; w < kWavesPerBlock; ++w) s += lds[w];
    return s;
}

// Fixed-order reduction of P per-block partial sums written by an earlier
// kernel; every block that calls it obtains the bitwise-same value.
template <typename T>
__device__ __forceinline__ T reduce_partials(const T* __restrict__ part, uint32_t P,
                                             T* lds) {
    T s = T(0);
    for (uint32_t i = threadIdx.x; i < P; i += kBlock) s += part[i];
    return block_sum_all(s, lds);
}

// Workgroups are dispatched round-robin over the XCDs (block b runs on XCD
// b % 8).  Remap so that XCD x owns one contiguous range of logical blocks:
// rows processed on one XCD are then contiguous and their x-gathers share
// that XCD's L2.  A bijection on [0, G).
__device__ __forceinline__ uint32_t xcd_logical_block(uint32_t b, uint32_t G) {
    const uint32_t x = b % kNumXcd, k = b / kNumXcd;
    uint32_t start = 0;
    for (uint32_t y = 0; y < x; ++y) start += (G - y + kNumXcd - 1) / kNumXcd;
    return start + k;
}

template <typename T>
__device__ __forceinline__ T ntload(const T* p) {
    return __builtin_nontemporal_load(p);
}

// Separately rounded multiply and add, so element-wise updates round exactly
// like a scalar x = x + (a * y) loop.  HIP's __dmul_rn / __dadd_rn are plain
// operators, so this holds because device code is compiled with
// -ffp-contract=off (no multiply-add fusion anywhere).
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float add_rn(float a, float b) { return __fadd_rn(a, b); }

// ---- CRS row helpers -------------------------------------------------

template <int LPR, typename T>
__device__ __forceinline__ T row_product(uint32_t row, const uint32_t* __restrict__ rp,
                                         const uint32_t* __restrict__ ci,
                                         const T* __restrict__ val,
                                         const T* __restrict__ x, int lane) {
    const uint32_t b = rp[row], e = rp[row + 1];
    T acc = T(0);
    for (uint32_t j = b + lane; j < e; j += LPR) {
        acc += ntload(val + j) * x[ntload(ci + j)];
    }
    return group_sum<LPR>(acc);
}

__device__ __forceinline__ void block_row_range(uint32_t m, uint32_t chunk, uint32_t L,
                                                uint32_t& r0, uint32_t& r1) {
    const uint64_t a = uint64_t(L) * chunk;
    r0 = a < m ? uint32_t(a) : m;
    const uint64_t e = uint64_t(r0) + chunk;
    r1 = e < m ? uint32_t(e) : m;
}

// ---- SpMV epilogues ----------------------------------------------------
//
// The engines hand every row sum to an epilogue, called by exactly one lane
// per row: epi.done(row, sum, pv), where pv = epi.pre(row) was fetched when
// the row's tile started, so that the latency of the per-row operands the
// epilogue reads (b, D^-1, d, ...) overlaps the row's nonzeros.

struct NoPre {
    __device__ __forceinline__ int operator()(uint32_t) const { return 0; }
};

template <typename Pre, typename Fin>
struct Epilogue {
    Pre pre;
    Fin fin;
    using PV = decltype(std::declval<Pre&>()(0u));
    template <typename T>
    __device__ __forceinline__ void done(uint32_t row, T sum, const PV& pv) {
        if constexpr (std::is_same<Pre, NoPre>::value) fin(row, sum);
        else fin(row, sum, pv);
    }
    static constexpr bool kPairs = false;  // see DqEpilogue
};

// fin(row, sum)
template <typename Fin>
__device__ __forceinline__ Epilogue<NoPre, Fin> plain_epi(Fin fin) {
    return Epilogue<NoPre, Fin>{NoPre{}, fin};
}
// pre(row) -> pv; fin(row, sum, pv)
template <typename Pre, typename Fin>
__device__ __forceinline__ Epilogue<Pre, Fin> prefetch_epi(Pre pre, Fin fin) {
    return Epilogue<Pre, Fin>{pre, fin};
}

// The CG SpMV step's epilogue: q[row] = sum, and sum * d[row] added to the
// lane's partial of d.q (left to right, row by row).  With the 2x2 block
// engine a lane's two rows form an aligned pair (2I, 2I + 1) on a whole
// matrix: d is then read and q written as one 2-vector each (one contiguous
// 1 KiB word run per wavefront instead of every other 8-byte word).
template <typename T>
struct DqEpilogue {
    const T* __restrict__ d;
    T* __restrict__ q;
    T& pdq;
    using PV = T;
    using V2 = T __attribute__((ext_vector_type(2)));
    static constexpr bool kPairs = true;
    __device__ __forceinline__ T pre(uint32_t row) const { return d[row]; }
    __device__ __forceinline__ void done(uint32_t row, T sum, T dr) {
        q[row] = sum;
        pdq += dr * sum;
    }
    __device__ __forceinline__ void pre2(uint32_t row, T& d0, T& d1) const {
        const V2 v = *reinterpret_cast<const V2*>(d + row);
        d0 = v.x;
        d1 = v.y;
    }
    __device__ __forceinline__ void done2(uint32_t row, T s0, T s1, T d0, T d1) {
        V2 v;
        v.x = s0;
        v.y = s1;
        *reinterpret_cast<V2*>(q + row) = v;
        pdq += d0 * s0;
        pdq += d1 * s1;
    }
};

template <typename T>
struct Pair {
    T a, b;
};

// ---- SpMV engines -----------------------------------------------------
//
// Both engines walk the rows owned by one workgroup and hand every row sum
// to an epilogue (above).
//
// Vector engine (RPW < 64 rows per wavefront): LPR = 64/RPW lanes per row,
// rows [L*sp, L*sp + sp) of the matrix for logical block L, partial products
// combined by a fixed xor butterfly.
//
// Stream engine (RPW = 64): one row per lane.  A wavefront owns a run of
// 64-row tiles, whose nonzeros form one contiguous range; it streams that
// range in chunks of 64*E entries, each lane loading E independent (val,
// colind) pairs and gathering x, stages the products in LDS and then lane l
// sums row l's products in column order.  Block L owns tiles [L*sp, L*sp+sp),
// split into four contiguous runs, one per wavefront.  Each row sum is
// sum_j fl(a_ij x_j) accumulated left to right with one rounding per add.
// With D16 the column indices are read as 16-bit segment offsets (Cols),
// 10 instead of 12 bytes per FP64 nonzero; the sums are the same.

template <int W>
struct VectorEngine {
    static constexpr int LPR = 64 / W;
    template <typename T, typename Epi>
    __device__ __forceinline__ static void run(uint32_t m, uint32_t sp, uint32_t L,
                                               const uint32_t* __restrict__ rp,
                                               const Cols& cols, const T* __restrict__ val,
                                               const T* __restrict__ x, T*, Epi& epi) {
        const uint32_t* __restrict__ ci = cols.ci;
        uint32_t r0, r1;
        block_row_range(m, sp, L, r0, r1);
        const int g = threadIdx.x / LPR, lane = threadIdx.x % LPR;
        for (uint32_t row = r0 + g; row < r1; row += kBlock / LPR) {
            const T s = row_product<LPR>(row, rp, ci, val, x, lane);
            if (lane == 0) epi.done(row, s, epi.pre(row));
        }
    }
};

// Logical blocks [lo, hi) that run on the same XCD as logical block L (the
// inverse view of xcd_logical_block).
__device__ __forceinline__ void xcd_block_range(uint32_t L, uint32_t G, uint32_t& lo,
                                                uint32_t& hi) {
    uint32_t start = 0;
    for (uint32_t y = 0; y < kNumXcd; ++y) {
        const uint32_t cnt = (G - y + kNumXcd - 1) / kNumXcd;
        if (L < start + cnt || y + 1 == kNumXcd) {
            lo = start;
            hi = start + cnt;
            return;
        }
        start += cnt;
    }
}

template <int E, bool NT, bool D16 = false, int RUN = 0>
struct StreamEngine {
    static_assert(64 * E <= int(kSegEntries), "a chunk must span at most two segments");

    template <typename T>
    __device__ __forceinline__ static T ld(const T* p) {
        if constexpr (NT) return ntload(p);
        else return *p;
    }

    // Loads the chunk of entries [base, base + 64E) (indices past J1 clamped
    // to J1 - 1, so every column read is a real one).  base and J1 are
    // wavefront-uniform.
    template <typename T>
    __device__ __forceinline__ static void load_chunk(const Cols& cols,
                                                      const T* __restrict__ val, uint32_t base,
                                                      uint32_t J1, int lane, uint32_t (&c)[E],
                                                      T (&a)[E]) {
        // column indices first: the gathers wait on them, not on the values
        if constexpr (D16) {
            // the clamped indices of this chunk lie in [q, q + 64E) with
            // q = min(base, J1 - 1), i.e. in segment s0 or s0 + 1
            const uint32_t q = base < J1 - 1u ? base : J1 - 1u;
            const uint32_t s0 = q >> kSegShift;
            const uint32_t b0 = cols.segb[s0], b1 = cols.segb[s0 + 1];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t j = base + uint32_t(e) * 64u + uint32_t(lane);
                const uint32_t jj = j < J1 ? j : J1 - 1u;
                c[e] = ((jj >> kSegShift) == s0 ? b0 : b1) + uint32_t(ld(cols.ci16 + jj));
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t j = base + uint32_t(e) * 64u + uint32_t(lane);
                c[e] = ld(cols.ci + (j < J1 ? j : J1 - 1u));
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t j = base + uint32_t(e) * 64u + uint32_t(lane);
            a[e] = ld(val + (j < J1 ? j : J1 - 1u));
        }
    }

    // Rows [R0, R1) of one wavefront: lane l owns row r0 + l of the current
    // 64-row tile; the wavefront's nonzeros [rp[R0], rp[R1]) are streamed as
    // one contiguous range in chunks of 64*E entries, the next chunk's loads
    // issued before the current chunk is summed.
    template <typename T, typename Epi>
    __device__ __forceinline__ static void run_rows(uint32_t R0, uint32_t R1,
                                                    const uint32_t* __restrict__ rp,
                                                    const Cols& cols, const T* __restrict__ val,
                                                    const T* __restrict__ x, T* __restrict__ st,
                                                    int lane, Epi& epi) {
        const uint32_t J1 = __builtin_amdgcn_readfirstlane(rp[R1]);
        uint32_t r0 = R0;
        uint32_t nrows = (R1 - r0) < 64u ? (R1 - r0) : 64u;
        uint32_t tile_end = rp[r0 + nrows];
        uint32_t rb = (uint32_t(lane) < nrows) ? rp[r0 + uint32_t(lane)] : tile_end;
        typename Epi::PV pv{};
        if (uint32_t(lane) < nrows) pv = epi.pre(r0 + uint32_t(lane));
        uint32_t re = __shfl_down(rb, 1, kWave);
        if (uint32_t(lane) + 1 >= nrows) re = tile_end;
        uint32_t base = __builtin_amdgcn_readfirstlane(__shfl(rb, 0, kWave));
        T acc = T(0);
        uint32_t c[E];
        T a[E];
        if (base < J1) load_chunk(cols, val, base, J1, lane, c, a);
        while (base < J1) {
            const uint32_t craw = base + 64u * E;
            const uint32_t cend = craw < J1 ? craw : J1;
            T p[E];
#pragma unroll
            for (int e = 0; e < E; ++e) p[e] = x[c[e]];
            // keep every gather ahead of the prefetch in the instruction stream:
            // loads return in order, so a gather issued late would make its
            // wait drain the prefetch too
            __builtin_amdgcn_sched_barrier(0);
            // next chunk, issued unconditionally (clamped to the last entry)
            // so that its loads overlap this chunk's products and row sums
            uint32_t cn[E];
            T an[E];
            load_chunk(cols, val, craw, J1, lane, cn, an);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t j = base + uint32_t(e) * 64u + uint32_t(lane);
                st[e * 64 + lane] = (j < J1) ? mul_rn(a[e], p[e]) : T(0);
            }
            __builtin_amdgcn_wave_barrier();
            // consume this chunk; it may complete several tiles
            for (;;) {
                const uint32_t lo = rb > base ? rb : base;
                const uint32_t hi = re < cend ? re : cend;
                uint32_t j = lo;
                for (; j + 4 <= hi; j += 4) {
                    const T v0 = st[j - base], v1 = st[j + 1 - base];
                    const T v2 = st[j + 2 - base], v3 = st[j + 3 - base];
                    acc = add_rn(add_rn(add_rn(add_rn(acc, v0), v1), v2), v3);
                }
                for (; j < hi; ++j) acc = add_rn(acc, st[j - base]);
                if (tile_end > cend) break;
                if (uint32_t(lane) < nrows) epi.done(r0 + uint32_t(lane), acc, pv);
                acc = T(0);
                r0 += 64u;
                if (r0 >= R1) return;
                nrows = (R1 - r0) < 64u ? (R1 - r0) : 64u;
                tile_end = rp[r0 + nrows];
                rb = (uint32_t(lane) < nrows) ? rp[r0 + uint32_t(lane)] : tile_end;
                re = __shfl_down(rb, 1, kWave);
                if (uint32_t(lane) + 1 >= nrows) re = tile_end;
                if (uint32_t(lane) < nrows) pv = epi.pre(r0 + uint32_t(lane));
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
            if (uint32_t(lane) < nrows) epi.done(r0 + uint32_t(lane), acc, pv);
            acc = T(0);
            r0 += 64u;
            if (r0 >= R1) return;
            nrows = (R1 - r0) < 64u ? (R1 - r0) : 64u;
            if (uint32_t(lane) < nrows) pv = epi.pre(r0 + uint32_t(lane));
        }
    }

    // Block L owns 64-row tiles [L*sp, L*sp + sp).
    // RUN = 0: each of its wavefronts takes one contiguous quarter of them.
    // RUN > 0: the tiles of all blocks on L's XCD are dealt out in runs of RUN
    // tiles, round-robin over that XCD's wavefronts, so that the wavefronts
    // of an XCD sweep its rows together and the x entries they gather (a
    // band around the rows in flight) stay in that XCD's L2.
    template <typename T, typename Epi>
    __device__ __forceinline__ static void run(uint32_t m, uint32_t sp, uint32_t L,
                                               const uint32_t* __restrict__ rp,
                                               const Cols& cols, const T* __restrict__ val,
                                               const T* __restrict__ x, T* stage, Epi& epi) {
        const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
        const uint64_t ntiles = (uint64_t(m) + 63) / 64;
        if constexpr (RUN > 0) {
            uint32_t lo, hi;
            xcd_block_range(L, gridDim.x, lo, hi);
            const uint64_t tb = uint64_t(lo) * sp < ntiles ? uint64_t(lo) * sp : ntiles;
            const uint64_t te = uint64_t(hi) * sp < ntiles ? uint64_t(hi) * sp : ntiles;
            const uint64_t nw = uint64_t(hi - lo) * kWavesPerBlock;
            const uint64_t v = uint64_t(L - lo) * kWavesPerBlock + uint64_t(w);
            for (uint64_t t0 = tb + v * RUN; t0 < te; t0 += nw * RUN) {
                const uint64_t t1 = t0 + RUN < te ? t0 + RUN : te;
                const uint32_t R1 = uint32_t(t1 * 64 < m ? t1 * 64 : m);
                run_rows(uint32_t(t0 * 64), R1, rp, cols, val, x, stage + w * (64 * E), lane,
                         epi);
            }
            return;
        }
        uint64_t tb = uint64_t(L) * sp;
        if (tb > ntiles) tb = ntiles;
        const uint64_t te = tb + sp < ntiles ? tb + sp : ntiles;
        const uint64_t q = (te - tb + kWavesPerBlock - 1) / kWavesPerBlock;
        uint64_t t0 = tb + uint64_t(w) * q;
        if (t0 > te) t0 = te;
        const uint64_t t1 = t0 + q < te ? t0 + q : te;
        if (t0 >= t1) return;
        const uint32_t R0 = uint32_t(t0 * 64);
        const uint32_t R1 = uint32_t(t1 * 64 < m ? t1 * 64 : m);
        run_rows(R0, R1, rp, cols, val, x, stage + w * (64 * E), lane, epi);
    }
};

// ---- 2x2 block stream engine ------------------------------------------
//
// For matrices in 2x2 block form: rows 2I and 2I+1 have the same columns,
// which come in pairs (2J, 2J+1) (LSCM normal equations: the u and v
// unknowns of a vertex).  Block row I stores its blocks in order, each as
// an upper pair (a(2I,2J), a(2I,2J+1)) and a lower pair (a(2I+1,2J),
// a(2I+1,2J+1)) in the layout of block_upper / block_lower (grouped by 64
// blocks in FP64, block by block in FP32), with
// the block column J as a 16-bit segment offset (Cols::ci16 / Cols::segb
// over the block stream; 8.5 bytes per FP64 nonzero) or, with D16 = false,
// as a 32-bit index in Cols::ci.  One block row per lane, 64 block rows per
// wavefront tile, chunks of 64*EB blocks; the four products of every block
// are staged in LDS and each lane sums its two rows left to right
// (a(2I,2J) x_2J, then a(2I,2J+1) x_2J+1, for J ascending) -- the order and
// roundings of the CRS row loop.  `rp` is the block row pointer (m/2 + 1
// entries), `val` the block values (block_vals_size), `m` the number of
// (scalar) rows.

template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec2<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};

template <int EB, bool D16 = true, int RUN = 0>
struct BlockStreamEngine {
    static_assert(64 * EB <= int(kSegEntries), "a chunk must span at most two segments");

    // block values of block jj (block_upper / block_lower): FP64 as two
    // aligned 2-vectors from the upper and lower pair groups, FP32 as one
    // aligned 4-vector
    template <typename T>
    __device__ __forceinline__ static void load_vals(const T* __restrict__ val, uint32_t jj,
                                                     T (&a)[4]) {
        if constexpr (sizeof(T) == 4) {
            using V4 = T __attribute__((ext_vector_type(4)));
            const V4 v =
                __builtin_nontemporal_load(reinterpret_cast<const V4*>(val + block_upper<T>(jj)));
            a[0] = v.x;
            a[1] = v.y;
            a[2] = v.z;
            a[3] = v.w;
            return;
        }
        using V2 = typename Vec2<T>::type;
        const V2 lo =
            __builtin_nontemporal_load(reinterpret_cast<const V2*>(val + block_upper<T>(jj)));
        const V2 hi =
            __builtin_nontemporal_load(reinterpret_cast<const V2*>(val + block_lower<T>(jj)));
        a[0] = lo.x;
        a[1] = lo.y;
        a[2] = hi.x;
        a[3] = hi.y;
    }

    template <typename T>
    __device__ __forceinline__ static void load_chunk(const Cols& cols,
                                                      const T* __restrict__ val, uint32_t base,
                                                      uint32_t J1, int lane, uint32_t (&c)[EB],
                                                      T (&a)[EB][4]) {
        if constexpr (D16) {
            const uint32_t q = base < J1 - 1u ? base : J1 - 1u;
            const uint32_t s0 = q >> kSegShift;
            const uint32_t b0 = cols.segb[s0], b1 = cols.segb[s0 + 1];
#pragma unroll
            for (int e = 0; e < EB; ++e) {
                const uint32_t j = base + uint32_t(e) * 64u + uint32_t(lane);
                const uint32_t jj = j < J1 ? j : J1 - 1u;
                c[e] = ((jj >> kSegShift) == s0 ? b0 : b1) + uint32_t(ntload(cols.ci16 + jj));
            }
        } else {  // 32-bit block columns in cols.ci
#pragma unroll
            for (int e = 0; e < EB; ++e) {
                const uint32_t j = base + uint32_t(e) * 64u + uint32_t(lane);
                c[e] = ntload(cols.ci + (j < J1 ? j : J1 - 1u));
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int e = 0; e < EB; ++e) {
            const uint32_t j = base + uint32_t(e) * 64u + uint32_t(lane);
            load_vals(val, j < J1 ? j : J1 - 1u, a[e]);
        }
    }

    // Block rows [R0, R1) of one wavefront (same structure as
    // StreamEngine::run_rows, one block row per lane).  Block row B holds
    // rows 2B + delta and 2B + 1 + delta, block column J columns 2J + delta
    // and 2J + 1 + delta (delta = 0 for a whole matrix; a batched system
    // may sit at any row offset).
    template <typename T, typename Epi>
    __device__ __forceinline__ static void run_rows(uint32_t R0, uint32_t R1,
                                                    const uint32_t* __restrict__ rp,
                                                    const Cols& cols, const T* __restrict__ val,
                                                    const T* __restrict__ x, T* __restrict__ st,
                                                    int lane, Epi& epi, int32_t delta = 0) {
        using V2 = typename Vec2<T>::type;
        const T* __restrict__ xs = x + delta;
        const bool aligned = (delta & 1) == 0;
        const uint32_t J1 = __builtin_amdgcn_readfirstlane(rp[R1]);
        uint32_t r0 = R0;
        uint32_t nrows = (R1 - r0) < 64u ? (R1 - r0) : 64u;
        uint32_t tile_end = rp[r0 + nrows];
        uint32_t rb = (uint32_t(lane) < nrows) ? rp[r0 + uint32_t(lane)] : tile_end;
        uint32_t re = __shfl_down(rb, 1, kWave);
        if (uint32_t(lane) + 1 >= nrows) re = tile_end;
        uint32_t base = __builtin_amdgcn_readfirstlane(__shfl(rb, 0, kWave));
        auto row_of = [&](uint32_t B) { return uint32_t(int64_t(2u * B) + delta); };
        typename Epi::PV pv0{}, pv1{};
        T acc0 = T(0), acc1 = T(0);
        // the epilogue operands of a lane's two rows (one 2-vector when the
        // epilogue takes pairs and the rows are an aligned pair)
        auto pre_pair = [&](uint32_t row) {
            if constexpr (Epi::kPairs) {
                if (aligned) {
                    epi.pre2(row, pv0, pv1);
                    return;
                }
            }
            pv0 = epi.pre(row);
            pv1 = epi.pre(row + 1u);
        };
        auto done_pair = [&](uint32_t row) {
            if constexpr (Epi::kPairs) {
                if (aligned) {
                    epi.done2(row, acc0, acc1, pv0, pv1);
                    return;
                }
            }
            epi.done(row, acc0, pv0);
            epi.done(row + 1u, acc1, pv1);
        };
        if (uint32_t(lane) < nrows) pre_pair(row_of(r0 + uint32_t(lane)));
        uint32_t c[EB];
        T a[EB][4];
        if (base < J1) load_chunk(cols, val, base, J1, lane, c, a);
        while (base < J1) {
            const uint32_t craw = base + 64u * EB;
            const uint32_t cend = craw < J1 ? craw : J1;
            V2 xp[EB];
            if (aligned) {
#pragma unroll
                for (int e = 0; e < EB; ++e)
                    xp[e] = *reinterpret_cast<const V2*>(xs + 2ull * c[e]);
            } else {
#pragma unroll
                for (int e = 0; e < EB; ++e) {
                    xp[e].x = xs[2ull * c[e]];
                    xp[e].y = xs[2ull * c[e] + 1];
                }
            }
            __builtin_amdgcn_sched_barrier(0);
            uint32_t cn[EB];
            T an[EB][4];
            load_chunk(cols, val, craw, J1, lane, cn, an);
#pragma unroll
            for (int e = 0; e < EB; ++e) {
                const uint32_t j = base + uint32_t(e) * 64u + uint32_t(lane);
                const bool ok = j < J1;
                T* o = st + 4 * (e * 64 + lane);
                o[0] = ok ? mul_rn(a[e][0], xp[e].x) : T(0);
                o[1] = ok ? mul_rn(a[e][1], xp[e].y) : T(0);
                o[2] = ok ? mul_rn(a[e][2], xp[e].x) : T(0);
                o[3] = ok ? mul_rn(a[e][3], xp[e].y) : T(0);
            }
            __builtin_amdgcn_wave_barrier();
            for (;;) {
                const uint32_t lo = rb > base ? rb : base;
                const uint32_t hi = re < cend ? re : cend;
                for (uint32_t j = lo; j < hi; ++j) {
                    const T* o = st + 4 * (j - base);
                    acc0 = add_rn(add_rn(acc0, o[0]), o[1]);
                    acc1 = add_rn(add_rn(acc1, o[2]), o[3]);
                }
                if (tile_end > cend) break;
                if (uint32_t(lane) < nrows) {
                    done_pair(row_of(r0 + uint32_t(lane)));
                }
                acc0 = acc1 = T(0);
                r0 += 64u;
                if (r0 >= R1) return;
                nrows = (R1 - r0) < 64u ? (R1 - r0) : 64u;
                tile_end = rp[r0 + nrows];
                rb = (uint32_t(lane) < nrows) ? rp[r0 + uint32_t(lane)] : tile_end;
                re = __shfl_down(rb, 1, kWave);
                if (uint32_t(lane) + 1 >= nrows) re = tile_end;
                if (uint32_t(lane) < nrows) {
                    pre_pair(row_of(r0 + uint32_t(lane)));
                }
            }
            __builtin_amdgcn_wave_barrier();
            base = craw;
#pragma unroll
            for (int e = 0; e < EB; ++e) {
                c[e] = cn[e];
#pragma unroll
                for (int k = 0; k < 4; ++k) a[e][k] = an[e][k];
            }
        }
        for (;;) {
            if (uint32_t(lane) < nrows) {
                done_pair(row_of(r0 + uint32_t(lane)));
            }
            acc0 = acc1 = T(0);
            r0 += 64u;
            if (r0 >= R1) return;
            nrows = (R1 - r0) < 64u ? (R1 - r0) : 64u;
            if (uint32_t(lane) < nrows) {
                pre_pair(row_of(r0 + uint32_t(lane)));
            }
        }
    }

    // Block L owns block-row tiles [L*sp, L*sp + sp) (tiles of 64 block
    // rows).  RUN = 0: each of its wavefronts takes a contiguous quarter of
    // them.  RUN > 0: the tiles of all blocks on L's XCD are dealt out in
    // runs of RUN tiles, round-robin over that XCD's wavefronts, so that at
    // any time the wavefronts of an XCD read and write one compact stretch
    // of the matrix and of y instead of one stretch each.
    template <typename T, typename Epi>
    __device__ __forceinline__ static void run(uint32_t m, uint32_t sp, uint32_t L,
                                               const uint32_t* __restrict__ rp,
                                               const Cols& cols, const T* __restrict__ val,
                                               const T* __restrict__ x, T* stage, Epi& epi) {
        const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
        const uint32_t mb = m / 2;
        const uint64_t ntiles = (uint64_t(mb) + 63) / 64;
        if constexpr (RUN > 0) {
            uint32_t lo, hi;
            xcd_block_range(L, gridDim.x, lo, hi);
            const uint64_t tb = uint64_t(lo) * sp < ntiles ? uint64_t(lo) * sp : ntiles;
            const uint64_t te = uint64_t(hi) * sp < ntiles ? uint64_t(hi) * sp : ntiles;
            const uint64_t nw = uint64_t(hi - lo) * kWavesPerBlock;
            const uint64_t v = uint64_t(L - lo) * kWavesPerBlock + uint64_t(w);
            for (uint64_t t0 = tb + v * RUN; t0 < te; t0 += nw * RUN) {
                const uint64_t t1 = t0 + RUN < te ? t0 + RUN : te;
                const uint32_t R1 = uint32_t(t1 * 64 < mb ? t1 * 64 : mb);
                run_rows(uint32_t(t0 * 64), R1, rp, cols, val, x, stage + w * (64 * EB * 4),
                         lane, epi);
            }
            return;
        }
        uint64_t tb = uint64_t(L) * sp;
        if (tb > ntiles) tb = ntiles;
        const uint64_t te = tb + sp < ntiles ? tb + sp : ntiles;
        const uint64_t q = (te - tb + kWavesPerBlock - 1) / kWavesPerBlock;
        uint64_t t0 = tb + uint64_t(w) * q;
        if (t0 > te) t0 = te;
        const uint64_t t1 = t0 + q < te ? t0 + q : te;
        if (t0 >= t1) return;
        const uint32_t R0 = uint32_t(t0 * 64);
        const uint32_t R1 = uint32_t(t1 * 64 < mb ? t1 * 64 : mb);
        run_rows(R0, R1, rp, cols, val, x, stage + w * (64 * EB * 4), lane, epi);
    }
};

constexpr int kBlockEB = 4;  // blocks per lane per chunk (16 nonzeros)

// Engine for a rows-per-wavefront code, and its LDS staging size.  Codes
// 64 and above are stream engines; 64 itself is resolved by the launchers
// (engine_rpw) to kStreamAuto or kStreamAutoPlain.
constexpr int kStreamE = 8;
constexpr int kStreamRun = 4;

template <int RPW>
struct EngineOf {
    using type = VectorEngine<RPW>;
};
// Stream-engine variants for launch-shape sweeps (chunk width, load kind,
// column format, tile order); all compute the same row sums.
template <>
struct EngineOf<65> {
    using type = StreamEngine<4, true>;
};
template <>
struct EngineOf<66> {
    using type = StreamEngine<8, false>;
};
template <>
struct EngineOf<67> {
    using type = StreamEngine<16, true>;
};
template <>
struct EngineOf<68> {  // 16-bit column offsets, contiguous quarters
    using type = StreamEngine<kStreamE, true, true>;
};
template <>
struct EngineOf<69> {  // 32-bit indices, XCD-interleaved runs
    using type = StreamEngine<kStreamE, true, false, kStreamRun>;
};
template <>
struct EngineOf<70> {  // 16-bit column offsets, XCD-interleaved runs
    using type = StreamEngine<kStreamE, true, true, kStreamRun>;
};
template <>
struct EngineOf<71> {  // 32-bit indices, contiguous quarters
    using type = StreamEngine<kStreamE, true>;
};
template <>
struct EngineOf<72> {  // 16-bit column offsets, 16 entries per lane per chunk
    using type = StreamEngine<16, true, true>;
};
template <>
struct EngineOf<73> {  // 2x2 blocks (rp, cols, val are the block arrays)
    using type = BlockStreamEngine<kBlockEB>;
};
template <>
struct EngineOf<74> {  // 2x2 blocks, 2 blocks per lane per chunk
    using type = BlockStreamEngine<2>;
};
template <>
struct EngineOf<75> {  // 2x2 blocks, 1 block per lane per chunk
    using type = BlockStreamEngine<1>;
};
template <>
struct EngineOf<76> {  // 2x2 blocks, 8 blocks per lane per chunk
    using type = BlockStreamEngine<8>;
};
template <>
struct EngineOf<77> {  // 2x2 blocks, XCD-interleaved runs of 4 tiles
    using type = BlockStreamEngine<kBlockEB, true, 4>;
};
template <>
struct EngineOf<78> {  // 2x2 blocks, 32-bit block columns, 4 per lane per chunk
    using type = BlockStreamEngine<kBlockEB, false>;
};
template <>
struct EngineOf<79> {  // 2x2 blocks, 32-bit block columns, 8 per lane per chunk
    using type = BlockStreamEngine<8, false>;
};

template <int RPW>
struct StageSize {
    static constexpr int value = (RPW == 64 || (RPW >= 68 && RPW <= 71))
                                     ? kWavesPerBlock * 64 * kStreamE
                                 : RPW == 65 ? kWavesPerBlock * 64 * 4
                                 : RPW == 66 ? kWavesPerBlock * 64 * 8
                                 : (RPW == 67 || RPW == 72) ? kWavesPerBlock * 64 * 16
                                 : (RPW == 73 || RPW == 77 || RPW == 78)
                                     ? kWavesPerBlock * 64 * kBlockEB * 4
                                 : RPW == 74 ? kWavesPerBlock * 64 * 2 * 4
                                 : RPW == 75 ? kWavesPerBlock * 64 * 1 * 4
                                 : (RPW == 76 || RPW == 79) ? kWavesPerBlock * 64 * 8 * 4
                                             : 1;
};

}  // namespace nlhip
