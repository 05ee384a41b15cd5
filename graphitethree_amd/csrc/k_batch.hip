// Batched Jacobi-PCG over many independent systems (the per-chart solves of
// a texture atlas, src/lib/OGF/mesh/commands/mesh_grob_surface_commands.cpp:
// 724-730), run as one fixed grid of workgroups per launch.
//
// Layout: the systems are packed as one block-diagonal CRS (global row and
// column indices), rows of system s in [row_off[s], row_off[s+1]).  Each
// system is cut into tiles of kTileRows rows that never straddle systems.
// Workgroups walk the device-side list of active tiles (a contiguous share
// each in the iteration kernels, grid stride elsewhere), so one launch
// covers every unconverged system whatever their sizes.
//
// Per iteration and per system the arithmetic is the CG of k_cg.hip, in the
// same three launches (direction, SpMV, update; see below).  System scalars
// are reduced over that system's tile partials in a fixed order, so a
// system's result depends only on the system itself: it is bitwise the same
// solved alone or in any batch.
#include <hip/hip_cooperative_groups.h>

#include "dev_common.h"
#include "nlhip_kernels.h"

namespace nlhip {

// Rows [r0, r1) of tile t: kTileRows rows, or kBlockTileRows for systems in
// 2x2 block form.  The tiling of a system depends only on that system.
template <typename T>
__device__ __forceinline__ void tile_rows(const BatchArgs<T>& a, uint32_t t, uint32_t& r0,
                                          uint32_t& r1) {
    const uint32_t s = a.tile_sys[t];
    r0 = a.tile_row0[t];
    const uint32_t e = a.row_off[s + 1];
    const uint32_t tr = (a.sys_pair0 && a.sys_pair0[s] != kNoBlocks) ? kBlockTileRows : kTileRows;
    r1 = (r0 + tr < e) ? r0 + tr : e;
}

// Blocks per lane per chunk of the batched block engine: 4 in the kernel
// built for block-form systems only, 2 in the kernel that also holds the
// scalar engine (whose registers would otherwise cut occupancy to 3
// wavefronts per SIMD); and the LDS staging per workgroup (elements),
// enough for any of them.
constexpr int kBatchEB = 2;
constexpr int kBatchEBBlock = 4;
constexpr int kBatchStageE = kStreamE > kBatchEBBlock * 4 ? kStreamE : kBatchEBBlock * 4;
constexpr int kBatchStage = kWavesPerBlock * 64 * kBatchStageE;

template <typename T>
__device__ __forceinline__ bool tile_blk(const BatchArgs<T>& a, uint32_t t) {
    return a.sys_pair0 && a.sys_pair0[a.tile_sys[t]] != kNoBlocks;
}

// Stream SpMV of passes [first, first + R) of a tile with rows [r0, r1) of
// system s (first row lo, first block row p0, or kNoBlocks for the scalar
// form), by one wavefront: pass i covers kTilePassRows rows from
// r0 + i kTilePassRows or, in 2x2 block form, kTilePassRows block rows
// (row pairs) from row r0 + 2 i kTilePassRows (r0 - lo and r1 - lo are even
// for such a system).  Returns the first row of the run.
//
// F (kFormAny, kFormBlock16, kFormBlock32) says which forms the kernel is
// built for: any, or only the 2x2 block form with 16-bit or 32-bit block
// columns (a batch whose systems all have that form; such a kernel holds
// one engine and needs fewer registers, so more wavefronts fit on a CU).
constexpr int kFormAny = 0, kFormBlock16 = 1, kFormBlock32 = 2;

template <typename T, int F = kFormAny, typename Epi>
__device__ __forceinline__ void wave_spmv_rows(const BatchArgs<T>& a, uint32_t r0, uint32_t r1,
                                               uint32_t lo, uint32_t p0, uint32_t first,
                                               uint32_t R, const T* x, T* stage, Epi& epi) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (F != kFormAny || p0 != kNoBlocks) {
        constexpr int EB = F == kFormAny ? kBatchEB : kBatchEBBlock;
        static_assert(EB * 4 * 64 * kWavesPerBlock <= kBatchStage, "staging area too small");
        const int32_t delta = int32_t(lo - 2u * p0);
        const uint32_t B0 = p0 + (r0 - lo) / 2 + kTilePassRows * first;
        const uint32_t Bend = p0 + (r1 - lo) / 2;
        if (B0 >= Bend) return;
        const uint32_t B1 = (B0 + kTilePassRows * R < Bend) ? B0 + kTilePassRows * R : Bend;
        const Cols bcols{a.bcol, a.bci16, a.bsegb};
        T* st = stage + w * (64 * EB * 4);
        if (F == kFormBlock16 || (F == kFormAny && a.bci16))
            BlockStreamEngine<EB, true>::run_rows(B0, B1, a.brp, bcols, a.bval, x, st, lane,
                                                  epi, delta);
        else
            BlockStreamEngine<EB, false>::run_rows(B0, B1, a.brp, bcols, a.bval, x, st, lane,
                                                   epi, delta);
        return;
    }
    if constexpr (F == kFormAny) {
        const uint32_t R0 = r0 + kTilePassRows * first;
        if (R0 >= r1) return;
        const uint32_t R1 = (R0 + kTilePassRows * R < r1) ? R0 + kTilePassRows * R : r1;
        const Cols cols{a.ci, a.ci16, a.segb};
        if (a.ci16)
            StreamEngine<kStreamE, true, true>::run_rows(R0, R1, a.rp, cols, a.val, x,
                                                         stage + w * (64 * kStreamE), lane, epi);
        else
            StreamEngine<kStreamE, true>::run_rows(R0, R1, a.rp, cols, a.val, x,
                                                   stage + w * (64 * kStreamE), lane, epi);
    }
}

// The whole tile by the workgroup: wavefront w takes pass w.
template <typename T, int F = kFormAny, typename Epi>
__device__ __forceinline__ void tile_spmv_rows(const BatchArgs<T>& a, uint32_t r0, uint32_t r1,
                                               uint32_t lo, uint32_t p0, const T* x, T* stage,
                                               Epi& epi) {
    wave_spmv_rows<T, F>(a, r0, r1, lo, p0, threadIdx.x / kWave, 1u, x, stage, epi);
}

template <typename T, typename Epi>
__device__ __forceinline__ void tile_spmv(const BatchArgs<T>& a, uint32_t t, uint32_t r0,
                                          uint32_t r1, bool blk, const T* x, T* stage,
                                          Epi& epi) {
    const uint32_t s = a.tile_sys[t];
    tile_spmv_rows(a, r0, r1, a.row_off[s], blk ? a.sys_pair0[s] : kNoBlocks, x, stage, epi);
}

// Metadata of one entry of the active-tile list.  In the iteration kernels
// lane j of every wavefront loads the entry k + j, so a workgroup fetches
// the metadata of 64 tiles in three dependent loads instead of three per
// tile, and reads one tile's with v_readlane.
struct TileDesc {
    uint32_t t;         // tile
    uint32_t s;         // its system
    uint32_t r0, r1;    // its rows
    uint32_t lo;        // first row of s
    uint32_t p0;        // first block row of s, or kNoBlocks
    uint32_t to0, to1;  // tiles of s
    uint32_t flags;     // 1: done[s], 2: stop[s] (when asked for)
};

template <typename T>
__device__ __forceinline__ TileDesc tile_desc_load(const BatchArgs<T>& a, uint32_t k,
                                                   uint32_t k1, bool with_stop) {
    TileDesc d{};
    if (k < k1) {
        d.t = a.act_tiles[k];
        d.s = a.tile_sys[d.t];
        d.r0 = a.tile_row0[d.t];
        d.lo = a.row_off[d.s];
        const uint32_t e = a.row_off[d.s + 1];
        d.p0 = a.sys_pair0 ? a.sys_pair0[d.s] : kNoBlocks;
        const uint32_t tr = d.p0 != kNoBlocks ? kBlockTileRows : kTileRows;
        d.r1 = (d.r0 + tr < e) ? d.r0 + tr : e;
        d.to0 = a.tile_off[d.s];
        d.to1 = a.tile_off[d.s + 1];
        d.flags = (a.done[d.s] ? 1u : 0u) | ((with_stop && a.stop[d.s]) ? 2u : 0u);
    }
    return d;
}

__device__ __forceinline__ uint32_t lane_value(uint32_t v, int j) {
    return uint32_t(__builtin_amdgcn_readlane(int(v), j));
}

__device__ __forceinline__ TileDesc tile_desc_at(const TileDesc& v, int j) {
    TileDesc d;
    d.t = lane_value(v.t, j);
    d.s = lane_value(v.s, j);
    d.r0 = lane_value(v.r0, j);
    d.r1 = lane_value(v.r1, j);
    d.lo = lane_value(v.lo, j);
    d.p0 = lane_value(v.p0, j);
    d.to0 = lane_value(v.to0, j);
    d.to1 = lane_value(v.to1, j);
    d.flags = lane_value(v.flags, j);
    return d;
}

// Rows of a tile per thread in the iteration kernels (loads of all of them
// are issued before the stores).
constexpr int kTileRowsPerThread = int(kBlockTileRows) / kBlock;
static_assert(kTileRows <= kBlockTileRows && kBlockTileRows % kBlock == 0, "tile shape");

template <typename T, bool PREC>
__global__ __launch_bounds__(kBlock) void k_b_init(BatchArgs<T> a) {
    __shared__ T stage[kBatchStage];
    __shared__ T lds[kWavesPerBlock];
    const uint32_t n = *a.n_act_tiles;
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const uint32_t t = a.act_tiles[k];
        uint32_t r0, r1;
        tile_rows(a, t, r0, r1);
        T prh = T(0), prr = T(0), pbb = T(0);
        auto epi = prefetch_epi(
            [&](uint32_t row) { return Pair<T>{a.b[row], PREC ? a.dinv[row] : T(1)}; },
            [&](uint32_t row, T s, const Pair<T>& pv) {
                const T bi = pv.a;
                const T ri = add_rn(s, mul_rn(T(-1), bi));
                const T hi = PREC ? mul_rn(pv.b, ri) : ri;
                a.r[row] = ri;
                a.h[row] = hi;
                prh += ri * hi;
                prr += ri * ri;
                pbb += bi * bi;
            });
        tile_spmv(a, t, r0, r1, tile_blk(a, t), a.x, stage, epi);
        prh = block_sum_all(prh, lds);
        prr = block_sum_all(prr, lds);
        pbb = block_sum_all(pbb, lds);
        if (threadIdx.x == 0) {
            a.part_rh[t] = prh;
            a.part_rr[t] = prr;
            a.part_bb[t] = pbb;
        }
    }
}

// One wavefront per system: sum of part[tile_off[s] .. tile_off[s+1]) in a
// fixed order (lane-strided, then the xor butterfly).
template <typename T>
__device__ __forceinline__ T tiles_sum(const T* part, uint32_t to0, uint32_t to1) {
    const int lane = threadIdx.x & (kWave - 1);
    T v = T(0);
    for (uint32_t t = to0 + uint32_t(lane); t < to1; t += kWave) v += part[t];
    return group_sum<kWave>(v);
}

template <typename T>
__device__ __forceinline__ T sys_sum(const BatchArgs<T>& a, const T* part, uint32_t s) {
    return tiles_sum(part, a.tile_off[s], a.tile_off[s + 1]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_sys_init(BatchArgs<T> a, T eps, int max_iter) {
    const uint32_t s = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (s >= a.nsys) return;
    const T rho = sys_sum(a, a.part_rh, s);
    const T err = sys_sum(a, a.part_rr, s);
    const T bb = sys_sum(a, a.part_bb, s);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const T thresh = (eps * eps) * bb;
        a.bb[s] = bb;
        a.thresh[s] = thresh;
        a.rho[s] = rho;
        a.rho_prev[s] = T(1);
        a.err[s] = err;
        a.beta[s] = T(0);
        a.alpha[s] = T(0);
        a.iters[s] = 0;
        a.stop[s] = 0;
        const bool skip = a.skip && a.skip[s];
        const int cap = a.cap && a.cap[s] < max_iter ? a.cap[s] : max_iter;
        a.done[s] = (skip || !(err > thresh) || cap <= 0) ? 1 : 0;
    }
}

// Three launches per iteration, as the single-system CG (k_cg.hip): the
// system scalars are reduced from the tile partials by every workgroup that
// needs them (sys_sum, a fixed order), and only the workgroup of a system's
// first tile writes them.
//
//   direction  rho, err from the r.h, r.r partials; x -= alpha d with the
//              previous alpha (the deferred update of x); stopping test
//              (raises stop[s]); beta = rho / rho_prev; d = beta d + h
//              (d = h at iteration 0, whose test ran in k_b_sys_init)
//   spmv       stop[s] -> done[s]; q = A d, tile partials of d.q
//   update     alpha = rho / sum(d.q); r -= alpha q; h = P r; partials of
//              r.h and r.r
// Each workgroup takes a contiguous share of the active-tile list.  The
// list keeps a system's tiles adjacent (k_b_compact appends them in order
// within each wavefront), so a workgroup meets a system once or twice per
// launch and reduces its scalars only then, instead of once per tile.
// Shares follow the XCD-contiguous logical block order, so the tiles of a
// system run on one or two XCDs and its SpMV gathers stay in their L2.
__device__ __forceinline__ void tile_share(uint32_t n, uint32_t& k0, uint32_t& k1) {
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    k0 = uint32_t(uint64_t(n) * L / gridDim.x);
    k1 = uint32_t(uint64_t(n) * (L + 1) / gridDim.x);
}

template <typename T>
__device__ __forceinline__ void b_dir_body(const BatchArgs<T>& a, int max_iter) {
    uint32_t k0, k1;
    tile_share(*a.n_act_tiles, k0, k1);
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t cs = ~0u;  // system whose scalars are held below
    int it = 0;
    bool stop = false;
    T rho = T(0), err = T(0), alpha = T(0), beta = T(0);
    for (uint32_t kb = k0; kb < k1; kb += kWave) {
        // stop[s] is raised in this launch, so it is not read here
        const TileDesc mine = tile_desc_load(a, kb + uint32_t(lane), k1, false);
        const int nb = int(k1 - kb < uint32_t(kWave) ? k1 - kb : uint32_t(kWave));
        for (int j = 0; j < nb; ++j) {
            const TileDesc td = tile_desc_at(mine, j);
            if (td.flags & 1u) continue;
            if (td.s != cs) {
                cs = td.s;
                it = a.iters[td.s];
                if (it > 0) {
                    rho = tiles_sum(a.part_rh, td.to0, td.to1);
                    err = tiles_sum(a.part_rr, td.to0, td.to1);
                    const int cap =
                        a.cap && a.cap[td.s] < max_iter ? a.cap[td.s] : max_iter;
                    stop = !(err > a.thresh[td.s]) || it >= cap;
                    alpha = a.alpha[td.s];
                    beta = rho / a.rho_prev[td.s];
                }
            }
            T xv[kTileRowsPerThread], dv[kTileRowsPerThread], hv[kTileRowsPerThread];
            if (it == 0) {
#pragma unroll
                for (int q = 0; q < kTileRowsPerThread; ++q) {
                    const uint32_t i = td.r0 + threadIdx.x + uint32_t(q * kBlock);
                    if (i < td.r1) hv[q] = a.h[i];
                }
#pragma unroll
                for (int q = 0; q < kTileRowsPerThread; ++q) {
                    const uint32_t i = td.r0 + threadIdx.x + uint32_t(q * kBlock);
                    if (i < td.r1) a.d[i] = hv[q];
                }
                continue;
            }
            if (td.t == td.to0 && threadIdx.x == 0) {
                a.err[td.s] = err;
                a.rho[td.s] = rho;
                if (stop) a.stop[td.s] = 1;
                else a.beta[td.s] = beta;
            }
#pragma unroll
            for (int q = 0; q < kTileRowsPerThread; ++q) {
                const uint32_t i = td.r0 + threadIdx.x + uint32_t(q * kBlock);
                if (i < td.r1) {
                    xv[q] = a.x[i];
                    dv[q] = a.d[i];
                    if (!stop) hv[q] = a.h[i];
                }
            }
#pragma unroll
            for (int q = 0; q < kTileRowsPerThread; ++q) {
                const uint32_t i = td.r0 + threadIdx.x + uint32_t(q * kBlock);
                if (i < td.r1) {
                    a.x[i] = add_rn(xv[q], mul_rn(-alpha, dv[q]));
                    if (!stop) a.d[i] = add_rn(mul_rn(beta, dv[q]), hv[q]);
                }
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_dir(BatchArgs<T> a, int max_iter) {
    b_dir_body<T>(a, max_iter);
}

// SpMV of the active tiles.  A wavefront streams R consecutive passes of one
// tile (R = 1, 2 or 4, chosen per launch from the number of active tiles):
// long runs keep a wavefront's loads in flight when tiles are plentiful,
// R = 1 spreads few tiles over all wavefronts.  4 / R wavefronts share a
// tile and a workgroup step covers R tiles.  Whatever R, a tile's d.q
// partial is formed in one order: per pass the sum over its lanes (each
// lane summing its rows in order), then the 4 pass sums in pass order, as
// block_sum_all does when wavefront w takes pass w.
template <typename T, int F>
__device__ __forceinline__ void b_spmv_body(const BatchArgs<T>& a, T* stage,
                                            T (*pass_sum)[4] /* [tile slot][pass] */) {
    const uint32_t n = *a.n_act_tiles;
    const uint32_t R = n >= a.spmv_r4_tiles ? 4u : (n >= a.spmv_r2_tiles ? 2u : 1u);
    const int lane = threadIdx.x & (kWave - 1);
    // wavefront index, known to the compiler as uniform across the wavefront
    const uint32_t w = uint32_t(__builtin_amdgcn_readfirstlane(int(threadIdx.x / kWave)));
    const uint32_t wpt = 4u / R;  // wavefronts per tile
    const uint32_t slot = w / wpt, first = (w % wpt) * R;
    // contiguous share of whole steps (R tiles each)
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint64_t steps = (uint64_t(n) + R - 1) / R;
    const uint32_t k0 = uint32_t(steps * L / gridDim.x) * R;
    const uint32_t k1e = uint32_t(steps * (L + 1) / gridDim.x) * R;
    const uint32_t k1 = k1e < n ? k1e : n;
    for (uint32_t kb = k0; kb < k1; kb += kWave) {
        const TileDesc mine = tile_desc_load(a, kb + uint32_t(lane), k1, true);
        const int nb = int(k1 - kb < uint32_t(kWave) ? k1 - kb : uint32_t(kWave));
        for (int j = 0; j < nb; j += int(R)) {
            const bool have = j + int(slot) < nb;
            TileDesc td{};
            if (have) td = tile_desc_at(mine, j + int(slot));
            const bool work = have && (td.flags & 3u) == 0u;
            if (have && td.flags == 2u && td.t == td.to0 && first == 0 && lane == 0)
                a.done[td.s] = 1;
            T acc0 = T(0), acc1 = T(0), acc2 = T(0), acc3 = T(0);
            if (work) {
                const bool blk = F != kFormAny || td.p0 != kNoBlocks;
                const uint32_t base = td.r0 + (blk ? 2u : 1u) * kTilePassRows * first;
                const uint32_t shift = blk ? 7u : 6u;  // log2 of rows per pass
                auto epi = prefetch_epi([&](uint32_t row) { return a.d[row]; },
                                        [&](uint32_t row, T v, T dr) {
                                            a.q[row] = v;
                                            // the accumulator of the row's pass
                                            // (selects, not branches, keep the
                                            // four in registers)
                                            const uint32_t pp = (row - base) >> shift;
                                            const T p = dr * v;
                                            acc0 = pp == 0 ? acc0 + p : acc0;
                                            acc1 = pp == 1 ? acc1 + p : acc1;
                                            acc2 = pp == 2 ? acc2 + p : acc2;
                                            acc3 = pp == 3 ? acc3 + p : acc3;
                                        });
                wave_spmv_rows<T, F>(a, td.r0, td.r1, td.lo, td.p0, first, R, a.d, stage, epi);
            }
            acc0 = group_sum<kWave>(acc0);
            if (R > 1) acc1 = group_sum<kWave>(acc1);
            if (R == 4) {
                acc2 = group_sum<kWave>(acc2);
                acc3 = group_sum<kWave>(acc3);
                if (work && lane == 0) a.part_dq[td.t] = ((acc0 + acc1) + acc2) + acc3;
            } else {
                if (work && lane == 0) {
                    pass_sum[slot][first] = acc0;
                    if (R == 2) pass_sum[slot][first + 1] = acc1;
                }
                __syncthreads();
                if (work && first == 0 && lane == 0)
                    a.part_dq[td.t] = ((pass_sum[slot][0] + pass_sum[slot][1]) +
                                       pass_sum[slot][2]) + pass_sum[slot][3];
                __syncthreads();
            }
        }
    }
}

template <typename T, int F>
__global__ __launch_bounds__(kBlock) void k_b_spmv(BatchArgs<T> a) {
    __shared__ T stage[kBatchStage];
    __shared__ T pass_sum[kWavesPerBlock][4];
    b_spmv_body<T, F>(a, stage, pass_sum);
}

template <typename T, bool PREC>
__device__ __forceinline__ void b_update_body(const BatchArgs<T>& a, T* lds) {
    uint32_t k0, k1;
    tile_share(*a.n_act_tiles, k0, k1);
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t cs = ~0u;
    T rho = T(0), alpha = T(0);
    for (uint32_t kb = k0; kb < k1; kb += kWave) {
        const TileDesc mine = tile_desc_load(a, kb + uint32_t(lane), k1, false);
        const int nb = int(k1 - kb < uint32_t(kWave) ? k1 - kb : uint32_t(kWave));
        for (int j = 0; j < nb; ++j) {
            const TileDesc td = tile_desc_at(mine, j);
            if (td.flags & 1u) continue;
            if (td.s != cs) {
                cs = td.s;
                rho = a.rho[td.s];
                alpha = rho / tiles_sum(a.part_dq, td.to0, td.to1);
            }
            T rv[kTileRowsPerThread], qv[kTileRowsPerThread], pv[kTileRowsPerThread];
#pragma unroll
            for (int q = 0; q < kTileRowsPerThread; ++q) {
                const uint32_t i = td.r0 + threadIdx.x + uint32_t(q * kBlock);
                if (i < td.r1) {
                    rv[q] = a.r[i];
                    qv[q] = a.q[i];
                    if (PREC) pv[q] = a.dinv[i];
                }
            }
            T prh = T(0), prr = T(0);
#pragma unroll
            for (int q = 0; q < kTileRowsPerThread; ++q) {
                const uint32_t i = td.r0 + threadIdx.x + uint32_t(q * kBlock);
                if (i < td.r1) {
                    const T ri = add_rn(rv[q], mul_rn(-alpha, qv[q]));
                    const T hi = PREC ? mul_rn(pv[q], ri) : ri;
                    a.r[i] = ri;
                    a.h[i] = hi;
                    prh += ri * hi;
                    prr += ri * ri;
                }
            }
            prh = block_sum_all(prh, lds);
            prr = block_sum_all(prr, lds);
            if (threadIdx.x == 0) {
                a.part_rh[td.t] = prh;
                a.part_rr[td.t] = prr;
                if (td.t == td.to0) {
                    a.alpha[td.s] = alpha;
                    a.rho_prev[td.s] = rho;
                    a.iters[td.s] = a.iters[td.s] + 1;
                }
            }
        }
    }
}

template <typename T, bool PREC>
__global__ __launch_bounds__(kBlock) void k_b_update(BatchArgs<T> a) {
    __shared__ T lds[kWavesPerBlock];
    b_update_body<T, PREC>(a, lds);
}

// Active-list rebuild: count reset, then append (order is irrelevant to the
// results: every tile and system is computed the same wherever it sits).
__global__ void k_b_reset_counts(uint32_t* n_act_tiles, uint32_t* n_act_sys) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *n_act_tiles = 0;
        *n_act_sys = 0;
    }
}

template <typename T>
__device__ __forceinline__ void b_compact_body(const BatchArgs<T>& a) {
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < a.ntiles; t += stride) {
        if (!a.done[a.tile_sys[t]]) {
            const uint32_t k = atomicAdd(a.n_act_tiles, 1u);
            a.act_tiles[k] = t;
        }
    }
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.nsys; s += stride) {
        if (!a.done[s]) {
            const uint32_t k = atomicAdd(a.n_act_sys, 1u);
            a.act_sys[k] = s;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_compact(BatchArgs<T> a) {
    b_compact_body<T>(a);
}

// The same iterations in one persistent, cooperative launch: the direction,
// SpMV and update phases of n_iter iterations and then the active-list
// rebuild, each phase ended by a grid-wide barrier instead of a kernel
// boundary.  Every workgroup runs the phase bodies of the three-launch
// kernels above over its share of the active tiles, so the results are
// theirs, bit for bit.  All workgroups take the same path: they leave
// together after n_iter iterations, or after a rebuild that leaves no tile
// active (read by every workgroup after the same barrier).
template <typename T, bool PREC, int F>
__global__ __launch_bounds__(kBlock) void k_b_persist(BatchArgs<T> a, int max_iter,
                                                      uint32_t n_iter) {
    __shared__ T stage[kBatchStage];
    __shared__ T pass_sum[kWavesPerBlock][4];
    __shared__ T lds[kWavesPerBlock];
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    for (uint32_t it = 0; it < n_iter; ++it) {
        b_dir_body<T>(a, max_iter);
        grid.sync();
        b_spmv_body<T, F>(a, stage, pass_sum);
        grid.sync();
        b_update_body<T, PREC>(a, lds);
        grid.sync();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.n_act_tiles = 0;
        *a.n_act_sys = 0;
    }
    grid.sync();
    b_compact_body<T>(a);
}

// All tiles / systems active (before the initial residual).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_all_active(BatchArgs<T> a) {
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < a.ntiles; t += stride)
        a.act_tiles[t] = t;
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.nsys; s += stride)
        a.act_sys[s] = s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.n_act_tiles = a.ntiles;
        *a.n_act_sys = a.nsys;
    }
}

// ------------------------------------------------ FP64 outer refinement

// r = b - A x per tile (FP64), partials of r.r and b.b.
__global__ __launch_bounds__(kBlock) void k_b_residual64(BatchArgs<double> a) {
    __shared__ double stage[kBatchStage];
    __shared__ double lds[kWavesPerBlock];
    for (uint32_t t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
        uint32_t r0, r1;
        tile_rows(a, t, r0, r1);
        double srr = 0.0, sbb = 0.0;
        auto epi = prefetch_epi([&](uint32_t row) { return a.b[row]; },
                                [&](uint32_t row, double s, double bi) {
                                    const double ri = bi - s;
                                    a.r[row] = ri;
                                    srr += ri * ri;
                                    sbb += bi * bi;
                                });
        tile_spmv(a, t, r0, r1, tile_blk(a, t), a.x, stage, epi);
        srr = block_sum_all(srr, lds);
        sbb = block_sum_all(sbb, lds);
        if (threadIdx.x == 0) {
            a.part_rr[t] = srr;
            a.part_bb[t] = sbb;
        }
    }
}

// Per system: rr, bb -> err[s], bb[s].
__global__ __launch_bounds__(kBlock) void k_b_sys_norms(BatchArgs<double> a) {
    const uint32_t s = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (s >= a.nsys) return;
    const double rr = sys_sum(a, a.part_rr, s);
    const double bb = sys_sum(a, a.part_bb, s);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        a.err[s] = rr;
        a.bb[s] = bb;
    }
}

// r32 = float(r64 * scale[s]) over all tiles; x32 = 0.
__global__ __launch_bounds__(kBlock) void k_b_scale_cast(BatchArgs<double> a64,
                                                         BatchArgs<float> a32,
                                                         const double* scale) {
    for (uint32_t t = blockIdx.x; t < a64.ntiles; t += gridDim.x) {
        uint32_t r0, r1;
        tile_rows(a64, t, r0, r1);
        const double sc = scale[a64.tile_sys[t]];
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
            const_cast<float*>(a32.b)[i] = float(a64.r[i] * sc);
            a32.x[i] = 0.0f;
        }
    }
}

// x64 += unscale[s] * double(x32) for systems that ran an inner solve.
__global__ __launch_bounds__(kBlock) void k_b_ir_update(BatchArgs<double> a64,
                                                        BatchArgs<float> a32,
                                                        const double* unscale) {
    for (uint32_t t = blockIdx.x; t < a64.ntiles; t += gridDim.x) {
        const uint32_t s = a64.tile_sys[t];
        if (a32.skip[s]) continue;
        uint32_t r0, r1;
        tile_rows(a64, t, r0, r1);
        const double us = unscale[s];
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBlock)
            a64.x[i] = a64.x[i] + us * double(a32.x[i]);
    }
}

// dst = src on the tiles of the systems with flag[s].
__global__ __launch_bounds__(kBlock) void k_b_copy_systems(BatchArgs<double> a,
                                                           double* __restrict__ dst,
                                                           const double* __restrict__ src,
                                                           const int32_t* __restrict__ flag) {
    for (uint32_t t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
        if (!flag[a.tile_sys[t]]) continue;
        uint32_t r0, r1;
        tile_rows(a, t, r0, r1);
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBlock) dst[i] = src[i];
    }
}

// ------------------------------------------------------------- launchers

template <typename T>
hipError_t batch_start(const BatchArgs<T>& a, bool precond, T eps, int max_iter, uint32_t grid,
                       hipStream_t s) {
    k_b_all_active<T><<<grid, kBlock, 0, s>>>(a);
    if (precond) k_b_init<T, true><<<grid, kBlock, 0, s>>>(a);
    else k_b_init<T, false><<<grid, kBlock, 0, s>>>(a);
    const uint32_t sg = (a.nsys + kWavesPerBlock - 1) / kWavesPerBlock;
    k_b_sys_init<T><<<sg, kBlock, 0, s>>>(a, eps, max_iter);
    k_b_reset_counts<<<1, 64, 0, s>>>(a.n_act_tiles, a.n_act_sys);
    k_b_compact<T><<<grid, kBlock, 0, s>>>(a);
    return hipGetLastError();
}

template <typename T>
hipError_t batch_iteration(const BatchArgs<T>& a, bool precond, int max_iter, uint32_t grid,
                           hipStream_t s) {
    k_b_dir<T><<<grid, kBlock, 0, s>>>(a, max_iter);
    if (a.all_blocks && a.bci16) k_b_spmv<T, kFormBlock16><<<grid, kBlock, 0, s>>>(a);
    else if (a.all_blocks) k_b_spmv<T, kFormBlock32><<<grid, kBlock, 0, s>>>(a);
    else k_b_spmv<T, kFormAny><<<grid, kBlock, 0, s>>>(a);
    if (precond) k_b_update<T, true><<<grid, kBlock, 0, s>>>(a);
    else k_b_update<T, false><<<grid, kBlock, 0, s>>>(a);
    return hipGetLastError();
}

// The persistent kernel for a batch: its function and the largest grid whose
// workgroups are all resident at once (what a cooperative launch needs).
template <typename T>
static const void* persist_fn(const BatchArgs<T>& a, bool precond) {
    if (a.all_blocks && a.bci16)
        return precond ? reinterpret_cast<const void*>(&k_b_persist<T, true, kFormBlock16>)
                       : reinterpret_cast<const void*>(&k_b_persist<T, false, kFormBlock16>);
    if (a.all_blocks)
        return precond ? reinterpret_cast<const void*>(&k_b_persist<T, true, kFormBlock32>)
                       : reinterpret_cast<const void*>(&k_b_persist<T, false, kFormBlock32>);
    return precond ? reinterpret_cast<const void*>(&k_b_persist<T, true, kFormAny>)
                   : reinterpret_cast<const void*>(&k_b_persist<T, false, kFormAny>);
}

template <typename T>
hipError_t batch_persistent_grid(const BatchArgs<T>& a, bool precond, uint32_t* grid) {
    int per_cu = 0, dev = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_fn(a, precond),
                                                                kBlock, 0);
    if (e != hipSuccess) return e;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if (per_cu <= 0 || cus <= 0) return hipErrorInvalidConfiguration;
    *grid = uint32_t(per_cu) * uint32_t(cus);
    return hipSuccess;
}

template <typename T>
hipError_t batch_persistent(const BatchArgs<T>& a, bool precond, int max_iter, uint32_t n_iter,
                            uint32_t grid, hipStream_t s) {
    BatchArgs<T> args = a;
    void* params[] = {&args, &max_iter, &n_iter};
    return hipLaunchCooperativeKernel(persist_fn(a, precond), dim3(grid), dim3(kBlock), params,
                                      0, s);
}

template <typename T>
hipError_t batch_compact(const BatchArgs<T>& a, uint32_t grid, hipStream_t s) {
    k_b_reset_counts<<<1, 64, 0, s>>>(a.n_act_tiles, a.n_act_sys);
    k_b_compact<T><<<grid, kBlock, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t batch_residual64(const BatchArgs<double>& a, uint32_t grid, hipStream_t s) {
    k_b_residual64<<<grid, kBlock, 0, s>>>(a);
    const uint32_t sg = (a.nsys + kWavesPerBlock - 1) / kWavesPerBlock;
    k_b_sys_norms<<<sg, kBlock, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t batch_scale_cast(const BatchArgs<double>& a64, const BatchArgs<float>& a32,
                            const double* scale, uint32_t grid, hipStream_t s) {
    k_b_scale_cast<<<grid, kBlock, 0, s>>>(a64, a32, scale);
    return hipGetLastError();
}

hipError_t batch_copy_systems(const BatchArgs<double>& a, double* dst, const double* src,
                              const int32_t* flag, uint32_t grid, hipStream_t s) {
    k_b_copy_systems<<<grid, kBlock, 0, s>>>(a, dst, src, flag);
    return hipGetLastError();
}

hipError_t batch_ir_update(const BatchArgs<double>& a64, const BatchArgs<float>& a32,
                           const double* unscale, uint32_t grid, hipStream_t s) {
    k_b_ir_update<<<grid, kBlock, 0, s>>>(a64, a32, unscale);
    return hipGetLastError();
}

template hipError_t batch_start<double>(const BatchArgs<double>&, bool, double, int, uint32_t,
                                        hipStream_t);
template hipError_t batch_start<float>(const BatchArgs<float>&, bool, float, int, uint32_t,
                                       hipStream_t);
template hipError_t batch_iteration<double>(const BatchArgs<double>&, bool, int, uint32_t,
                                            hipStream_t);
template hipError_t batch_iteration<float>(const BatchArgs<float>&, bool, int, uint32_t,
                                           hipStream_t);
template hipError_t batch_compact<double>(const BatchArgs<double>&, uint32_t, hipStream_t);
template hipError_t batch_persistent_grid<double>(const BatchArgs<double>&, bool, uint32_t*);
template hipError_t batch_persistent_grid<float>(const BatchArgs<float>&, bool, uint32_t*);
template hipError_t batch_persistent<double>(const BatchArgs<double>&, bool, int, uint32_t,
                                             uint32_t, hipStream_t);
template hipError_t batch_persistent<float>(const BatchArgs<float>&, bool, int, uint32_t,
                                            uint32_t, hipStream_t);
template hipError_t batch_compact<float>(const BatchArgs<float>&, uint32_t, hipStream_t);

}  // namespace nlhip
