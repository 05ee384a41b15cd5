// Batched Jacobi-PCG over many independent systems (the per-chart solves of
// a texture atlas, src/lib/OGF/mesh/commands/mesh_grob_surface_commands.cpp:
// 724-730), run as one fixed grid of persistent workgroups.
//
// Layout: the systems are packed as one block-diagonal CRS (global row and
// column indices), rows of system s in [row_off[s], row_off[s+1]).  Each
// system is cut into tiles of kTileRows rows that never straddle systems.
// A workgroup walks the device-side list of active tiles (grid stride), so
// one launch covers every unconverged system whatever their sizes.
//
// Per iteration and per system the arithmetic is the CG of k_cg.hip:
//   direction d = beta d + h;  spmv q = A d (+ tile partials of d.q);
//   reduce-A alpha = rho / sum(d.q);  update x, r, h (+ partials r.h, r.r);
//   reduce-B rho', err, stopping test, beta = rho' / rho.
// System scalars are reduced by one wavefront per system over that system's
// tile partials in a fixed order, so a system's result depends only on the
// system itself: it is bitwise the same solved alone or in any batch.
#include "dev_common.h"
#include "nlhip_kernels.h"

namespace nlhip {

// Rows [r0, r1) of tile t: kTileRows rows, or kBlockTileRows for systems in
// 2x2 block form (a tile is one pass of the workgroup's 4 wavefronts).  The
// tiling of a system depends only on that system.
template <typename T>
__device__ __forceinline__ void tile_rows(const BatchArgs<T>& a, uint32_t t, uint32_t& r0,
                                          uint32_t& r1) {
    const uint32_t s = a.tile_sys[t];
    r0 = a.tile_row0[t];
    const uint32_t e = a.row_off[s + 1];
    const uint32_t tr = (a.sys_pair0 && a.sys_pair0[s] != kNoBlocks) ? kBlockTileRows : kTileRows;
    r1 = (r0 + tr < e) ? r0 + tr : e;
}

// Blocks per lane per chunk of the batched block engine (the single-system
// default, 4, was slower on the atlas's small charts), and the LDS staging
// per workgroup (elements), enough for either engine.
constexpr int kBatchEB = 2;
constexpr int kBatchStage =
    kWavesPerBlock * 64 * (kStreamE > kBatchEB * 4 ? kStreamE : kBatchEB * 4);

template <typename T>
__device__ __forceinline__ bool tile_blk(const BatchArgs<T>& a, uint32_t t) {
    return a.sys_pair0 && a.sys_pair0[a.tile_sys[t]] != kNoBlocks;
}

// Stream SpMV of one tile: wavefront w takes rows [r0 + 64 w, r0 + 64 w + 64)
// or, for a system in 2x2 block form (blk), the block rows of its row pairs
// [r0 + 128 w, r0 + 128 w + 128) (r0 - row_off[s] and r1 - row_off[s] are
// even for such a system).
template <typename T, typename Epi>
__device__ __forceinline__ void tile_spmv(const BatchArgs<T>& a, uint32_t t, uint32_t r0,
                                          uint32_t r1, bool blk, const T* x, T* stage,
                                          Epi& epi) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (blk) {
        static_assert(kBatchEB * 4 * 64 * kWavesPerBlock <= kBatchStage, "staging area too small");
        const uint32_t s = a.tile_sys[t], lo = a.row_off[s], p0 = a.sys_pair0[s];
        const int32_t delta = int32_t(lo - 2u * p0);
        const uint32_t B0 = p0 + (r0 - lo) / 2 + 64u * uint32_t(w), Bend = p0 + (r1 - lo) / 2;
        if (B0 >= Bend) return;
        const uint32_t B1 = (B0 + 64u < Bend) ? B0 + 64u : Bend;
        const Cols bcols{a.bcol, a.bci16, a.bsegb};
        T* st = stage + w * (64 * kBatchEB * 4);
        if (a.bci16)
            BlockStreamEngine<kBatchEB, true>::run_rows(B0, B1, a.brp, bcols, a.bval, x, st,
                                                        lane, epi, delta);
        else
            BlockStreamEngine<kBatchEB, false>::run_rows(B0, B1, a.brp, bcols, a.bval, x, st,
                                                         lane, epi, delta);
        return;
    }
    const uint32_t R0 = r0 + 64u * uint32_t(w);
    if (R0 >= r1) return;
    const uint32_t R1 = (R0 + 64u < r1) ? R0 + 64u : r1;
    const Cols cols{a.ci, a.ci16, a.segb};
    if (a.ci16)
        StreamEngine<kStreamE, true, true>::run_rows(R0, R1, a.rp, cols, a.val, x,
                                                     stage + w * (64 * kStreamE), lane, epi);
    else
        StreamEngine<kStreamE, true>::run_rows(R0, R1, a.rp, cols, a.val, x,
                                               stage + w * (64 * kStreamE), lane, epi);
}

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
__device__ __forceinline__ T sys_sum(const BatchArgs<T>& a, const T* part, uint32_t s) {
    const int lane = threadIdx.x & (kWave - 1);
    T v = T(0);
    for (uint32_t t = a.tile_off[s] + uint32_t(lane); t < a.tile_off[s + 1]; t += kWave)
        v += part[t];
    return group_sum<kWave>(v);
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
        a.err[s] = err;
        a.beta[s] = T(0);
        a.alpha[s] = T(0);
        a.iters[s] = 0;
        const bool skip = a.skip && a.skip[s];
        a.done[s] = (skip || !(err > thresh) || max_iter <= 0) ? 1 : 0;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_dir(BatchArgs<T> a) {
    const uint32_t n = *a.n_act_tiles;
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const uint32_t t = a.act_tiles[k];
        const uint32_t s = a.tile_sys[t];
        if (a.done[s]) continue;
        uint32_t r0, r1;
        tile_rows(a, t, r0, r1);
        const int it = a.iters[s];
        const T beta = a.beta[s];
        if (it == 0) {
            for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBlock) a.d[i] = a.h[i];
        } else {
            for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBlock)
                a.d[i] = add_rn(mul_rn(beta, a.d[i]), a.h[i]);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_spmv(BatchArgs<T> a) {
    __shared__ T stage[kBatchStage];
    __shared__ T lds[kWavesPerBlock];
    const uint32_t n = *a.n_act_tiles;
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const uint32_t t = a.act_tiles[k];
        const uint32_t s = a.tile_sys[t];
        if (a.done[s]) continue;
        uint32_t r0, r1;
        tile_rows(a, t, r0, r1);
        T pdq = T(0);
        auto epi = prefetch_epi([&](uint32_t row) { return a.d[row]; },
                                [&](uint32_t row, T v, T dr) {
                                    a.q[row] = v;
                                    pdq += dr * v;
                                });
        tile_spmv(a, t, r0, r1, tile_blk(a, t), a.d, stage, epi);
        pdq = block_sum_all(pdq, lds);
        if (threadIdx.x == 0) a.part_dq[t] = pdq;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_reduce_alpha(BatchArgs<T> a) {
    const uint32_t ks = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (ks >= *a.n_act_sys) return;
    const uint32_t s = a.act_sys[ks];
    if (a.done[s]) return;
    const T dq = sys_sum(a, a.part_dq, s);
    if ((threadIdx.x & (kWave - 1)) == 0) a.alpha[s] = a.rho[s] / dq;
}

template <typename T, bool PREC>
__global__ __launch_bounds__(kBlock) void k_b_update(BatchArgs<T> a) {
    __shared__ T lds[kWavesPerBlock];
    const uint32_t n = *a.n_act_tiles;
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const uint32_t t = a.act_tiles[k];
        const uint32_t s = a.tile_sys[t];
        if (a.done[s]) continue;
        uint32_t r0, r1;
        tile_rows(a, t, r0, r1);
        const T alpha = a.alpha[s];
        T prh = T(0), prr = T(0);
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
            const T xi = add_rn(a.x[i], mul_rn(-alpha, a.d[i]));
            const T ri = add_rn(a.r[i], mul_rn(-alpha, a.q[i]));
            const T hi = PREC ? mul_rn(a.dinv[i], ri) : ri;
            a.x[i] = xi;
            a.r[i] = ri;
            a.h[i] = hi;
            prh += ri * hi;
            prr += ri * ri;
        }
        prh = block_sum_all(prh, lds);
        prr = block_sum_all(prr, lds);
        if (threadIdx.x == 0) {
            a.part_rh[t] = prh;
            a.part_rr[t] = prr;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_reduce_beta(BatchArgs<T> a, int max_iter) {
    const uint32_t ks = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (ks >= *a.n_act_sys) return;
    const uint32_t s = a.act_sys[ks];
    if (a.done[s]) return;
    const T rho_new = sys_sum(a, a.part_rh, s);
    const T err = sys_sum(a, a.part_rr, s);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const int it = a.iters[s] + 1;
        a.iters[s] = it;
        a.err[s] = err;
        if (!(err > a.thresh[s]) || it >= max_iter) {
            a.done[s] = 1;
        } else {
            a.beta[s] = rho_new / a.rho[s];
            a.rho[s] = rho_new;
        }
    }
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
__global__ __launch_bounds__(kBlock) void k_b_compact(BatchArgs<T> a) {
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
    const uint32_t sg = (a.nsys + kWavesPerBlock - 1) / kWavesPerBlock;
    k_b_dir<T><<<grid, kBlock, 0, s>>>(a);
    k_b_spmv<T><<<grid, kBlock, 0, s>>>(a);
    k_b_reduce_alpha<T><<<sg, kBlock, 0, s>>>(a);
    if (precond) k_b_update<T, true><<<grid, kBlock, 0, s>>>(a);
    else k_b_update<T, false><<<grid, kBlock, 0, s>>>(a);
    k_b_reduce_beta<T><<<sg, kBlock, 0, s>>>(a, max_iter);
    return hipGetLastError();
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
template hipError_t batch_compact<float>(const BatchArgs<float>&, uint32_t, hipStream_t);

}  // namespace nlhip
