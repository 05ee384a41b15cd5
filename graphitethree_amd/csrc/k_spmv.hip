// CRS sparse matrix-vector kernels for gfx950.
//
// Layout: rowptr[m+1] (u32), colind[nnz] (u32), val[nnz] (f64 or f32).
// Two engines (dev_common.h), selected by rows per wavefront (RPW):
//   RPW in {2,4,8,16}: 64/RPW lanes per row, xor-butterfly row sums;
//   RPW = 64: one row per lane, nonzeros streamed 8 per lane per chunk with
//   non-temporal loads, products staged in LDS, rows summed left to right.
// Workgroups own contiguous row ranges, remapped so each XCD owns one
// contiguous slice of the matrix (xcd_logical_block): the x gathers of
// neighbouring rows then hit the same L2.
//
// Follows nlMultMatrixVector on CRS matrices (y_i = sum_j a_ij x_j),
// called at src/lib/OGF/scene_graph/NL/matrix.cpp:303-305.
#include "dev_common.h"
#include "nlhip_kernels.h"

namespace nlhip {

template <int RPW, typename T>
__global__ __launch_bounds__(kBlock) void k_spmv(uint32_t m, uint32_t sp,
                                                 const uint32_t* __restrict__ rp, Cols cols,
                                                 const T* __restrict__ val,
                                                 const T* __restrict__ x, T* __restrict__ y) {
    __shared__ T stage[StageSize<RPW>::value];
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    auto epi = plain_epi([&](uint32_t row, T s) { y[row] = s; });
    EngineOf<RPW>::type::run(m, sp, L, rp, cols, val, x, stage, epi);
}

// r = b - A x (FP64), with per-block partial sums of r.r and b.b written to
// part_rr[L], part_bb[L] (L = logical block).
template <int RPW>
__global__ __launch_bounds__(kBlock) void k_residual(
    uint32_t m, uint32_t sp, const uint32_t* __restrict__ rp, Cols cols,
    const double* __restrict__ val, const double* __restrict__ x,
    const double* __restrict__ b, double* __restrict__ r, double* __restrict__ part_rr,
    double* __restrict__ part_bb) {
    __shared__ double stage[StageSize<RPW>::value];
    __shared__ double lds[kWavesPerBlock];
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    double srr = 0.0, sbb = 0.0;
    auto epi = prefetch_epi([&](uint32_t row) { return b[row]; },
                            [&](uint32_t row, double s, double bi) {
                                const double ri = bi - s;
                                if (r) r[row] = ri;
                                srr += ri * ri;
                                sbb += bi * bi;
                            });
    EngineOf<RPW>::type::run(m, sp, L, rp, cols, val, x, stage, epi);
    srr = block_sum_all(srr, lds);
    sbb = block_sum_all(sbb, lds);
    if (threadIdx.x == 0) {
        part_rr[L] = srr;
        part_bb[L] = sbb;
    }
}

// dinv[i] = 1 / a_ii (1 when the row has no or a zero diagonal entry);
// duplicate diagonal entries are summed.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_jacobi(uint32_t m, const uint32_t* __restrict__ rp,
                                                   const uint32_t* __restrict__ ci,
                                                   const double* __restrict__ val,
                                                   T* __restrict__ dinv) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    double d = 0.0;
    for (uint32_t j = rp[i]; j < rp[i + 1]; ++j)
        if (ci[j] == i) d += val[j];
    const T dt = T(d);
    dinv[i] = (dt != T(0)) ? T(1) / dt : T(1);
}

__global__ __launch_bounds__(kBlock) void k_cast_f32(uint64_t n, const double* __restrict__ a,
                                                     float* __restrict__ b) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        b[i] = float(a[i]);
}

// ---------------------------------------------------------------- launchers

uint32_t spmv_work_per_block(uint32_t m, uint32_t grid, int rpw) {
    if (rpw >= 64) {
        const uint64_t rows = rpw >= 73 ? m / 2 : m;  // 73..79: tiles of 64 block rows
        const uint64_t tiles = (rows + 63) / 64;
        uint64_t t = (tiles + grid - 1) / grid;
        return uint32_t(t ? t : 1);
    }
    const uint32_t rows_per_pass = uint32_t(kWavesPerBlock * rpw);
    uint64_t c = (uint64_t(m) + grid - 1) / grid;
    c = (c + rows_per_pass - 1) / rows_per_pass * rows_per_pass;
    if (c == 0) c = rows_per_pass;
    return uint32_t(c);
}

template <typename T>
hipError_t spmv(uint32_t m, const uint32_t* rp, const Cols& cols, const T* val, const T* x,
                T* y, int rpw, uint32_t grid, hipStream_t s) {
    if (m == 0) return hipSuccess;
    const uint32_t sp = spmv_work_per_block(m, grid, rpw);
    rpw = engine_rpw(rpw, cols);
    NLHIP_RPW_SWITCH(rpw, (k_spmv<RPW, T><<<grid, kBlock, 0, s>>>(m, sp, rp, cols, val, x, y)));
    return hipGetLastError();
}
template hipError_t spmv<double>(uint32_t, const uint32_t*, const Cols&, const double*,
                                 const double*, double*, int, uint32_t, hipStream_t);
template hipError_t spmv<float>(uint32_t, const uint32_t*, const Cols&, const float*,
                                const float*, float*, int, uint32_t, hipStream_t);

hipError_t residual64(uint32_t m, const uint32_t* rp, const Cols& cols, const double* val,
                      const double* x, const double* b, double* r, double* part_rr,
                      double* part_bb, int rpw, uint32_t grid, hipStream_t s) {
    const uint32_t sp = spmv_work_per_block(m, grid, rpw);
    rpw = engine_rpw(rpw, cols);
    NLHIP_RPW_SWITCH(rpw, (k_residual<RPW><<<grid, kBlock, 0, s>>>(m, sp, rp, cols, val, x, b,
                                                                   r, part_rr, part_bb)));
    return hipGetLastError();
}

// segb[s] = min column of segment s; *span = max over segments of the
// column range.  One workgroup per segment.
__global__ __launch_bounds__(kBlock) void k_seg_bases(uint64_t nnz, const uint32_t* __restrict__ ci,
                                                      uint32_t* __restrict__ segb,
                                                      uint32_t* __restrict__ span) {
    __shared__ uint32_t lo_s[kWavesPerBlock], hi_s[kWavesPerBlock];
    const uint64_t j0 = uint64_t(blockIdx.x) << kSegShift;
    uint32_t lo = 0xffffffffu, hi = 0;
    for (uint64_t j = j0 + threadIdx.x; j < j0 + kSegEntries && j < nnz; j += kBlock) {
        const uint32_t c = ci[j];
        lo = c < lo ? c : lo;
        hi = c > hi ? c : hi;
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const uint32_t l2 = __shfl_xor(lo, o, kWave), h2 = __shfl_xor(hi, o, kWave);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) {
        lo_s[w] = lo;
        hi_s[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kWavesPerBlock; ++k) {
            lo = lo_s[k] < lo ? lo_s[k] : lo;
            hi = hi_s[k] > hi ? hi_s[k] : hi;
        }
        segb[blockIdx.x] = lo;
        atomicMax(span, hi - lo);
    }
}

__global__ __launch_bounds__(kBlock) void k_seg_encode(uint64_t nnz, const uint32_t* __restrict__ ci,
                                                       const uint32_t* __restrict__ segb,
                                                       uint16_t* __restrict__ ci16) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t j = uint64_t(blockIdx.x) * kBlock + threadIdx.x; j < nnz; j += stride)
        ci16[j] = uint16_t(ci[j] - segb[j >> kSegShift]);
}

// 2x2 block form (see BlockStreamEngine): *bad stays 0 when every row pair
// (2I, 2I+1) has the same even-length column list made of pairs (2J, 2J+1).
__global__ __launch_bounds__(kBlock) void k_block_check(uint32_t nb,
                                                        const uint32_t* __restrict__ rp,
                                                        const uint32_t* __restrict__ ci,
                                                        uint32_t* __restrict__ bad) {
    const uint32_t I = blockIdx.x * kBlock + threadIdx.x;
    if (I >= nb) return;
    const uint32_t b0 = rp[2 * I], b1 = rp[2 * I + 1], e1 = rp[2 * I + 2];
    const uint32_t len = b1 - b0;
    uint32_t fail = (e1 - b1 != len) || (len & 1u);
    for (uint32_t q = 0; !fail && q < len; q += 2) {
        const uint32_t c = ci[b0 + q];
        fail = (c & 1u) || ci[b0 + q + 1] != c + 1u || ci[b1 + q] != c ||
               ci[b1 + q + 1] != c + 1u;
    }
    if (fail) atomicOr(bad, 1u);
}

// brp[I] = rp[2I] / 4; block (I, q/2): bcol = ci[rp[2I] + q] / 2, values
// (a(2I,c), a(2I,c+1), a(2I+1,c), a(2I+1,c+1)).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_block_pack(uint32_t nb,
                                                       const uint32_t* __restrict__ rp,
                                                       const uint32_t* __restrict__ ci,
                                                       const double* __restrict__ val,
                                                       uint32_t* __restrict__ brp,
                                                       uint32_t* __restrict__ bcol,
                                                       T* __restrict__ bval) {
    const uint32_t I = blockIdx.x * kBlock + threadIdx.x;
    if (I > nb) return;
    const uint32_t b0 = rp[2 * I];
    brp[I] = b0 / 4;
    if (I == nb) return;
    const uint32_t b1 = rp[2 * I + 1], len = b1 - b0;
    for (uint32_t q = 0; q < len; q += 2) {
        const uint64_t k = b0 / 4 + q / 2;
        if (bcol) bcol[k] = ci[b0 + q] / 2;
        bval[block_upper<T>(k)] = T(val[b0 + q]);
        bval[block_upper<T>(k) + 1] = T(val[b0 + q + 1]);
        bval[block_lower<T>(k)] = T(val[b1 + q]);
        bval[block_lower<T>(k) + 1] = T(val[b1 + q + 1]);
    }
}

// FP32 block values (block by block) from FP64 ones (grouped), nblk blocks.
__global__ __launch_bounds__(kBlock) void k_block_vals_f32(uint64_t nblk,
                                                           const double* __restrict__ src,
                                                           float* __restrict__ dst) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t k = uint64_t(blockIdx.x) * kBlock + threadIdx.x; k < nblk; k += stride) {
        const uint64_t u = block_upper<double>(k), l = block_lower<double>(k);
        dst[block_upper<float>(k)] = float(src[u]);
        dst[block_upper<float>(k) + 1] = float(src[u + 1]);
        dst[block_lower<float>(k)] = float(src[l]);
        dst[block_lower<float>(k) + 1] = float(src[l + 1]);
    }
}

hipError_t block_vals_f32(uint64_t nblk, const double* src, float* dst, hipStream_t s) {
    if (nblk == 0) return hipSuccess;
    const uint64_t g = (nblk + kBlock - 1) / kBlock;
    k_block_vals_f32<<<uint32_t(g < 65536 ? g : 65536), kBlock, 0, s>>>(nblk, src, dst);
    return hipGetLastError();
}

hipError_t block_check(uint32_t nb, const uint32_t* rp, const uint32_t* ci, uint32_t* bad,
                       hipStream_t s) {
    if (nb == 0) return hipSuccess;
    k_block_check<<<(nb + kBlock - 1) / kBlock, kBlock, 0, s>>>(nb, rp, ci, bad);
    return hipGetLastError();
}

template <typename T>
hipError_t block_pack(uint32_t nb, const uint32_t* rp, const uint32_t* ci, const double* val,
                      uint32_t* brp, uint32_t* bcol, T* bval, hipStream_t s) {
    k_block_pack<T><<<(nb + 1 + kBlock - 1) / kBlock, kBlock, 0, s>>>(nb, rp, ci, val, brp,
                                                                     bcol, bval);
    return hipGetLastError();
}
template hipError_t block_pack<double>(uint32_t, const uint32_t*, const uint32_t*,
                                       const double*, uint32_t*, uint32_t*, double*,
                                       hipStream_t);
template hipError_t block_pack<float>(uint32_t, const uint32_t*, const uint32_t*,
                                      const double*, uint32_t*, uint32_t*, float*, hipStream_t);

hipError_t seg_bases(uint64_t nnz, const uint32_t* ci, uint32_t* segb, uint32_t* span,
                     hipStream_t s) {
    const uint64_t nseg = (nnz + kSegEntries - 1) >> kSegShift;
    if (nseg == 0 || nseg > 0x7fffffffu) return hipErrorInvalidValue;
    k_seg_bases<<<uint32_t(nseg), kBlock, 0, s>>>(nnz, ci, segb, span);
    return hipGetLastError();
}

hipError_t seg_encode(uint64_t nnz, const uint32_t* ci, const uint32_t* segb, uint16_t* ci16,
                      hipStream_t s) {
    if (nnz == 0) return hipSuccess;
    uint64_t g = (nnz + kBlock - 1) / kBlock;
    if (g > 16384) g = 16384;
    k_seg_encode<<<uint32_t(g), kBlock, 0, s>>>(nnz, ci, segb, ci16);
    return hipGetLastError();
}

template <typename T>
hipError_t jacobi_dinv(uint32_t m, const uint32_t* rp, const uint32_t* ci, const double* val,
                       T* dinv, hipStream_t s) {
    if (m == 0) return hipSuccess;
    k_jacobi<T><<<(m + kBlock - 1) / kBlock, kBlock, 0, s>>>(m, rp, ci, val, dinv);
    return hipGetLastError();
}
template hipError_t jacobi_dinv<double>(uint32_t, const uint32_t*, const uint32_t*,
                                        const double*, double*, hipStream_t);
template hipError_t jacobi_dinv<float>(uint32_t, const uint32_t*, const uint32_t*,
                                       const double*, float*, hipStream_t);

hipError_t cast_f64_f32(uint64_t n, const double* a, float* b, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t g = (n + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    k_cast_f32<<<uint32_t(g), kBlock, 0, s>>>(n, a, b);
    return hipGetLastError();
}

}  // namespace nlhip
