// SSOR preconditioner on the device (nlNewSSORPreconditioner(M, omega);
// Matrix::solve_iterative passes omega = 1.5 at
// src/lib/OGF/scene_graph/NL/matrix.cpp:366-368):
//   w_i = (x_i - sum_{j<i} a_ij w_j) * omega / d_i      rows in increasing order
//   y_i = (w_i (d_i / omega) - sum_{j>i} a_ij y_j) * omega / d_i
//                                                       rows in decreasing order
//   y_i = (2 - omega) y_i
// Each sweep is a triangular solve.  The host sorts the rows into levels
// (a row's level is one more than the deepest row it reads), so the rows of
// one level are independent: a level with many rows is one launch over the
// whole grid, and a run of consecutive small levels is one workgroup that
// steps through them with a barrier between levels.  Each sweep keeps, in
// level order, the entries each row reads, packed in CRS order.  Every row
// sums them in that order with one rounding per product and per addition,
// so the result is that of the sequential sweep, bit for bit.
#include "dev_common.h"
#include "nlhip_internal.h"

namespace nlhip {

namespace {

// d_i = sum of row i's diagonal entries (CRS order), dscale_i = d_i / omega;
// *bad = 1 when a diagonal is zero or not finite.
__global__ void k_ssor_diag(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* val,
                            double omega, double* diag, double* dscale, uint32_t* bad) {
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t flag = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double d = 0.0;
        for (uint32_t e = rp[i]; e < rp[i + 1]; ++e)
            if (ci[e] == i) d = add_rn(d, val[e]);
        diag[i] = d;
        dscale[i] = d / omega;
        flag |= (d == 0.0 || !isfinite(d)) ? 1u : 0u;
    }
    if (flag) atomicOr(bad, 1u);
}

constexpr int kSsorU = 8;  // packed entries loaded (and gathers issued) at a time

// The row at position k of a sweep: its packed entries come in groups of
// kSsorU (all loads of a group, then all gathers, then the additions in
// order), so a row costs one round trip per group instead of one per entry.
template <bool UPPER>
__device__ __forceinline__ void ssor_row(const SsorArgs& a, const SsorSweepArgs& sw, uint32_t k) {
    const uint32_t i = sw.rows[k];
    const uint32_t e0 = sw.eoff[k], e1 = sw.eoff[k + 1];
    const double* v = UPPER ? a.y : a.w;
    const double xi = UPPER ? mul_rn(a.w[i], a.dscale[i]) : a.x[i];
    const double di = a.diag[i];
    double S = 0.0;
    for (uint32_t e = e0; e < e1; e += kSsorU) {
        uint32_t c[kSsorU];
        double av[kSsorU], g[kSsorU];
#pragma unroll
        for (int u = 0; u < kSsorU; ++u) {
            const uint32_t q = e + u < e1 ? e + u : e1 - 1u;
            c[u] = sw.col[q];
            av[u] = sw.val[q];
        }
#pragma unroll
        for (int u = 0; u < kSsorU; ++u) g[u] = v[c[u]];
#pragma unroll
        for (int u = 0; u < kSsorU; ++u)
            if (e + uint32_t(u) < e1) S = add_rn(S, mul_rn(av[u], g[u]));
    }
    const double t = mul_rn(add_rn(xi, -S), a.omega);
    (UPPER ? a.y : a.w)[i] = t / di;
}

// Levels [l0, l1): one level over the whole grid, or (one workgroup) several
// levels in order with a barrier after each.
template <bool UPPER>
__global__ __launch_bounds__(kBlock) void k_ssor_sweep(SsorArgs a, uint32_t l0, uint32_t l1) {
    const SsorSweepArgs& sw = UPPER ? a.up : a.lo;
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t l = l0; l < l1; ++l) {
        const uint32_t b = sw.off[l], e = sw.off[l + 1];
        for (uint32_t k = b + blockIdx.x * kBlock + threadIdx.x; k < e; k += stride)
            ssor_row<UPPER>(a, sw, k);
        if (l + 1 < l1) __syncthreads();
    }
}

__global__ void k_ssor_scale(SsorArgs a) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride)
        a.y[i] = mul_rn(a.post, a.y[i]);
}

uint32_t blocks_for(uint64_t n) {
    const uint64_t g = (n + kBlock - 1) / kBlock;
    return uint32_t(g < 1 ? 1 : (g > 65535 ? 65535 : g));
}

}  // namespace

hipError_t ssor_diag(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* val,
                     double omega, double* diag, double* dscale, uint32_t* bad, hipStream_t s) {
    if (n) k_ssor_diag<<<blocks_for(n), kBlock, 0, s>>>(n, rp, ci, val, omega, diag, dscale, bad);
    return hipGetLastError();
}

hipError_t ssor_sweep(const SsorArgs& a, bool upper, uint32_t l0, uint32_t l1, uint32_t rows,
                      hipStream_t s) {
    // several levels in one launch only on one workgroup (the barrier orders them)
    const uint32_t grid = (l1 - l0 > 1) ? 1u : blocks_for(rows);
    if (upper) k_ssor_sweep<true><<<grid, kBlock, 0, s>>>(a, l0, l1);
    else k_ssor_sweep<false><<<grid, kBlock, 0, s>>>(a, l0, l1);
    return hipGetLastError();
}

hipError_t ssor_scale(const SsorArgs& a, hipStream_t s) {
    if (a.n) k_ssor_scale<<<blocks_for(a.n), kBlock, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace nlhip
