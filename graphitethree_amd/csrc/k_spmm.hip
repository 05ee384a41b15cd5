// Multi right-hand-side Jacobi-PCG: k independent CG recurrences (one per
// right-hand side, as OpenNL's NL_NB_SYSTEMS loop solves them) that share
// every pass over the matrix.  Used for mesh_smooth's x, y, z solves
// (src/lib/OGF/mesh/commands/mesh_grob_surface_commands.cpp:703).
//
// Layout: vectors are n x k row-major ("interleaved"), element (i, j) at
// i*k + j, so the gather of x[col] for all right-hand sides is one
// contiguous k*8-byte segment.  A row is handled by a group of G lanes
// (G = k rounded up to a power of two), lane j of the group owning right-hand
// side j; the group walks the row's entries left to right, U at a time, so
// (val, colind) are loaded once per group and each lane sums its row in
// column order with one rounding per product and per add.
//
// Per iteration: direction D = beta_j D + H; SpMM Q = A D with per-workgroup
// partials of d_j.q_j; a reduction launch (alpha_j); update X, R, H with
// partials of r_j.h_j and r_j.r_j; a reduction launch (stopping test, beta_j).
// Columns whose stopping test held are left untouched.
#include "dev_common.h"
#include "nlhip_kernels.h"

namespace nlhip {

constexpr int kSpmmU = 4;  // entries per lane per step

template <int G, typename T, typename Epi>
__device__ __forceinline__ void spmm_rows(uint32_t m, uint32_t k, uint32_t rows_per_block,
                                          uint32_t L, const uint32_t* __restrict__ rp,
                                          const uint32_t* __restrict__ ci,
                                          const T* __restrict__ val, const T* __restrict__ X,
                                          Epi& epi) {
    const uint32_t g = threadIdx.x / G, j = threadIdx.x % G;
    const uint32_t groups = kBlock / G;
    const uint64_t rb = uint64_t(L) * rows_per_block;
    const uint32_t r0 = rb < m ? uint32_t(rb) : m;
    const uint32_t r1 = (uint64_t(r0) + rows_per_block < m) ? r0 + rows_per_block : m;
    const bool active_col = j < k;
    for (uint32_t row = r0 + g; row < r1; row += groups) {
        const uint32_t b = rp[row], e = rp[row + 1];
        T acc = T(0);
        for (uint32_t p = b; p < e; p += kSpmmU) {
            uint32_t c[kSpmmU];
            T a[kSpmmU], xv[kSpmmU];
#pragma unroll
            for (int u = 0; u < kSpmmU; ++u) {
                const uint32_t q = p + u < e ? p + u : e - 1u;
                c[u] = ci[q];
                a[u] = val[q];
            }
#pragma unroll
            for (int u = 0; u < kSpmmU; ++u)
                xv[u] = active_col ? X[uint64_t(c[u]) * k + j] : T(0);
#pragma unroll
            for (int u = 0; u < kSpmmU; ++u)
                if (p + u < e) acc = add_rn(acc, mul_rn(a[u], xv[u]));
        }
        if (active_col) epi(row, j, acc);
    }
}

// Per-workgroup partial sums of k columns: lds[kBlock] scratch; the group
// lanes holding column j combine through LDS in a fixed order.
template <typename T>
__device__ __forceinline__ void column_partials(T v, uint32_t k, uint32_t G, T* lds,
                                                T* out /* k values */) {
    __syncthreads();
    lds[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x < k) {
        T s = T(0);
        for (uint32_t t = threadIdx.x; t < kBlock; t += G) s += lds[t];
        out[threadIdx.x] = s;
    }
}

template <int G, typename T, bool PREC>
__global__ __launch_bounds__(kBlock) void k_m_init(MultiArgs<T> a) {
    __shared__ T lds[kBlock];
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t k = a.k;
    T prh = T(0), prr = T(0), pbb = T(0);
    auto epi = [&](uint32_t row, uint32_t j, T s) {
        const uint64_t i = uint64_t(row) * k + j;
        const T bi = a.b[i];
        const T ri = add_rn(s, mul_rn(T(-1), bi));
        const T hi = PREC ? mul_rn(a.dinv[row], ri) : ri;
        a.r[i] = ri;
        a.h[i] = hi;
        prh += ri * hi;
        prr += ri * ri;
        pbb += bi * bi;
    };
    spmm_rows<G>(a.m, k, a.rows_per_block, L, a.rp, a.ci, a.val, a.x, epi);
    column_partials(prh, k, G, lds, a.part_rh + size_t(L) * k);
    column_partials(prr, k, G, lds, a.part_rr + size_t(L) * k);
    column_partials(pbb, k, G, lds, a.part_bb + size_t(L) * k);
}

// One wavefront per column: sum of P partials (stride k), fixed order.
template <typename T>
__device__ __forceinline__ T col_sum(const T* part, uint32_t P, uint32_t k, uint32_t j) {
    const int lane = threadIdx.x & (kWave - 1);
    T v = T(0);
    for (uint32_t p = uint32_t(lane); p < P; p += kWave) v += part[size_t(p) * k + j];
    return group_sum<kWave>(v);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_m_sys_init(MultiArgs<T> a, T eps, int max_iter) {
    const uint32_t j = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (j >= a.k) return;
    const T rho = col_sum(a.part_rh, a.P, a.k, j);
    const T err = col_sum(a.part_rr, a.P, a.k, j);
    const T bb = col_sum(a.part_bb, a.P, a.k, j);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const T thresh = (eps * eps) * bb;
        a.bb[j] = bb;
        a.thresh[j] = thresh;
        a.rho[j] = rho;
        a.err[j] = err;
        a.beta[j] = T(0);
        a.alpha[j] = T(0);
        a.iters[j] = 0;
        const bool skip = a.skip && a.skip[j];
        a.done[j] = (skip || !(err > thresh) || max_iter <= 0) ? 1 : 0;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_m_dir(MultiArgs<T> a) {
    const uint32_t k = a.k;
    const uint64_t total = uint64_t(a.m) * k;
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < total; i += stride) {
        const uint32_t j = uint32_t(i % k);
        if (a.done[j]) continue;
        a.d[i] = (a.iters[j] == 0) ? a.h[i] : add_rn(mul_rn(a.beta[j], a.d[i]), a.h[i]);
    }
}

template <int G, typename T>
__global__ __launch_bounds__(kBlock) void k_m_spmm(MultiArgs<T> a) {
    __shared__ T lds[kBlock];
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t k = a.k;
    T pdq = T(0);
    auto epi = [&](uint32_t row, uint32_t j, T s) {
        const uint64_t i = uint64_t(row) * k + j;
        a.q[i] = s;
        pdq += a.d[i] * s;
    };
    spmm_rows<G>(a.m, k, a.rows_per_block, L, a.rp, a.ci, a.val, a.d, epi);
    column_partials(pdq, k, G, lds, a.part_dq + size_t(L) * k);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_m_reduce_alpha(MultiArgs<T> a) {
    const uint32_t j = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (j >= a.k || a.done[j]) return;
    const T dq = col_sum(a.part_dq, a.P, a.k, j);
    if ((threadIdx.x & (kWave - 1)) == 0) a.alpha[j] = a.rho[j] / dq;
}

template <typename T, bool PREC>
__global__ __launch_bounds__(kBlock) void k_m_update(MultiArgs<T> a) {
    __shared__ T lds[kBlock];
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t k = a.k;
    // element-wise over the same rows as the SpMM block, G lanes per row
    const uint32_t G = a.G;
    const uint32_t g = threadIdx.x / G, j = threadIdx.x % G;
    const uint64_t rb = uint64_t(L) * a.rows_per_block;
    const uint32_t r0 = rb < a.m ? uint32_t(rb) : a.m;
    const uint32_t r1 = (uint64_t(r0) + a.rows_per_block < a.m) ? r0 + a.rows_per_block : a.m;
    T prh = T(0), prr = T(0);
    if (j < k && !a.done[j]) {
        const T alpha = a.alpha[j];
        for (uint32_t row = r0 + g; row < r1; row += kBlock / G) {
            const uint64_t i = uint64_t(row) * k + j;
            const T xi = add_rn(a.x[i], mul_rn(-alpha, a.d[i]));
            const T ri = add_rn(a.r[i], mul_rn(-alpha, a.q[i]));
            const T hi = PREC ? mul_rn(a.dinv[row], ri) : ri;
            a.x[i] = xi;
            a.r[i] = ri;
            a.h[i] = hi;
            prh += ri * hi;
            prr += ri * ri;
        }
    }
    column_partials(prh, k, G, lds, a.part_rh + size_t(L) * k);
    column_partials(prr, k, G, lds, a.part_rr + size_t(L) * k);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_m_reduce_beta(MultiArgs<T> a, int max_iter) {
    const uint32_t j = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (j >= a.k || a.done[j]) return;
    const T rho_new = col_sum(a.part_rh, a.P, a.k, j);
    const T err = col_sum(a.part_rr, a.P, a.k, j);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const int it = a.iters[j] + 1;
        a.iters[j] = it;
        a.err[j] = err;
        if (!(err > a.thresh[j]) || it >= max_iter) {
            a.done[j] = 1;
        } else {
            a.beta[j] = rho_new / a.rho[j];
            a.rho[j] = rho_new;
        }
    }
}

// active column count -> *a.n_active (single thread, after reduce_beta)
template <typename T>
__global__ void k_m_count(MultiArgs<T> a) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint32_t act = 0;
        for (uint32_t c = 0; c < a.k; ++c) act += a.done[c] ? 0u : 1u;
        *a.n_active = act;
    }
}

// ---------------------------------------------------- layout / refinement

// interleaved[i*k + j] = colmajor[j*n + i] (and back)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_to_interleaved(uint32_t n, uint32_t k,
                                                           const T* __restrict__ cm,
                                                           T* __restrict__ il) {
    const uint64_t total = uint64_t(n) * k;
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t t = uint64_t(blockIdx.x) * kBlock + threadIdx.x; t < total; t += stride) {
        const uint64_t i = t / k, j = t % k;
        il[t] = cm[j * n + i];
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_to_colmajor(uint32_t n, uint32_t k,
                                                        const T* __restrict__ il,
                                                        T* __restrict__ cm) {
    const uint64_t total = uint64_t(n) * k;
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t t = uint64_t(blockIdx.x) * kBlock + threadIdx.x; t < total; t += stride) {
        const uint64_t j = t / n, i = t % n;
        cm[t] = il[i * k + j];
    }
}

// R = B - A X (FP64, interleaved), per-column partial |r|^2 and |b|^2.
template <int G>
__global__ __launch_bounds__(kBlock) void k_m_residual64(MultiArgs<double> a) {
    __shared__ double lds[kBlock];
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t k = a.k;
    double prr = 0.0, pbb = 0.0;
    auto epi = [&](uint32_t row, uint32_t j, double s) {
        const uint64_t i = uint64_t(row) * k + j;
        const double bi = a.b[i];
        const double ri = bi - s;
        a.r[i] = ri;
        prr += ri * ri;
        pbb += bi * bi;
    };
    spmm_rows<G>(a.m, k, a.rows_per_block, L, a.rp, a.ci, a.val, a.x, epi);
    column_partials(prr, k, G, lds, a.part_rr + size_t(L) * k);
    column_partials(pbb, k, G, lds, a.part_bb + size_t(L) * k);
}

__global__ __launch_bounds__(kBlock) void k_m_norms64(MultiArgs<double> a) {
    const uint32_t j = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (j >= a.k) return;
    const double rr = col_sum(a.part_rr, a.P, a.k, j);
    const double bb = col_sum(a.part_bb, a.P, a.k, j);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        a.err[j] = rr;
        a.bb[j] = bb;
    }
}

// b32 = float(r64 * scale[j]); x32 = 0
__global__ __launch_bounds__(kBlock) void k_m_scale_cast(MultiArgs<double> a64,
                                                         MultiArgs<float> a32,
                                                         const double* scale) {
    const uint32_t k = a64.k;
    const uint64_t total = uint64_t(a64.m) * k;
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < total; i += stride) {
        const_cast<float*>(a32.b)[i] = float(a64.r[i] * scale[i % k]);
        a32.x[i] = 0.0f;
    }
}

__global__ __launch_bounds__(kBlock) void k_m_ir_update(MultiArgs<double> a64,
                                                        MultiArgs<float> a32,
                                                        const double* unscale) {
    const uint32_t k = a64.k;
    const uint64_t total = uint64_t(a64.m) * k;
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < total; i += stride) {
        const uint32_t j = uint32_t(i % k);
        if (a32.skip[j]) continue;
        a64.x[i] = a64.x[i] + unscale[j] * double(a32.x[i]);
    }
}

// ------------------------------------------------------------- launchers

#define NLHIP_G_SWITCH(G_, CALL)                         \
    switch (G_) {                                        \
        case 1: { constexpr int G = 1; CALL; } break;    \
        case 2: { constexpr int G = 2; CALL; } break;    \
        case 4: { constexpr int G = 4; CALL; } break;    \
        case 8: { constexpr int G = 8; CALL; } break;    \
        case 16: { constexpr int G = 16; CALL; } break;  \
        case 32: { constexpr int G = 32; CALL; } break;  \
        default: return hipErrorInvalidValue;            \
    }

static uint32_t ew_grid_m(uint64_t n) {
    uint64_t g = (n + kBlock - 1) / kBlock;
    return uint32_t(g > 4096 ? 4096 : (g == 0 ? 1 : g));
}

template <typename T>
hipError_t multi_start(const MultiArgs<T>& a, bool precond, T eps, int max_iter, hipStream_t s) {
    if (precond) {
        NLHIP_G_SWITCH(a.G, (k_m_init<G, T, true><<<a.P, kBlock, 0, s>>>(a)));
    } else {
        NLHIP_G_SWITCH(a.G, (k_m_init<G, T, false><<<a.P, kBlock, 0, s>>>(a)));
    }
    k_m_sys_init<T><<<(a.k + kWavesPerBlock - 1) / kWavesPerBlock, kBlock, 0, s>>>(a, eps,
                                                                                   max_iter);
    k_m_count<T><<<1, 64, 0, s>>>(a);
    return hipGetLastError();
}

template <typename T>
hipError_t multi_iteration(const MultiArgs<T>& a, bool precond, int max_iter, hipStream_t s) {
    const uint32_t cg = (a.k + kWavesPerBlock - 1) / kWavesPerBlock;
    k_m_dir<T><<<ew_grid_m(uint64_t(a.m) * a.k), kBlock, 0, s>>>(a);
    NLHIP_G_SWITCH(a.G, (k_m_spmm<G, T><<<a.P, kBlock, 0, s>>>(a)));
    k_m_reduce_alpha<T><<<cg, kBlock, 0, s>>>(a);
    if (precond) k_m_update<T, true><<<a.P, kBlock, 0, s>>>(a);
    else k_m_update<T, false><<<a.P, kBlock, 0, s>>>(a);
    k_m_reduce_beta<T><<<cg, kBlock, 0, s>>>(a, max_iter);
    k_m_count<T><<<1, 64, 0, s>>>(a);
    return hipGetLastError();
}

template <typename T>
hipError_t multi_layout(uint32_t n, uint32_t k, const T* src, T* dst, bool to_interleaved,
                        hipStream_t s) {
    const uint32_t g = ew_grid_m(uint64_t(n) * k);
    if (to_interleaved) k_to_interleaved<T><<<g, kBlock, 0, s>>>(n, k, src, dst);
    else k_to_colmajor<T><<<g, kBlock, 0, s>>>(n, k, src, dst);
    return hipGetLastError();
}

hipError_t multi_residual64(const MultiArgs<double>& a, hipStream_t s) {
    NLHIP_G_SWITCH(a.G, (k_m_residual64<G><<<a.P, kBlock, 0, s>>>(a)));
    k_m_norms64<<<(a.k + kWavesPerBlock - 1) / kWavesPerBlock, kBlock, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t multi_scale_cast(const MultiArgs<double>& a64, const MultiArgs<float>& a32,
                            const double* scale, hipStream_t s) {
    k_m_scale_cast<<<ew_grid_m(uint64_t(a64.m) * a64.k), kBlock, 0, s>>>(a64, a32, scale);
    return hipGetLastError();
}

hipError_t multi_ir_update(const MultiArgs<double>& a64, const MultiArgs<float>& a32,
                           const double* unscale, hipStream_t s) {
    k_m_ir_update<<<ew_grid_m(uint64_t(a64.m) * a64.k), kBlock, 0, s>>>(a64, a32, unscale);
    return hipGetLastError();
}

template hipError_t multi_start<double>(const MultiArgs<double>&, bool, double, int, hipStream_t);
template hipError_t multi_start<float>(const MultiArgs<float>&, bool, float, int, hipStream_t);
template hipError_t multi_iteration<double>(const MultiArgs<double>&, bool, int, hipStream_t);
template hipError_t multi_iteration<float>(const MultiArgs<float>&, bool, int, hipStream_t);
template hipError_t multi_layout<double>(uint32_t, uint32_t, const double*, double*, bool,
                                         hipStream_t);

}  // namespace nlhip
