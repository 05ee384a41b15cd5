// Multi right-hand-side Jacobi-PCG: k independent CG recurrences (one per
// right-hand side, as OpenNL's NL_NB_SYSTEMS loop solves them) that share
// every pass over the matrix.  Used for mesh_smooth's x, y, z solves
// (src/lib/OGF/mesh/commands/mesh_grob_surface_commands.cpp:703).
//
// Layout: vectors are n x k row-major ("interleaved"), element (i, j) at
// i*k + j, so the gather of x[col] for all right-hand sides is one
// contiguous k*8-byte segment.  The SpMM-shaped launches (SpMM step, start,
// residual) sum each row in column order with one rounding per product and
// per add, with one of two engines (dev_spmm.h), same sums:
// * k <= kMultiStreamMaxK: the stream engine, one row per lane, the matrix
//   streamed in 64-row tiles and each entry's k values gathered as one copy;
//   Pdq workgroups of sp tiles, as many as the device holds at once;
// * larger k: G-lane row groups (G = k rounded up to a power of two, lane j
//   owning right-hand side j), (val, colind) staged in LDS once per pass.
//
// Per iteration, 5 launches: direction D = beta_j D + H; SpMM Q = A D with
// per-workgroup partials of d_j.q_j; a reduction launch (alpha_j); update
// X, R, H with partials of r_j.h_j and r_j.r_j; a reduction launch
// (stopping test, beta_j).  Columns whose stopping test held are left
// untouched; once every column has stopped, the launches exit at once.
#include <algorithm>

#include "dev_common.h"
#include "dev_spmm.h"
#include "nlhip_internal.h"
#include "nlhip_kernels.h"

namespace nlhip {

// Rows [r0, r1) of logical block L for the element-wise kernels, with the
// same G-lane groups as spmm_rows: thread t handles column t % G of rows
// r0 + t / G, r0 + t / G + kBlock / G, ...
__device__ __forceinline__ void block_rows(const uint32_t m, const uint32_t rows_per_block,
                                           uint32_t L, uint32_t& r0, uint32_t& r1) {
    const uint64_t rb = uint64_t(L) * rows_per_block;
    r0 = rb < m ? uint32_t(rb) : m;
    r1 = (uint64_t(r0) + rows_per_block < m) ? r0 + rows_per_block : m;
}

// Whether any column still iterates (its done flag, written by an earlier
// launch, is clear); the same answer in every thread.
template <typename T>
__device__ __forceinline__ bool any_active(const MultiArgs<T>& a) {
    for (uint32_t j = 0; j < a.k; ++j)
        if (!a.done[j]) return true;
    return false;
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

// The same for V columns per lane (lane l of a G-lane row group holds the
// partials of columns l*V .. l*V + V - 1); lds holds kBlock * V values.
template <int V, typename T>
__device__ __forceinline__ void column_partials_v(const T (&v)[V], uint32_t k, uint32_t G,
                                                  T* lds, T* out /* k values */) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < V; ++c) lds[threadIdx.x * V + c] = v[c];
    __syncthreads();
    if (threadIdx.x < k) {
        const uint32_t lane = threadIdx.x / V, c = threadIdx.x % V;
        T s = T(0);
        for (uint32_t t = lane; t < kBlock; t += G) s += lds[t * V + c];
        out[threadIdx.x] = s;
    }
}

// Columns per lane of the row-group SpMM kernels for k right-hand sides
// (G = k rounded up to a power of two lanes per row with one): two, gathered
// as one 16-byte copy per entry, for FP64 when k is even and at least 6 (4-8 %
// faster FP64 solves; FP32 pairs, 8 bytes, measured no faster).
template <typename T>
bool multi_two_cols(uint32_t k) {
    return std::is_same<T, double>::value && k % 2 == 0 && k >= 6;
}
template <int G>
constexpr int two_cols_v() {
    return G >= 8 ? 2 : 1;
}

template <int G, int V, bool ONE, typename T, bool PREC>
__global__ __launch_bounds__(kBlock) void k_m_init(MultiArgs<T> a) {
    __shared__ T lds[kBlock * V];
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t k = a.k;
    T prh[V], prr[V], pbb[V];
#pragma unroll
    for (int c = 0; c < V; ++c) prh[c] = prr[c] = pbb[c] = T(0);
    auto pre = [&](uint32_t row, uint32_t j) {
        return Pair<T>{a.b[uint64_t(row) * k + j], PREC ? a.dinv[row] : T(1)};
    };
    auto fin = [&](uint32_t row, uint32_t j, T s, const Pair<T>& pv) {
        const uint64_t i = uint64_t(row) * k + j;
        const T bi = pv.a;
        const T ri = add_rn(s, mul_rn(T(-1), bi));
        const T hi = PREC ? mul_rn(pv.b, ri) : ri;
        a.r[i] = ri;
        a.h[i] = hi;
        const uint32_t c = j % V;  // the lane's columns are j - c .. j - c + V - 1
        prh[c] += ri * hi;
        prr[c] += ri * ri;
        pbb[c] += bi * bi;
    };
    const bool zero_x = a.x_nonzero && *a.x_nonzero == 0;  // X = 0: A X = 0 (block-uniform)
    spmm_rows_v<G / V, V, ONE>(a.m, k, a.rows_per_block, L, a.rp, a.ci, a.val, a.x, pre, fin,
                               nullptr, zero_x);
    column_partials_v<V>(prh, k, G / V, lds, a.part_rh + size_t(L) * k);
    column_partials_v<V>(prr, k, G / V, lds, a.part_rr + size_t(L) * k);
    column_partials_v<V>(pbb, k, G / V, lds, a.part_bb + size_t(L) * k);
}

// One workgroup per column: sum of the P partials of column j (stride k) in
// a fixed order (thread t takes partials t, t + 256, ..., then the block
// sums the threads); every thread gets the total.
template <typename T>
__device__ __forceinline__ T col_sum(const T* part, uint32_t P, uint32_t k, uint32_t j,
                                     T* lds) {
    T v = T(0);
    uint32_t p = threadIdx.x;
    for (; p + 3 * kBlock < P; p += 4 * kBlock) {
        const T a = part[size_t(p) * k + j], b = part[size_t(p + kBlock) * k + j];
        const T c = part[size_t(p + 2 * kBlock) * k + j];
        const T d = part[size_t(p + 3 * kBlock) * k + j];
        v = v + a + b + c + d;
    }
    for (; p < P; p += kBlock) v += part[size_t(p) * k + j];
    return block_sum_all(v, lds);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_m_sys_init(MultiArgs<T> a, T eps, int max_iter) {
    __shared__ T lds[kWavesPerBlock];
    const uint32_t j = blockIdx.x;
    // the start's partials: one per workgroup of the SpMM-shaped launch
    const T rho = col_sum(a.part_rh, a.Pdq, a.k, j, lds);
    const T err = col_sum(a.part_rr, a.Pdq, a.k, j, lds);
    const T bb = col_sum(a.part_bb, a.Pdq, a.k, j, lds);
    if (threadIdx.x == 0) {
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
    if (!any_active(a)) return;  // every column has stopped
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t k = a.k, G = a.G;
    const uint32_t g = threadIdx.x / G, j = threadIdx.x % G;
    if (j >= k || a.done[j]) return;
    uint32_t r0, r1;
    block_rows(a.m, a.rows_per_block, L, r0, r1);
    if (a.iters[j] == 0) {
        for (uint32_t row = r0 + g; row < r1; row += kBlock / G) {
            const uint64_t i = uint64_t(row) * k + j;
            a.d[i] = a.h[i];
        }
    } else {
        const T beta = a.beta[j];
        for (uint32_t row = r0 + g; row < r1; row += kBlock / G) {
            const uint64_t i = uint64_t(row) * k + j;
            a.d[i] = add_rn(mul_rn(beta, a.d[i]), a.h[i]);
        }
    }
}

template <int G, int V, bool ONE, typename T>
__global__ __launch_bounds__(kBlock) void k_m_spmm(MultiArgs<T> a) {
    __shared__ T lds[kBlock * V];
    if (!any_active(a)) return;  // every column has stopped (block-uniform)
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t k = a.k;
    T pdq[V];
#pragma unroll
    for (int c = 0; c < V; ++c) pdq[c] = T(0);
    auto pre = [&](uint32_t row, uint32_t j) { return a.d[uint64_t(row) * k + j]; };
    auto fin = [&](uint32_t row, uint32_t j, T s, T di) {
        a.q[uint64_t(row) * k + j] = s;
        pdq[j % V] += di * s;
    };
    spmm_rows_v<G / V, V, ONE>(a.m, k, a.rows_per_block, L, a.rp, a.ci, a.val, a.d, pre, fin,
                               a.done);
    column_partials_v<V>(pdq, k, G / V, lds, a.part_dq + size_t(L) * k);
}

// Wavefront entries per chunk / 64 of the stream kernels.
constexpr int kMultiStreamE = 4;

// The stream SpMM's epilogue: q[row, :] = sums, and d[row, j] * sums[j]
// added to column j's partial of d.q (row by row).
template <int K, typename T>
struct MultiDqEpilogue {
    const T* __restrict__ d;
    T* __restrict__ q;
    T (&pdq)[K];
    struct PV {
        T v[K];
    };
    __device__ __forceinline__ PV pre(uint32_t row) const {
        PV p;
#pragma unroll
        for (int j = 0; j < K; ++j) p.v[j] = d[uint64_t(row) * K + j];
        return p;
    }
    __device__ __forceinline__ void done(uint32_t row, const T (&s)[K], const PV& pv) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            q[uint64_t(row) * K + j] = s[j];
            pdq[j] += pv.v[j] * s[j];
        }
    }
};

// Per-workgroup partials of K columns: each wavefront's sum (xor butterfly),
// then the wavefronts in index order; out[j] for j < K.  Every thread of
// the block calls it (it synchronises the block).
template <int K, typename T>
__device__ __forceinline__ void stream_partials(const T (&v)[K], T* red, T* out) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();  // red of an earlier call has been read
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const T s = group_sum<kWave>(v[j]);
        if (lane == 0) red[w * K + j] = s;
    }
    __syncthreads();
    if (threadIdx.x < uint32_t(K)) {
        T s = red[threadIdx.x];
        for (int x = 1; x < kWavesPerBlock; ++x) s += red[x * K + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// SpMM step for k = K <= kMultiStreamMaxK with the stream engine: Q = A D
// for the columns still iterating, Pdq workgroups of sp tiles.
template <int K, bool D16, typename T>
__global__ __launch_bounds__(kBlock) void k_m_spmm_stream(MultiArgs<T> a) {
    using Eng = MultiStreamEngine<K, kMultiStreamE, D16>;
    __shared__ T stage[kWavesPerBlock * Eng::kStage];
    __shared__ T red[kWavesPerBlock * K];
    if (!any_active(a)) return;  // every column has stopped (block-uniform)
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    bool act[K];
    T pdq[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        act[j] = !a.done[j];
        pdq[j] = T(0);
    }
    MultiDqEpilogue<K, T> epi{a.d, a.q, pdq};
    Eng::run(a.m, a.sp, L, a.rp, a.cols, a.val, a.d, act, stage, epi);
    stream_partials<K>(pdq, red, a.part_dq + size_t(L) * K);
}

// The start's epilogue (k_m_init's, row by row): r = A x - b, h = D^-1 r,
// partials of r.h, r.r and b.b per column.
template <int K, typename T, bool PREC>
struct MultiInitEpilogue {
    const MultiArgs<T>& a;
    T (&prh)[K];
    T (&prr)[K];
    T (&pbb)[K];
    struct PV {
        T b[K];
        T dinv;
    };
    __device__ __forceinline__ PV pre(uint32_t row) const {
        PV p;
#pragma unroll
        for (int j = 0; j < K; ++j) p.b[j] = a.b[uint64_t(row) * K + j];
        p.dinv = PREC ? a.dinv[row] : T(1);
        return p;
    }
    __device__ __forceinline__ void done(uint32_t row, const T (&s)[K], const PV& pv) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint64_t i = uint64_t(row) * K + j;
            const T bi = pv.b[j];
            const T ri = add_rn(s[j], mul_rn(T(-1), bi));
            const T hi = PREC ? mul_rn(pv.dinv, ri) : ri;
            a.r[i] = ri;
            a.h[i] = hi;
            prh[j] += ri * hi;
            prr[j] += ri * ri;
            pbb[j] += bi * bi;
        }
    }
};

// k_m_init with the stream engine (k = K <= kMultiStreamMaxK): Pdq
// workgroups, Pdq partials of each kind.
template <int K, bool D16, typename T, bool PREC>
__global__ __launch_bounds__(kBlock) void k_m_init_stream(MultiArgs<T> a) {
    using Eng = MultiStreamEngine<K, kMultiStreamE, D16>;
    __shared__ T stage[kWavesPerBlock * Eng::kStage];
    __shared__ T red[kWavesPerBlock * K];
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    bool act[K];
    T prh[K], prr[K], pbb[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        act[j] = true;
        prh[j] = prr[j] = pbb[j] = T(0);
    }
    MultiInitEpilogue<K, T, PREC> epi{a, prh, prr, pbb};
    const bool zero_x = a.x_nonzero && *a.x_nonzero == 0;  // X = 0: A X = 0 (block-uniform)
    Eng::run(a.m, a.sp, L, a.rp, a.cols, a.val, a.x, act, stage, epi, zero_x);
    stream_partials<K>(prh, red, a.part_rh + size_t(L) * K);
    stream_partials<K>(prr, red, a.part_rr + size_t(L) * K);
    stream_partials<K>(pbb, red, a.part_bb + size_t(L) * K);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_m_reduce_alpha(MultiArgs<T> a) {
    __shared__ T lds[kWavesPerBlock];
    const uint32_t j = blockIdx.x;
    if (a.done[j]) return;
    const T dq = col_sum(a.part_dq, a.Pdq, a.k, j, lds);
    if (threadIdx.x == 0) a.alpha[j] = a.rho[j] / dq;
}

template <typename T, bool PREC>
__global__ __launch_bounds__(kBlock) void k_m_update(MultiArgs<T> a) {
    __shared__ T lds[kBlock];
    if (!any_active(a)) return;  // every column has stopped (block-uniform)
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t k = a.k;
    // element-wise over the same rows as the SpMM block, G lanes per row
    const uint32_t G = a.G;
    const uint32_t g = threadIdx.x / G, j = threadIdx.x % G;
    uint32_t r0, r1;
    block_rows(a.m, a.rows_per_block, L, r0, r1);
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
    __shared__ T lds[kWavesPerBlock];
    const uint32_t j = blockIdx.x;
    if (a.done[j]) return;
    const T rho_new = col_sum(a.part_rh, a.P, a.k, j, lds);
    const T err = col_sum(a.part_rr, a.P, a.k, j, lds);
    if (threadIdx.x == 0) {
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

// Re-opens the columns flagged in `again` (their loop stopped on the
// stopping test; the host has lowered their thresholds): the reduction the
// stopping launch skipped (rho_new from the r.h partials, beta = rho_new /
// rho, rho = rho_new) and done = 0.  The next launches continue those
// columns exactly as if their threshold had been the lower one from the
// start.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_m_reopen(MultiArgs<T> a, const int32_t* again) {
    __shared__ T lds[kWavesPerBlock];
    const uint32_t j = blockIdx.x;
    if (!again[j]) return;  // block-uniform
    const T rho_new = col_sum(a.part_rh, a.P, a.k, j, lds);
    if (threadIdx.x == 0) {
        a.beta[j] = rho_new / a.rho[j];
        a.rho[j] = rho_new;
        a.done[j] = 0;
    }
}


// ---------------------------------------------------- layout / refinement

// interleaved[i*k + j] = colmajor[j*n + i] (and back), G-lane groups per
// row (G = k rounded up to a power of two), grid-stride over rows.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_to_interleaved(uint32_t n, uint32_t k, uint32_t G,
                                                           const T* __restrict__ cm,
                                                           T* __restrict__ il,
                                                           int32_t* nonzero) {
    const uint32_t g = threadIdx.x / G, j = threadIdx.x % G, per = kBlock / G;
    if (j >= k) return;
    bool nz = false;
    for (uint64_t i = uint64_t(blockIdx.x) * per + g; i < n; i += uint64_t(gridDim.x) * per) {
        const T v = cm[uint64_t(j) * n + i];
        il[i * k + j] = v;
        nz |= !(v == T(0));
    }
    if (nonzero && nz) *nonzero = 1;  // every writer stores 1
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_to_colmajor(uint32_t n, uint32_t k, uint32_t G,
                                                        const T* __restrict__ il,
                                                        T* __restrict__ cm) {
    const uint32_t g = threadIdx.x / G, j = threadIdx.x % G, per = kBlock / G;
    if (j >= k) return;
    for (uint64_t i = uint64_t(blockIdx.x) * per + g; i < n; i += uint64_t(gridDim.x) * per)
        cm[uint64_t(j) * n + i] = il[i * k + j];
}

// R = B - A X (FP64, interleaved), per-column partial |r|^2 and |b|^2.
template <int G, int V, bool ONE>
__global__ __launch_bounds__(kBlock) void k_m_residual64(MultiArgs<double> a) {
    __shared__ double lds[kBlock * V];
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t k = a.k;
    double prr[V], pbb[V];
#pragma unroll
    for (int c = 0; c < V; ++c) prr[c] = pbb[c] = 0.0;
    auto pre = [&](uint32_t row, uint32_t j) { return a.b[uint64_t(row) * k + j]; };
    auto fin = [&](uint32_t row, uint32_t j, double s, double bi) {
        const double ri = bi - s;
        a.r[uint64_t(row) * k + j] = ri;
        prr[j % V] += ri * ri;
        pbb[j % V] += bi * bi;
    };
    spmm_rows_v<G / V, V, ONE>(a.m, k, a.rows_per_block, L, a.rp, a.ci, a.val, a.x, pre, fin);
    column_partials_v<V>(prr, k, G / V, lds, a.part_rr + size_t(L) * k);
    column_partials_v<V>(pbb, k, G / V, lds, a.part_bb + size_t(L) * k);
}

// k_m_residual64's epilogue for the stream engine: r = b - A x, partials of
// r.r and b.b per column.
template <int K>
struct MultiResidualEpilogue {
    const MultiArgs<double>& a;
    double (&prr)[K];
    double (&pbb)[K];
    struct PV {
        double b[K];
    };
    __device__ __forceinline__ PV pre(uint32_t row) const {
        PV p;
#pragma unroll
        for (int j = 0; j < K; ++j) p.b[j] = a.b[uint64_t(row) * K + j];
        return p;
    }
    __device__ __forceinline__ void done(uint32_t row, const double (&s)[K], const PV& pv) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const double bi = pv.b[j];
            const double ri = bi - s[j];
            a.r[uint64_t(row) * K + j] = ri;
            prr[j] += ri * ri;
            pbb[j] += bi * bi;
        }
    }
};

template <int K, bool D16>
__global__ __launch_bounds__(kBlock) void k_m_residual64_stream(MultiArgs<double> a) {
    using Eng = MultiStreamEngine<K, kMultiStreamE, D16>;
    __shared__ double stage[kWavesPerBlock * Eng::kStage];
    __shared__ double red[kWavesPerBlock * K];
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    bool act[K];
    double prr[K], pbb[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        act[j] = true;
        prr[j] = pbb[j] = 0.0;
    }
    MultiResidualEpilogue<K> epi{a, prr, pbb};
    Eng::run(a.m, a.sp, L, a.rp, a.cols, a.val, a.x, act, stage, epi);
    stream_partials<K>(prr, red, a.part_rr + size_t(L) * K);
    stream_partials<K>(pbb, red, a.part_bb + size_t(L) * K);
}

__global__ __launch_bounds__(kBlock) void k_m_norms64(MultiArgs<double> a) {
    __shared__ double lds[kWavesPerBlock];
    const uint32_t j = blockIdx.x;
    const double rr = col_sum(a.part_rr, a.Pdq, a.k, j, lds);
    const double bb = col_sum(a.part_bb, a.Pdq, a.k, j, lds);
    if (threadIdx.x == 0) {
        a.err[j] = rr;
        a.bb[j] = bb;
    }
}

// b32 = float(r64 * scale[j]); x32 = 0
__global__ __launch_bounds__(kBlock) void k_m_scale_cast(MultiArgs<double> a64,
                                                         MultiArgs<float> a32,
                                                         const double* scale) {
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t k = a64.k, G = a64.G;
    const uint32_t g = threadIdx.x / G, j = threadIdx.x % G;
    if (j >= k) return;
    const double sj = scale[j];
    uint32_t r0, r1;
    block_rows(a64.m, a64.rows_per_block, L, r0, r1);
    for (uint32_t row = r0 + g; row < r1; row += kBlock / G) {
        const uint64_t i = uint64_t(row) * k + j;
        const_cast<float*>(a32.b)[i] = float(a64.r[i] * sj);
        a32.x[i] = 0.0f;
    }
}

__global__ __launch_bounds__(kBlock) void k_m_ir_update(MultiArgs<double> a64,
                                                        MultiArgs<float> a32,
                                                        const double* unscale) {
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t k = a64.k, G = a64.G;
    const uint32_t g = threadIdx.x / G, j = threadIdx.x % G;
    if (j >= k || a32.skip[j]) return;
    const double uj = unscale[j];
    uint32_t r0, r1;
    block_rows(a64.m, a64.rows_per_block, L, r0, r1);
    for (uint32_t row = r0 + g; row < r1; row += kBlock / G) {
        const uint64_t i = uint64_t(row) * k + j;
        a64.x[i] = a64.x[i] + uj * double(a32.x[i]);
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

// K = k (1 .. kMultiStreamMaxK) and D16 as constants for the stream kernels.
#define NLHIP_KD_SWITCH(K_, D16_, CALL)                                                   \
    switch (K_) {                                                                          \
        case 1: { constexpr int K = 1; if (D16_) { constexpr bool D16 = true; CALL; }      \
                  else { constexpr bool D16 = false; CALL; } } break;                      \
        case 2: { constexpr int K = 2; if (D16_) { constexpr bool D16 = true; CALL; }      \
                  else { constexpr bool D16 = false; CALL; } } break;                      \
        case 3: { constexpr int K = 3; if (D16_) { constexpr bool D16 = true; CALL; }      \
                  else { constexpr bool D16 = false; CALL; } } break;                      \
        case 4: { constexpr int K = 4; if (D16_) { constexpr bool D16 = true; CALL; }      \
                  else { constexpr bool D16 = false; CALL; } } break;                      \
        default: return hipErrorInvalidValue;                                              \
    }
static_assert(kMultiStreamMaxK == 4, "one case per k up to kMultiStreamMaxK");

static uint32_t ew_grid_m(uint64_t n) {
    uint64_t g = (n + kBlock - 1) / kBlock;
    return uint32_t(g > 4096 ? 4096 : (g == 0 ? 1 : g));
}

namespace {

__global__ void k_window_max(uint32_t m, const uint32_t* rp, uint32_t pass, uint32_t* out) {
    const uint32_t nw = (m + 7) / 8;
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t best = 0;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += stride) {
        const uint32_t a = 8 * w;
        const uint32_t e = (uint64_t(a) + pass < m) ? a + pass : m;
        const uint32_t c = rp[e] - rp[a];
        best = c > best ? c : best;
    }
    atomicMax(out, best);
}

}  // namespace

// The largest number of entries in `pass` consecutive rows starting at a
// multiple of 8 rows: every pass spmm_rows takes starts at such a row
// (blocks hold multiples of 8 rows for every G).
bool spmm_one_chunk(uint32_t m, const uint32_t* rp, uint32_t G, hipStream_t s) {
    const uint32_t pass = std::max<uint32_t>(kBlock / G, kSpmmPass);
    DevBuf<uint32_t> out(1);
    NLHIP_CHECK(hipMemsetAsync(out.get(), 0, 4, s));
    k_window_max<<<ew_grid_m((uint64_t(m) + 7) / 8), kBlock, 0, s>>>(m, rp, pass, out.get());
    NLHIP_CHECK(hipGetLastError());
    uint32_t h = 0;
    NLHIP_CHECK(hipMemcpyAsync(&h, out.get(), 4, hipMemcpyDeviceToHost, s));
    NLHIP_CHECK(hipStreamSynchronize(s));
    return h <= kSpmmChunk;
}

namespace {
template <typename T>
hipError_t stream_spmm_kernel(uint32_t k, bool d16, const void*& f) {
    NLHIP_KD_SWITCH(k, d16, (f = reinterpret_cast<const void*>(&k_m_spmm_stream<K, D16, T>)));
    return hipSuccess;
}
}  // namespace

// Sized for the SpMM step, the launch of every iteration; the start and the
// residual (once per solve, more registers for their epilogues at large k)
// may hold fewer workgroups per CU and then run a second, partial round.
template <typename T>
uint32_t multi_stream_per_cu(uint32_t k, bool d16) {
    const void* f = nullptr;
    int per = 0;
    if (stream_spmm_kernel<T>(k, d16, f) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, kBlock, 0) != hipSuccess || per < 1)
        return 1;
    return uint32_t(per);
}
template uint32_t multi_stream_per_cu<double>(uint32_t, bool);
template uint32_t multi_stream_per_cu<float>(uint32_t, bool);

template <typename T>
hipError_t multi_start(const MultiArgs<T>& a, bool precond, T eps, int max_iter, hipStream_t s) {
    const bool d16 = a.cols.ci16 != nullptr;
    const bool v2 = multi_two_cols<T>(a.k);
    if (a.sp && precond) {
        NLHIP_KD_SWITCH(a.k, d16, (k_m_init_stream<K, D16, T, true><<<a.Pdq, kBlock, 0, s>>>(a)));
    } else if (a.sp) {
        NLHIP_KD_SWITCH(a.k, d16, (k_m_init_stream<K, D16, T, false><<<a.Pdq, kBlock, 0, s>>>(a)));
    } else if (precond && a.one_chunk) {
        NLHIP_G_SWITCH(a.G, (v2 ? k_m_init<G, two_cols_v<G>(), true, T, true><<<a.P, kBlock, 0, s>>>(a)
                               : k_m_init<G, 1, true, T, true><<<a.P, kBlock, 0, s>>>(a)));
    } else if (precond) {
        NLHIP_G_SWITCH(a.G, (v2 ? k_m_init<G, two_cols_v<G>(), false, T, true><<<a.P, kBlock, 0, s>>>(a)
                               : k_m_init<G, 1, false, T, true><<<a.P, kBlock, 0, s>>>(a)));
    } else if (a.one_chunk) {
        NLHIP_G_SWITCH(a.G, (v2 ? k_m_init<G, two_cols_v<G>(), true, T, false><<<a.P, kBlock, 0, s>>>(a)
                               : k_m_init<G, 1, true, T, false><<<a.P, kBlock, 0, s>>>(a)));
    } else {
        NLHIP_G_SWITCH(a.G, (v2 ? k_m_init<G, two_cols_v<G>(), false, T, false><<<a.P, kBlock, 0, s>>>(a)
                               : k_m_init<G, 1, false, T, false><<<a.P, kBlock, 0, s>>>(a)));
    }
    k_m_sys_init<T><<<a.k, kBlock, 0, s>>>(a, eps, max_iter);
    return hipGetLastError();
}

template <typename T>
hipError_t multi_iteration(const MultiArgs<T>& a, bool precond, int max_iter, hipStream_t s) {
    const uint32_t cg = a.k;  // one workgroup per column for the reductions
    const bool v2 = multi_two_cols<T>(a.k);
    k_m_dir<T><<<a.P, kBlock, 0, s>>>(a);
    if (a.sp) {
        NLHIP_KD_SWITCH(a.k, a.cols.ci16 != nullptr,
                        (k_m_spmm_stream<K, D16, T><<<a.Pdq, kBlock, 0, s>>>(a)));
    } else if (a.one_chunk) {
        NLHIP_G_SWITCH(a.G, (v2 ? k_m_spmm<G, two_cols_v<G>(), true, T><<<a.P, kBlock, 0, s>>>(a)
                               : k_m_spmm<G, 1, true, T><<<a.P, kBlock, 0, s>>>(a)));
    } else {
        NLHIP_G_SWITCH(a.G, (v2 ? k_m_spmm<G, two_cols_v<G>(), false, T><<<a.P, kBlock, 0, s>>>(a)
                               : k_m_spmm<G, 1, false, T><<<a.P, kBlock, 0, s>>>(a)));
    }
    k_m_reduce_alpha<T><<<cg, kBlock, 0, s>>>(a);
    if (precond) k_m_update<T, true><<<a.P, kBlock, 0, s>>>(a);
    else k_m_update<T, false><<<a.P, kBlock, 0, s>>>(a);
    k_m_reduce_beta<T><<<cg, kBlock, 0, s>>>(a, max_iter);
    return hipGetLastError();
}

template <typename T>
hipError_t multi_layout(uint32_t n, uint32_t k, const T* src, T* dst, bool to_interleaved,
                        hipStream_t s, int32_t* nonzero) {
    if (k == 0 || k > 32) return hipErrorInvalidValue;
    uint32_t G = 1;
    while (G < k) G <<= 1;
    const uint32_t g = ew_grid_m(uint64_t(n) * G);
    if (to_interleaved) k_to_interleaved<T><<<g, kBlock, 0, s>>>(n, k, G, src, dst, nonzero);
    else k_to_colmajor<T><<<g, kBlock, 0, s>>>(n, k, G, src, dst);
    return hipGetLastError();
}

hipError_t multi_reopen(const MultiArgs<double>& a, const int32_t* again, hipStream_t s) {
    k_m_reopen<double><<<a.k, kBlock, 0, s>>>(a, again);
    return hipGetLastError();
}

hipError_t multi_residual64(const MultiArgs<double>& a, hipStream_t s) {
    const bool v2 = multi_two_cols<double>(a.k);
    if (a.sp) {
        NLHIP_KD_SWITCH(a.k, a.cols.ci16 != nullptr,
                        (k_m_residual64_stream<K, D16><<<a.Pdq, kBlock, 0, s>>>(a)));
    } else if (a.one_chunk) {
        NLHIP_G_SWITCH(a.G, (v2 ? k_m_residual64<G, two_cols_v<G>(), true><<<a.P, kBlock, 0, s>>>(a)
                               : k_m_residual64<G, 1, true><<<a.P, kBlock, 0, s>>>(a)));
    } else {
        NLHIP_G_SWITCH(a.G, (v2 ? k_m_residual64<G, two_cols_v<G>(), false><<<a.P, kBlock, 0, s>>>(a)
                               : k_m_residual64<G, 1, false><<<a.P, kBlock, 0, s>>>(a)));
    }
    k_m_norms64<<<a.k, kBlock, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t multi_scale_cast(const MultiArgs<double>& a64, const MultiArgs<float>& a32,
                            const double* scale, hipStream_t s) {
    k_m_scale_cast<<<a64.P, kBlock, 0, s>>>(a64, a32, scale);
    return hipGetLastError();
}

hipError_t multi_ir_update(const MultiArgs<double>& a64, const MultiArgs<float>& a32,
                           const double* unscale, hipStream_t s) {
    k_m_ir_update<<<a64.P, kBlock, 0, s>>>(a64, a32, unscale);
    return hipGetLastError();
}

// dst[:, j] = src[:, j] for the columns j with flag[j] (interleaved n x k).
__global__ __launch_bounds__(kBlock) void k_m_copy_cols(uint64_t nk, uint32_t k,
                                                        double* __restrict__ dst,
                                                        const double* __restrict__ src,
                                                        const int32_t* __restrict__ flag) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < nk; i += stride)
        if (flag[i % k]) dst[i] = src[i];
}

hipError_t multi_copy_cols(uint32_t n, uint32_t k, double* dst, const double* src,
                           const int32_t* flag, hipStream_t s) {
    const uint64_t nk = uint64_t(n) * k;
    if (nk == 0) return hipSuccess;
    const uint64_t g = (nk + kBlock - 1) / kBlock;
    k_m_copy_cols<<<uint32_t(g < 16384 ? g : 16384), kBlock, 0, s>>>(nk, k, dst, src, flag);
    return hipGetLastError();
}

template hipError_t multi_start<double>(const MultiArgs<double>&, bool, double, int, hipStream_t);
template hipError_t multi_start<float>(const MultiArgs<float>&, bool, float, int, hipStream_t);
template hipError_t multi_iteration<double>(const MultiArgs<double>&, bool, int, hipStream_t);
template hipError_t multi_iteration<float>(const MultiArgs<float>&, bool, int, hipStream_t);
template hipError_t multi_layout<double>(uint32_t, uint32_t, const double*, double*, bool,
                                         hipStream_t, int32_t*);

}  // namespace nlhip
