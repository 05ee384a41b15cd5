// NLBlas level-1 kernels (scene_graph/NL/blas.cpp:86,96,115,125,144 call
// Dcopy, Dscal, Ddot, Dnrm2, Daxpy with stride 1).
//
// Element-wise updates use separately rounded multiply and add, so y = y + a x
// rounds like the scalar host loop.  Dot products are two-stage: each block
// sums one contiguous chunk (threads strided, then a fixed wave butterfly and
// wave order), and a single block sums the per-block partials in index order;
// the result is bitwise reproducible for a given grid size.
#include "dev_common.h"
#include "nlhip_kernels.h"

namespace nlhip {

__global__ __launch_bounds__(kBlock) void k_dscal(uint64_t n, double a, double* __restrict__ x) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        x[i] = mul_rn(a, x[i]);
}

__global__ __launch_bounds__(kBlock) void k_daxpy(uint64_t n, double a,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ y) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        y[i] = add_rn(y[i], mul_rn(a, x[i]));
}

__global__ __launch_bounds__(kBlock) void k_dmul(uint64_t n, const double* __restrict__ a,
                                                 const double* __restrict__ x,
                                                 double* __restrict__ y) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        y[i] = mul_rn(a[i], x[i]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(uint64_t n, T v, T* __restrict__ x) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) x[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_dot_partial(uint64_t n, uint64_t chunk,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ y,
                                                        double* __restrict__ part) {
    __shared__ double lds[kWavesPerBlock];
    const uint64_t b = uint64_t(blockIdx.x) * chunk;
    const uint64_t e = (b + chunk < n) ? b + chunk : n;
    double s = 0.0;
    for (uint64_t i = b + threadIdx.x; i < e; i += kBlock) s += x[i] * y[i];
    s = block_sum_all(s, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[0] = sum(part[0..P)); out[1] = sqrt(out[0]) (for Dnrm2).
__global__ __launch_bounds__(kBlock) void k_reduce_final(const double* __restrict__ part,
                                                         uint32_t P, double* __restrict__ out) {
    __shared__ double lds[kWavesPerBlock];
    const double s = reduce_partials(part, P, lds);
    if (threadIdx.x == 0) {
        out[0] = s;
        out[1] = sqrt(s);
    }
}

static uint32_t elementwise_grid(uint64_t n) {
    uint64_t g = (n + kBlock - 1) / kBlock;
    if (g > 4096) g = 4096;
    if (g == 0) g = 1;
    return uint32_t(g);
}

hipError_t blas_dscal(uint64_t n, double a, double* x, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_dscal<<<elementwise_grid(n), kBlock, 0, s>>>(n, a, x);
    return hipGetLastError();
}

hipError_t blas_daxpy(uint64_t n, double a, const double* x, double* y, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_daxpy<<<elementwise_grid(n), kBlock, 0, s>>>(n, a, x, y);
    return hipGetLastError();
}

hipError_t blas_dmul(uint64_t n, const double* a, const double* x, double* y, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_dmul<<<elementwise_grid(n), kBlock, 0, s>>>(n, a, x, y);
    return hipGetLastError();
}

template <typename T>
hipError_t fill(uint64_t n, T v, T* x, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_fill<T><<<elementwise_grid(n), kBlock, 0, s>>>(n, v, x);
    return hipGetLastError();
}
template hipError_t fill<double>(uint64_t, double, double*, hipStream_t);
template hipError_t fill<float>(uint64_t, float, float*, hipStream_t);
template hipError_t fill<int32_t>(uint64_t, int32_t, int32_t*, hipStream_t);

hipError_t blas_ddot(uint64_t n, const double* x, const double* y, double* part,
                     uint32_t grid, double* out, hipStream_t s) {
    if (grid == 0) return hipErrorInvalidValue;
    uint64_t chunk = (n + grid - 1) / grid;
    if (chunk == 0) chunk = 1;
    k_dot_partial<<<grid, kBlock, 0, s>>>(n, chunk, x, y, part);
    k_reduce_final<<<1, kBlock, 0, s>>>(part, grid, out);
    return hipGetLastError();
}

// Plain device copy in 16-byte words (the bandwidth reference for the other
// kernels): every thread moves four words per trip, loads before stores.
__global__ __launch_bounds__(kBlock) void k_copy16(uint64_t n16, const uint4* __restrict__ src,
                                                   uint4* __restrict__ dst) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride],
                    d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n16; i += stride) dst[i] = src[i];
}

hipError_t copy16(uint64_t n16, const void* src, void* dst, uint32_t grid, hipStream_t s) {
    if (n16 == 0) return hipSuccess;
    if (grid == 0) return hipErrorInvalidValue;
    k_copy16<<<grid, kBlock, 0, s>>>(n16, static_cast<const uint4*>(src), static_cast<uint4*>(dst));
    return hipGetLastError();
}

// Read-only stream (the bandwidth reference for read-dominated kernels such
// as the SpMV): every thread loads eight 16-byte words per trip, all loads
// before their use, and folds them into one word written at the end (so the
// loads are not removed); NT: non-temporal loads.
template <bool NT>
__global__ __launch_bounds__(kBlock) void k_read16(uint64_t n16, const uint4* __restrict__ src,
                                                  uint4* __restrict__ out) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    uint4 acc = make_uint4(0u, 0u, 0u, 0u);
    for (; i + 7 * stride < n16; i += 8 * stride) {
        uint4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (NT) {
                using V4 = unsigned int __attribute__((ext_vector_type(4)));
                const V4 w = __builtin_nontemporal_load(reinterpret_cast<const V4*>(src + i + u * stride));
                v[u] = make_uint4(w.x, w.y, w.z, w.w);
            } else {
                v[u] = src[i + u * stride];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            acc.x ^= v[u].x;
            acc.y ^= v[u].y;
            acc.z ^= v[u].z;
            acc.w ^= v[u].w;
        }
    }
    for (; i < n16; i += stride) {
        const uint4 v = src[i];
        acc.x ^= v.x;
        acc.y ^= v.y;
        acc.z ^= v.z;
        acc.w ^= v.w;
    }
    out[uint64_t(blockIdx.x) * kBlock + threadIdx.x] = acc;
}

hipError_t read16(uint64_t n16, const void* src, void* out, uint32_t grid, bool nt,
                  hipStream_t s) {
    if (n16 == 0) return hipSuccess;
    if (grid == 0) return hipErrorInvalidValue;
    if (nt)
        k_read16<true><<<grid, kBlock, 0, s>>>(n16, static_cast<const uint4*>(src),
                                               static_cast<uint4*>(out));
    else
        k_read16<false><<<grid, kBlock, 0, s>>>(n16, static_cast<const uint4*>(src),
                                                static_cast<uint4*>(out));
    return hipGetLastError();
}

// The same read-only stream with each wavefront reading one contiguous run
// (a multiple of 256 words; wavefront v of the launch takes the v-th run),
// four words per lane per trip, non-temporal loads.
__global__ __launch_bounds__(kBlock) void k_read16_runs(uint64_t n16,
                                                        const uint4* __restrict__ src,
                                                        uint4* __restrict__ out) {
    using V4 = unsigned int __attribute__((ext_vector_type(4)));
    const V4* __restrict__ p = reinterpret_cast<const V4*>(src);
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t W = uint64_t(gridDim.x) * kWavesPerBlock;
    const uint64_t v = uint64_t(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
    const uint64_t per = ((n16 + W - 1) / W + 255) / 256 * 256;
    const uint64_t b = v * per < n16 ? v * per : n16;
    const uint64_t e = b + per < n16 ? b + per : n16;
    V4 acc = {0u, 0u, 0u, 0u};
    uint64_t i = b + lane;
    for (; i + 192 < e; i += 256) {
        V4 w[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) w[u] = __builtin_nontemporal_load(p + i + 64 * u);
#pragma unroll
        for (int u = 0; u < 4; ++u) acc ^= w[u];
    }
    for (; i < e; i += 64) acc ^= __builtin_nontemporal_load(p + i);
    out[uint64_t(blockIdx.x) * kBlock + threadIdx.x] = make_uint4(acc.x, acc.y, acc.z, acc.w);
}

hipError_t read16_runs(uint64_t n16, const void* src, void* out, uint32_t grid, hipStream_t s) {
    if (n16 == 0) return hipSuccess;
    if (grid == 0) return hipErrorInvalidValue;
    k_read16_runs<<<grid, kBlock, 0, s>>>(n16, static_cast<const uint4*>(src),
                                          static_cast<uint4*>(out));
    return hipGetLastError();
}

// The SpMV's memory traffic without its arithmetic: per 2x2 block, 32 bytes
// of values (its upper and lower pairs, in the grouped layout of
// block_upper / block_lower) and a 2-byte column offset streamed
// (non-temporal) and one 16-byte word gathered from a vector that stays in
// L2 (near the block's row); per row pair (every `bpr` blocks) one 16-byte
// word read and, with WRITE, one written.  Each wavefront streams one
// contiguous run of blocks, a multiple of 256 (wavefront v of the launch
// takes the v-th), 4 blocks per lane per trip; the row words of a trip are
// those of the row pairs whose last block it reads.
template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_stream_mix(uint64_t nblk, uint32_t bpr,
                                                       const double2* __restrict__ vals,
                                                       const uint16_t* __restrict__ off,
                                                       const double2* __restrict__ xg,
                                                       const double2* __restrict__ rin,
                                                       double2* __restrict__ rout,
                                                       double2* __restrict__ out) {
    using D2 = double __attribute__((ext_vector_type(2)));
    const D2* __restrict__ pv = reinterpret_cast<const D2*>(vals);
    const D2* __restrict__ px = reinterpret_cast<const D2*>(xg);
    const D2* __restrict__ pr = reinterpret_cast<const D2*>(rin);
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t nrows = nblk / bpr;
    const uint64_t W = uint64_t(gridDim.x) * kWavesPerBlock;
    const uint64_t v = uint64_t(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
    const uint64_t per = ((nblk + W - 1) / W + 255) / 256 * 256;
    const uint64_t b = v * per < nblk ? v * per : nblk;
    const uint64_t e = b + per < nblk ? b + per : nblk;
    D2 acc = {0.0, 0.0};
    for (uint64_t i0 = b; i0 < e; i0 += 256) {
        D2 lo[4], hi[4];
        uint32_t c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            uint64_t k = i0 + 64 * u + lane;
            k = k < e ? k : e - 1;
            c[u] = __builtin_nontemporal_load(off + k);
            lo[u] = __builtin_nontemporal_load(pv + block_upper<double>(k) / 2);
            hi[u] = __builtin_nontemporal_load(pv + block_lower<double>(k) / 2);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t k = i0 + 64 * u + lane;
            const D2 xv = px[(k / bpr + (c[u] & 3u)) % nrows];
            acc += (lo[u] + hi[u]) * xv;
        }
        const uint64_t i1 = i0 + 256 < e ? i0 + 256 : e;
        for (uint64_t r = i0 / bpr + lane; r < i1 / bpr && r < nrows; r += kWave) {
            const D2 rv = pr[r];
            if (WRITE) rout[r] = make_double2(acc.x * rv.x, acc.y * rv.y);
            else acc += rv;
        }
    }
    if (acc.x == 1.2345e300) out[0] = make_double2(acc.x, acc.y);
}

hipError_t stream_mix(uint64_t nblk, uint32_t bpr, const double* vals, const uint16_t* off,
                      const double* xg, const double* rin, double* rout, double* out,
                      bool write, uint32_t grid, hipStream_t s) {
    if (nblk == 0) return hipSuccess;
    if (grid == 0 || bpr == 0 || nblk < bpr) return hipErrorInvalidValue;
    auto* v = reinterpret_cast<const double2*>(vals);
    auto* x = reinterpret_cast<const double2*>(xg);
    auto* ri = reinterpret_cast<const double2*>(rin);
    auto* ro = reinterpret_cast<double2*>(rout);
    auto* o = reinterpret_cast<double2*>(out);
    if (write) k_stream_mix<true><<<grid, kBlock, 0, s>>>(nblk, bpr, v, off, x, ri, ro, o);
    else k_stream_mix<false><<<grid, kBlock, 0, s>>>(nblk, bpr, v, off, x, ri, ro, o);
    return hipGetLastError();
}

hipError_t reduce_final(const double* part, uint32_t P, double* out, hipStream_t s) {
    k_reduce_final<<<1, kBlock, 0, s>>>(part, P, out);
    return hipGetLastError();
}

}  // namespace nlhip
