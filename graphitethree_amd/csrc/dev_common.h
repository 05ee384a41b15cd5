// Device-side helpers shared by the HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

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

// Separately rounded multiply and add (no contraction), so element-wise
// updates round exactly like a scalar x = x + (a * y) loop.
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

}  // namespace nlhip
