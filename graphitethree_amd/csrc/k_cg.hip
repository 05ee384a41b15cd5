// Fused Jacobi-preconditioned CG for gfx950 (FP64, and FP32 for the inner
// solve of mixed-precision refinement).
//
// Follows the CG of nlSolveSystemIterative(..., NL_CG, ...)
// (src/lib/OGF/scene_graph/NL/matrix.cpp:390-393) with the residual
// convention r = A x - b and the operation order
//   r = A x - b; h = P r; d = h; rho = r.h; err = r.r
//   while err > eps^2 b.b and it < max_iter:
//     Ad = A d; alpha = rho / d.Ad; x -= alpha d; r -= alpha Ad; h = P r
//     rho' = r.h; beta = rho'/rho; rho = rho'; d = beta d + h; err = r.r
// and the stopping rule |Ax-b|/|b| < eps of matrix.h:177-179.
//
// One iteration is three launches and no host round trip:
//   direction  reduce the r.h and r.r partials -> rho, err; x -= alpha d
//              with the previous iteration's alpha and d (the update of x
//              is deferred to here, where d is read anyway); stopping test;
//              beta = rho / rho_prev; d = beta d + h (d = h at it 0)
//   spmv       q = A d and partials of d.q (up to kSpmvGridMul times the
//              workgroups of the other launches: more waves in flight for
//              the gathers)
//   update     reduce the d.q partials -> alpha; r -= alpha q; h = P r;
//              partials of r.h and r.r
// Per iteration that moves 12 nnz + 100 n bytes in FP64 with 32-bit column
// indices, 10 nnz + 100 n with 16-bit offsets (values, column indices,
// rowptr, d gathered once, and the vector streams).  Every block reduces the
// partials of the previous launch itself, in a fixed order, so all blocks
// hold bitwise-identical scalars.  When the stopping test holds, the
// direction launch still applies the pending x update and raises `stop`; the
// next launch turns it into `done`, after which every launch exits at once,
// so the host can replay a captured graph of many iterations and read the
// status afterwards.
#include "dev_common.h"
#include "nlhip_kernels.h"

namespace nlhip {

template <typename T>
__device__ __forceinline__ bool block_done(const CGStatus<T>* st, int* s_flag) {
    if (threadIdx.x == 0)
        *s_flag = __hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return *s_flag != 0;
}

template <int RPW, typename T, bool PREC>
__global__ __launch_bounds__(kBlock) void k_cg_init(CGArgs<T> a) {
    __shared__ T stage[StageSize<RPW>::value];
    __shared__ T lds[kWavesPerBlock];
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
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
    EngineOf<RPW>::type::run(a.m, a.sp, L, a.rp, a.cols, a.val, a.x, stage, epi);
    prh = block_sum_all(prh, lds);
    prr = block_sum_all(prr, lds);
    pbb = block_sum_all(pbb, lds);
    if (threadIdx.x == 0) {
        a.part_rh[L] = prh;
        a.part_rr[L] = prr;
        a.part_bb[L] = pbb;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_init_finish(CGArgs<T> a, T eps, int max_iter) {
    __shared__ T lds[kWavesPerBlock];
    const T bb = reduce_partials(a.part_bb, a.P, lds);
    if (threadIdx.x == 0) {
        CGStatus<T>* st = a.st;
        st->bb = bb;
        st->thresh = (eps * eps) * bb;
        st->rho = T(0);
        st->rho_prev = T(1);
        st->err = T(0);
        st->alpha = T(0);
        st->beta = T(0);
        st->iter_a = 0;
        st->iter_b = 0;
        st->iters = 0;
        st->max_iter = max_iter;
        st->stop = 0;
        __hip_atomic_store(&st->done, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_dir(CGArgs<T> a) {
    __shared__ T lds[kWavesPerBlock];
    __shared__ int s_flag;
    CGStatus<T>* st = a.st;
    // `done` is only raised by a later launch, never by this one, so every
    // block of the launch that stops still applies its part of the x update
    if (block_done(st, &s_flag)) return;
    const int it = st->iter_a;
    const T rho = reduce_partials(a.part_rh, a.P, lds);
    const T err = reduce_partials(a.part_rr, a.P, lds);
    const bool stop = !(err > st->thresh) || it >= st->max_iter;
    const T alpha = st->alpha;  // of iteration it - 1
    const T beta = (it > 0) ? rho / st->rho_prev : T(0);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rho = rho;
        st->err = err;
        if (stop) {
            st->iters = it;
            st->stop = 1;
        } else {
            st->beta = beta;
            st->iter_b = it;
        }
    }
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    uint32_t r0, r1;
    block_row_range(a.m, a.chunk, L, r0, r1);
    if (it == 0) {
        if (!stop)
            for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBlock) a.d[i] = a.h[i];
    } else if (stop) {
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += kBlock)
            a.x[i] = add_rn(a.x[i], mul_rn(-alpha, a.d[i]));
    } else {
        // four rows per thread at a time: every load before any store (the
        // arrays may alias as far as the compiler knows)
        uint32_t i = r0 + threadIdx.x;
        for (; i + 3 * kBlock < r1; i += 4 * kBlock) {
            T dv[4], xv[4], hv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                dv[q] = a.d[i + q * kBlock];
                xv[q] = a.x[i + q * kBlock];
                hv[q] = a.h[i + q * kBlock];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a.x[i + q * kBlock] = add_rn(xv[q], mul_rn(-alpha, dv[q]));
                a.d[i + q * kBlock] = add_rn(mul_rn(beta, dv[q]), hv[q]);
            }
        }
        for (; i < r1; i += kBlock) {
            const T di = a.d[i];
            a.x[i] = add_rn(a.x[i], mul_rn(-alpha, di));
            a.d[i] = add_rn(mul_rn(beta, di), a.h[i]);
        }
    }
}

template <int RPW, typename T>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(CGArgs<T> a) {
    __shared__ T stage[StageSize<RPW>::value];
    __shared__ T lds[kWavesPerBlock];
    __shared__ int s_flag;
    if (threadIdx.x == 0) {
        CGStatus<T>* st = a.st;
        int f = __hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!f && st->stop) {
            f = 1;
            if (blockIdx.x == 0)
                __hip_atomic_store(&st->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_flag = f;
    }
    __syncthreads();
    if (s_flag) return;
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    const T* __restrict__ d = a.d;
    T pdq = T(0);
    DqEpilogue<T> epi{d, a.q, pdq};
    EngineOf<RPW>::type::run(a.m, a.sps, L, a.rp, a.cols, a.val, d, stage, epi);
    pdq = block_sum_all(pdq, lds);
    if (threadIdx.x == 0) a.part_dq[L] = pdq;
}

template <typename T, bool PREC>
__global__ __launch_bounds__(kBlock) void k_cg_update(CGArgs<T> a) {
    __shared__ T lds[kWavesPerBlock];
    __shared__ int s_flag;
    CGStatus<T>* st = a.st;
    if (block_done(st, &s_flag)) return;
    const int it = st->iter_b;
    const T dq = reduce_partials(a.part_dq, a.Ps, lds);
    const T rho = st->rho;
    const T alpha = rho / dq;
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    uint32_t r0, r1;
    block_row_range(a.m, a.chunk, L, r0, r1);
    T prh = T(0), prr = T(0);
    // four rows per thread at a time, loads first; the partial sums take the
    // rows in the same order as one at a time
    uint32_t i = r0 + threadIdx.x;
    for (; i + 3 * kBlock < r1; i += 4 * kBlock) {
        T rv[4], qv[4], dv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            rv[q] = a.r[i + q * kBlock];
            qv[q] = a.q[i + q * kBlock];
            dv[q] = PREC ? a.dinv[i + q * kBlock] : T(1);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const T ri = add_rn(rv[q], mul_rn(-alpha, qv[q]));
            const T hi = PREC ? mul_rn(dv[q], ri) : ri;
            a.r[i + q * kBlock] = ri;
            a.h[i + q * kBlock] = hi;
            prh += ri * hi;
            prr += ri * ri;
        }
    }
    for (; i < r1; i += kBlock) {
        const T ri = add_rn(a.r[i], mul_rn(-alpha, a.q[i]));
        const T hi = PREC ? mul_rn(a.dinv[i], ri) : ri;
        a.r[i] = ri;
        a.h[i] = hi;
        prh += ri * hi;
        prr += ri * ri;
    }
    prh = block_sum_all(prh, lds);
    prr = block_sum_all(prr, lds);
    if (threadIdx.x == 0) {
        a.part_rh[L] = prh;
        a.part_rr[L] = prr;
        if (blockIdx.x == 0) {
            st->rho_prev = rho;
            st->alpha = alpha;
            st->iter_a = it + 1;
        }
    }
}

// ---------------------------------------------------- refinement helpers

// out = float(r * s)
__global__ __launch_bounds__(kBlock) void k_scale_cast(uint64_t n, const double* __restrict__ r,
                                                       double s, float* __restrict__ out) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        out[i] = float(r[i] * s);
}

// x += s * double(c)
__global__ __launch_bounds__(kBlock) void k_ir_update(uint64_t n, double* __restrict__ x,
                                                      const float* __restrict__ c, double s) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        x[i] = x[i] + s * double(c[i]);
}

// ------------------------------------------------------------- launchers

template <typename T>
hipError_t cg_start(const CGArgs<T>& a, int rpw, bool precond, T eps, int max_iter,
                    hipStream_t s) {
    if (a.m == 0 || a.P == 0) return hipErrorInvalidValue;
    rpw = engine_rpw(rpw, a.cols);
    if (precond) {
        NLHIP_RPW_SWITCH(rpw, (k_cg_init<RPW, T, true><<<a.P, kBlock, 0, s>>>(a)));
    } else {
        NLHIP_RPW_SWITCH(rpw, (k_cg_init<RPW, T, false><<<a.P, kBlock, 0, s>>>(a)));
    }
    k_cg_init_finish<T><<<1, kBlock, 0, s>>>(a, eps, max_iter);
    return hipGetLastError();
}

template <typename T>
hipError_t cg_step_part(const CGArgs<T>& a, int rpw, bool precond, int part, hipStream_t s) {
    if (part == 0) {
        k_cg_dir<T><<<a.P, kBlock, 0, s>>>(a);
    } else if (part == 1) {
        rpw = engine_rpw(rpw, a.cols);
        NLHIP_RPW_SWITCH(rpw, (k_cg_spmv<RPW, T><<<a.Ps, kBlock, 0, s>>>(a)));
    } else if (precond) {
        k_cg_update<T, true><<<a.P, kBlock, 0, s>>>(a);
    } else {
        k_cg_update<T, false><<<a.P, kBlock, 0, s>>>(a);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t cg_step(const CGArgs<T>& a, int rpw, bool precond, hipStream_t s) {
    for (int part = 0; part < 3; ++part) {
        hipError_t e = cg_step_part<T>(a, rpw, precond, part, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template hipError_t cg_start<double>(const CGArgs<double>&, int, bool, double, int, hipStream_t);
template hipError_t cg_start<float>(const CGArgs<float>&, int, bool, float, int, hipStream_t);
template hipError_t cg_step<double>(const CGArgs<double>&, int, bool, hipStream_t);
template hipError_t cg_step<float>(const CGArgs<float>&, int, bool, hipStream_t);
template hipError_t cg_step_part<double>(const CGArgs<double>&, int, bool, int, hipStream_t);
template hipError_t cg_step_part<float>(const CGArgs<float>&, int, bool, int, hipStream_t);

static uint32_t ew_grid(uint64_t n) {
    uint64_t g = (n + kBlock - 1) / kBlock;
    return uint32_t(g > 4096 ? 4096 : (g == 0 ? 1 : g));
}

hipError_t ir_scale_cast(uint64_t n, const double* r, double s, float* out, hipStream_t st) {
    k_scale_cast<<<ew_grid(n), kBlock, 0, st>>>(n, r, s, out);
    return hipGetLastError();
}

hipError_t ir_update(uint64_t n, double* x, const float* c, double s, hipStream_t st) {
    k_ir_update<<<ew_grid(n), kBlock, 0, st>>>(n, x, c, s);
    return hipGetLastError();
}

}  // namespace nlhip
