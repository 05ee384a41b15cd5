// Kernels of the generalized symmetric eigensolver (A x = lambda M x, M a
// positive diagonal): block shift-invert subspace iteration with
// Rayleigh-Ritz, the GPU counterpart of the ARPACK shift-invert solves
// behind spectral LSCM and manifold harmonics (src/lib/OGF/mesh/commands/
// mesh_grob_surface_commands.cpp:761-763, mesh_grob_spectral_commands.cpp:
// 74-116).  Blocks of P vectors are column-major n x P.  Every reduction has
// a fixed order (per workgroup in row order, then workgroups in index
// order), so results do not depend on timing.
#include "dev_common.h"
#include "nlhip_internal.h"

namespace nlhip {

namespace {

constexpr uint32_t kGramTile = 64;  // rows per LDS tile of the Gram kernel

// val += -sigma mass[row] on each row's diagonal entry; *missing = 1 when a
// row has none.
__global__ void k_eig_shift_diag(uint32_t n, const uint32_t* rp, const uint32_t* ci,
                                 double* val, const double* mass, double sigma,
                                 uint32_t* missing) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        bool found = false;
        for (uint32_t k = rp[r]; k < rp[r + 1]; ++k) {
            if (ci[k] == r) {
                val[k] = add_rn(val[k], mul_rn(-sigma, mass ? mass[r] : 1.0));
                found = true;
                break;
            }
        }
        if (!found) atomicOr(missing, 1u);
    }
}

// Y[:, j] = M X[:, j]
__global__ void k_eig_scale_rows(uint32_t n, uint32_t p, const double* mass, const double* X,
                                 double* Y) {
    const uint64_t total = uint64_t(n) * p;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride)
        Y[i] = mass ? mul_rn(mass[i % n], X[i]) : X[i];
}

// Per workgroup b: part[b][i][j] = sum over its rows r (in order) of
// Y[r, i] Z[r, j].  Rows of workgroup b: [b rpb, (b + 1) rpb).
template <int P>
__global__ __launch_bounds__(kBlock) void k_eig_gram(uint32_t n, const double* Y, const double* Z,
                                                     uint32_t rpb, double* part) {
    __shared__ double ys[kGramTile][P];
    __shared__ double zs[kGramTile][P];
    constexpr int E = (P * P + kBlock - 1) / kBlock;  // entries per thread
    double acc[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = 0.0;
    const uint32_t r0 = blockIdx.x * rpb;
    const uint32_t r1 = (r0 + rpb < n) ? r0 + rpb : n;
    for (uint32_t t0 = r0; t0 < r1; t0 += kGramTile) {
        const uint32_t rows = (t0 + kGramTile < r1) ? kGramTile : r1 - t0;
        for (uint32_t idx = threadIdx.x; idx < kGramTile * P; idx += kBlock) {
            const uint32_t col = idx / kGramTile, rr = idx % kGramTile;
            const bool in = rr < rows;
            ys[rr][col] = in ? Y[size_t(col) * n + t0 + rr] : 0.0;
            zs[rr][col] = in ? Z[size_t(col) * n + t0 + rr] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e < P * P) {
                const int i = e / P, j = e % P;
                double a = acc[q];
                for (uint32_t rr = 0; rr < rows; ++rr) a = add_rn(a, mul_rn(ys[rr][i], zs[rr][j]));
                acc[q] = a;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int e = threadIdx.x + q * kBlock;
        if (e < P * P) part[size_t(blockIdx.x) * P * P + e] = acc[q];
    }
}

// G[e] = sum over workgroups b (in order) of part[b][e].
__global__ void k_eig_reduce(uint32_t nblk, uint32_t m, const double* part, double* G) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (uint32_t b = 0; b < nblk; ++b) s = add_rn(s, part[size_t(b) * m + e]);
        G[e] = s;
    }
}

// X[r, j] = sum over i (in order) of Y[r, i] C[i][j]  (C row-major P x P).
template <int P>
__global__ __launch_bounds__(kBlock) void k_eig_combine(uint32_t n, const double* Y,
                                                        const double* C, double* X) {
    __shared__ double cs[P * P];
    for (int e = threadIdx.x; e < P * P; e += kBlock) cs[e] = C[e];
    __syncthreads();
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        double y[P];
#pragma unroll
        for (int i = 0; i < P; ++i) y[i] = Y[size_t(i) * n + r];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < P; ++i) s = add_rn(s, mul_rn(y[i], cs[i * P + j]));
            X[size_t(j) * n + r] = s;
        }
    }
}

// Per workgroup b and column j (rows [b rpb, (b + 1) rpb), in order):
// part[b][j] = sum (AX - theta_j M X)^2, part[b][P + j] = sum AX^2,
// part[b][2P + j] = sum (M X)^2.
template <int P>
__global__ __launch_bounds__(kBlock) void k_eig_residual(uint32_t n, const double* AX,
                                                         const double* X, const double* mass,
                                                         const double* theta, uint32_t rpb,
                                                         double* part) {
    __shared__ double lds[kWavesPerBlock];
    const uint32_t r0 = blockIdx.x * rpb;
    const uint32_t r1 = (r0 + rpb < n) ? r0 + rpb : n;
    for (int j = 0; j < P; ++j) {
        double sr = 0.0, sa = 0.0, sm = 0.0;
        const double th = theta[j];
        for (uint32_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
            const double ax = AX[size_t(j) * n + r];
            const double mx = mass ? mul_rn(mass[r], X[size_t(j) * n + r]) : X[size_t(j) * n + r];
            const double d = add_rn(ax, -mul_rn(th, mx));
            sr = add_rn(sr, mul_rn(d, d));
            sa = add_rn(sa, mul_rn(ax, ax));
            sm = add_rn(sm, mul_rn(mx, mx));
        }
        sr = block_sum_all(sr, lds);
        sa = block_sum_all(sa, lds);
        sm = block_sum_all(sm, lds);
        if (threadIdx.x == 0) {
            part[size_t(blockIdx.x) * 3 * P + j] = sr;
            part[size_t(blockIdx.x) * 3 * P + P + j] = sa;
            part[size_t(blockIdx.x) * 3 * P + 2 * P + j] = sm;
        }
    }
}

uint32_t blocks_for(uint64_t n) {
    const uint64_t g = (n + kBlock - 1) / kBlock;
    return uint32_t(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

}  // namespace

void eig_shift_diag(uint32_t n, const uint32_t* rp, const uint32_t* ci, double* val,
                    const double* mass, double sigma, hipStream_t s) {
    DevBuf<uint32_t> missing(1);
    NLHIP_CHECK(hipMemsetAsync(missing.get(), 0, 4, s));
    k_eig_shift_diag<<<blocks_for(n), kBlock, 0, s>>>(n, rp, ci, val, mass, sigma,
                                                      missing.get());
    NLHIP_CHECK(hipGetLastError());
    uint32_t h = 0;
    NLHIP_CHECK(hipMemcpyAsync(&h, missing.get(), 4, hipMemcpyDeviceToHost, s));
    NLHIP_CHECK(hipStreamSynchronize(s));
    if (h) throw Error(NLHIP_ERR_INVALID, "a row of the matrix has no diagonal entry");
}

void eig_scale_rows(uint32_t n, uint32_t p, const double* mass, const double* X, double* Y,
                    hipStream_t s) {
    k_eig_scale_rows<<<blocks_for(uint64_t(n) * p), kBlock, 0, s>>>(n, p, mass, X, Y);
    NLHIP_CHECK(hipGetLastError());
}

// Rows per workgroup of the blocked reductions: whole tiles, at most 1024
// workgroups.
static uint32_t rows_per_group(uint32_t n) {
    uint32_t rpb = (n + 1023) / 1024;
    rpb = (rpb + kGramTile - 1) / kGramTile * kGramTile;
    return rpb < kGramTile ? kGramTile : rpb;
}

void eig_gram(uint32_t n, uint32_t p, const double* Y, const double* Z, double* G,
              DevBuf<double>& part, hipStream_t s) {
    const uint32_t rpb = rows_per_group(n);
    const uint32_t nblk = (n + rpb - 1) / rpb;
    part.ensure(size_t(nblk) * p * p);
    switch (p) {
        case 8: k_eig_gram<8><<<nblk, kBlock, 0, s>>>(n, Y, Z, rpb, part.get()); break;
        case 16: k_eig_gram<16><<<nblk, kBlock, 0, s>>>(n, Y, Z, rpb, part.get()); break;
        case 32: k_eig_gram<32><<<nblk, kBlock, 0, s>>>(n, Y, Z, rpb, part.get()); break;
        default: throw Error(NLHIP_ERR_INVALID, "block size must be 8, 16 or 32");
    }
    k_eig_reduce<<<blocks_for(p * p), kBlock, 0, s>>>(nblk, p * p, part.get(), G);
    NLHIP_CHECK(hipGetLastError());
}

void eig_combine(uint32_t n, uint32_t p, const double* Y, const double* C, double* X,
                 hipStream_t s) {
    const uint32_t g = blocks_for(n) < 4096 ? blocks_for(n) : 4096;
    switch (p) {
        case 8: k_eig_combine<8><<<g, kBlock, 0, s>>>(n, Y, C, X); break;
        case 16: k_eig_combine<16><<<g, kBlock, 0, s>>>(n, Y, C, X); break;
        case 32: k_eig_combine<32><<<g, kBlock, 0, s>>>(n, Y, C, X); break;
        default: throw Error(NLHIP_ERR_INVALID, "block size must be 8, 16 or 32");
    }
    NLHIP_CHECK(hipGetLastError());
}

void eig_residual(uint32_t n, uint32_t p, const double* AX, const double* X,
                  const double* mass, const double* theta, double* out,
                  DevBuf<double>& part, hipStream_t s) {
    const uint32_t rpb = rows_per_group(n);
    const uint32_t nblk = (n + rpb - 1) / rpb;
    part.ensure(size_t(nblk) * 3 * p);
    switch (p) {
        case 8: k_eig_residual<8><<<nblk, kBlock, 0, s>>>(n, AX, X, mass, theta, rpb, part.get()); break;
        case 16: k_eig_residual<16><<<nblk, kBlock, 0, s>>>(n, AX, X, mass, theta, rpb, part.get()); break;
        case 32: k_eig_residual<32><<<nblk, kBlock, 0, s>>>(n, AX, X, mass, theta, rpb, part.get()); break;
        default: throw Error(NLHIP_ERR_INVALID, "block size must be 8, 16 or 32");
    }
    k_eig_reduce<<<blocks_for(3 * p), kBlock, 0, s>>>(nblk, 3 * p, part.get(), out);
    NLHIP_CHECK(hipGetLastError());
}

}  // namespace nlhip
