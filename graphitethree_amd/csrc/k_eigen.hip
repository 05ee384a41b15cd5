// Kernels of the generalized symmetric eigensolver (A x = lambda M x, M a
// positive diagonal): Chebyshev-filtered subspace iteration with
// Rayleigh-Ritz, the GPU counterpart of the ARPACK shift-invert solves
// behind spectral LSCM and manifold harmonics (src/lib/OGF/mesh/commands/
// mesh_grob_surface_commands.cpp:761-763, mesh_grob_spectral_commands.cpp:
// 74-116).  Blocks of P vectors are interleaved n x P (row r's P entries
// contiguous), the layout of the multi-vector SpMM.  Every reduction has a
// fixed order (per workgroup in row order, then workgroups in index order),
// so results do not depend on timing.
#include "dev_common.h"
#include "dev_spmm.h"
#include "nlhip_internal.h"

namespace nlhip {

namespace {

constexpr uint32_t kGramTile = 64;  // rows per LDS tile of the Gram kernel

// AY = A Y.
template <int P, bool ONE>
__global__ __launch_bounds__(kBlock) void k_eig_spmm(uint32_t n, uint32_t rpb, const uint32_t* rp,
                                                     const uint32_t* ci, const double* val,
                                                     const double* Y, double* AY) {
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    auto pre = [](uint32_t, uint32_t) { return 0; };
    auto fin = [&](uint32_t row, uint32_t j, double s, int) { AY[size_t(row) * P + j] = s; };
    // two columns per lane: P / 2 lanes per row, one 16-byte gather per entry
    spmm_rows_v<P / 2, 2, ONE>(n, P, rpb, L, rp, ci, val, Y, pre, fin);
}

// One step of the Chebyshev recurrence on T = M^-1 A:
//   Ynew = (T Y - c Y) s1 - s2 Xold       (s2 = 0 and no Xold on the first)
// The row's minv, Y and Xold are loaded with its gathers.
struct ChebPre {
    double minv, y, xold;
};

template <int P, bool ONE>
__global__ __launch_bounds__(kBlock) void k_eig_cheb(uint32_t n, uint32_t rpb, const uint32_t* rp,
                                                     const uint32_t* ci, const double* val,
                                                     const double* minv, const double* Y,
                                                     const double* Xold, double* Ynew, double c,
                                                     double s1, double s2) {
    const uint32_t L = xcd_logical_block(blockIdx.x, gridDim.x);
    auto pre = [&](uint32_t row, uint32_t j) {
        const size_t i = size_t(row) * P + j;
        return ChebPre{minv ? minv[row] : 1.0, Y[i], Xold ? Xold[i] : 0.0};
    };
    auto fin = [&](uint32_t row, uint32_t j, double s, const ChebPre& pv) {
        const double t = minv ? mul_rn(s, pv.minv) : s;
        double v = mul_rn(add_rn(t, -mul_rn(c, pv.y)), s1);
        if (Xold) v = add_rn(v, -mul_rn(s2, pv.xold));
        Ynew[size_t(row) * P + j] = v;
    };
    // two columns per lane: P / 2 lanes per row, one 16-byte gather per entry
    spmm_rows_v<P / 2, 2, ONE>(n, P, rpb, L, rp, ci, val, Y, pre, fin);
}

// Per workgroup b: part[b][i][j] = sum over its rows r (in order) of
// Y[r, i] (M Z)[r, j]  (mass null: M = I).  Rows of b: [b rpb, (b+1) rpb).
template <int P>
__global__ __launch_bounds__(kBlock) void k_eig_gram(uint32_t n, const double* Y, const double* Z,
                                                     const double* mass, uint32_t rpb,
                                                     double* part) {
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
            const uint32_t rr = idx / P, col = idx % P;
            const bool in = rr < rows;
            const size_t i = size_t(t0 + rr) * P + col;
            ys[rr][col] = in ? Y[i] : 0.0;
            zs[rr][col] = in ? (mass ? mul_rn(mass[t0 + rr], Z[i]) : Z[i]) : 0.0;
        }
        __syncthreads();
        if constexpr (P == 32) {
            // entries (i0, j0), (i0, j0 + 16), (i0 + 16, j0), (i0 + 16, j0 + 16):
            // four LDS reads per row for four products
            const int i0 = threadIdx.x / 16, j0 = threadIdx.x % 16;
            for (uint32_t rr = 0; rr < rows; ++rr) {
                const double y0 = ys[rr][i0], y1 = ys[rr][i0 + 16];
                const double z0 = zs[rr][j0], z1 = zs[rr][j0 + 16];
                acc[0] = add_rn(acc[0], mul_rn(y0, z0));
                acc[1] = add_rn(acc[1], mul_rn(y0, z1));
                acc[2] = add_rn(acc[2], mul_rn(y1, z0));
                acc[3] = add_rn(acc[3], mul_rn(y1, z1));
            }
        } else {
#pragma unroll
            for (int q = 0; q < E; ++q) {
                const int e = threadIdx.x + q * kBlock;
                if (e < P * P) {
                    const int i = e / P, j = e % P;
                    double a = acc[q];
                    for (uint32_t rr = 0; rr < rows; ++rr)
                        a = add_rn(a, mul_rn(ys[rr][i], zs[rr][j]));
                    acc[q] = a;
                }
            }
        }
        __syncthreads();
    }
    if constexpr (P == 32) {
        const int i0 = threadIdx.x / 16, j0 = threadIdx.x % 16;
        double* out = part + size_t(blockIdx.x) * P * P;
        out[i0 * P + j0] = acc[0];
        out[i0 * P + j0 + 16] = acc[1];
        out[(i0 + 16) * P + j0] = acc[2];
        out[(i0 + 16) * P + j0 + 16] = acc[3];
    } else {
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e < P * P) part[size_t(blockIdx.x) * P * P + e] = acc[q];
        }
    }
}

// G[e] = sum over workgroups b of part[b][e] in a fixed order: a workgroup
// takes 16 consecutive entries; its thread (u, e) adds partials b = u, u + 16,
// u + 32, ... in order (16 consecutive entries per load: coalesced), then
// the 16 sums of an entry are added in u order.
__global__ __launch_bounds__(kBlock) void k_eig_reduce(uint32_t nblk, uint32_t m,
                                                       const double* part, double* G) {
    __shared__ double lds[16][16];
    const uint32_t el = threadIdx.x % 16, u = threadIdx.x / 16;
    const uint32_t e = blockIdx.x * 16 + el;
    double s = 0.0;
    if (e < m) {
        uint32_t b = u;
        for (; b + 48 < nblk; b += 64) {
            const double v0 = part[size_t(b) * m + e], v1 = part[size_t(b + 16) * m + e];
            const double v2 = part[size_t(b + 32) * m + e], v3 = part[size_t(b + 48) * m + e];
            s = add_rn(add_rn(add_rn(add_rn(s, v0), v1), v2), v3);
        }
        for (; b < nblk; b += 16) s = add_rn(s, part[size_t(b) * m + e]);
    }
    lds[u][el] = s;
    __syncthreads();
    if (u == 0 && e < m) {
        double t = 0.0;
        for (int q = 0; q < 16; ++q) t = add_rn(t, lds[q][el]);
        G[e] = t;
    }
}

// X[r, j] = sum over i (in order) of Y[r, i] C[i][j]  (C row-major P x P).
// Lane j of a group of P lanes holds column j of C in registers; the
// workgroup stages a tile of Y rows in LDS (coalesced), and each group reads
// its row's Y[r, i] from LDS (one address for the group) and writes X[r, j]
// (coalesced).
template <int P>
__global__ __launch_bounds__(kBlock) void k_eig_combine(uint32_t n, const double* Y,
                                                        const double* C, double* X) {
    constexpr uint32_t kTile = 64;  // rows staged at a time
    __shared__ double ys[kTile * P];
    const uint32_t j = threadIdx.x % P, g = threadIdx.x / P;
    constexpr uint32_t groups = kBlock / P;
    double c[P];
#pragma unroll
    for (int i = 0; i < P; ++i) c[i] = C[i * P + j];
    for (uint32_t t0 = blockIdx.x * kTile; t0 < n; t0 += gridDim.x * kTile) {
        const uint32_t rows = (t0 + kTile < n) ? kTile : n - t0;
        __syncthreads();  // the previous tile has been consumed
        for (uint32_t e = threadIdx.x; e < rows * P; e += kBlock) ys[e] = Y[size_t(t0) * P + e];
        __syncthreads();
        for (uint32_t rr = g; rr < rows; rr += groups) {
            const double* y = ys + rr * P;
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < P; ++i) acc = add_rn(acc, mul_rn(y[i], c[i]));
            X[size_t(t0 + rr) * P + j] = acc;
        }
    }
}

// Per workgroup b and column j (rows in order): part[b][j] = sum (AX -
// theta_j M X)^2, part[b][P + j] = sum AX^2, part[b][2P + j] = sum (M X)^2.
// Thread t takes column t % P of rows r0 + t / P, r0 + t / P + kBlock / P,
// ... (coalesced); the kBlock / P threads of a column are then added in
// thread order.
template <int P>
__global__ __launch_bounds__(kBlock) void k_eig_residual(uint32_t n, const double* AX,
                                                         const double* X, const double* mass,
                                                         const double* theta, uint32_t rpb,
                                                         double* part) {
    __shared__ double lds[3][kBlock];
    const uint32_t r0 = blockIdx.x * rpb;
    const uint32_t r1 = (r0 + rpb < n) ? r0 + rpb : n;
    const uint32_t j = threadIdx.x % P;
    const double th = theta[j];
    double sr = 0.0, sa = 0.0, sm = 0.0;
    for (uint32_t r = r0 + threadIdx.x / P; r < r1; r += kBlock / P) {
        const size_t i = size_t(r) * P + j;
        const double ax = AX[i];
        const double mx = mass ? mul_rn(mass[r], X[i]) : X[i];
        const double d = add_rn(ax, -mul_rn(th, mx));
        sr = add_rn(sr, mul_rn(d, d));
        sa = add_rn(sa, mul_rn(ax, ax));
        sm = add_rn(sm, mul_rn(mx, mx));
    }
    lds[0][threadIdx.x] = sr;
    lds[1][threadIdx.x] = sa;
    lds[2][threadIdx.x] = sm;
    __syncthreads();
    if (threadIdx.x < 3 * P) {
        const uint32_t q = threadIdx.x / P, c = threadIdx.x % P;
        double t = 0.0;
        for (uint32_t u = c; u < kBlock; u += P) t = add_rn(t, lds[q][u]);
        part[size_t(blockIdx.x) * 3 * P + q * P + c] = t;
    }
}

// Gershgorin bound of the spectrum of M^-1 A: max over rows of
// sum |a_ij| / m_i (the maximum of non-negative doubles, taken on their bit
// patterns, which order the same way).
__global__ void k_eig_bound(uint32_t n, const uint32_t* rp, const double* val, const double* minv,
                            unsigned long long* out) {
    unsigned long long best = 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (uint32_t k = rp[r]; k < rp[r + 1]; ++k) s = add_rn(s, fabs(val[k]));
        if (minv) s = mul_rn(s, minv[r]);
        const unsigned long long b = __double_as_longlong(s);
        best = b > best ? b : best;
    }
    if (best) atomicMax(out, best);
}

__global__ void k_eig_inverse(uint32_t n, const double* m, double* minv) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
        minv[r] = 1.0 / m[r];
}

uint32_t blocks_for(uint64_t n) {
    const uint64_t g = (n + kBlock - 1) / kBlock;
    return uint32_t(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

// Rows per workgroup of the blocked reductions: whole tiles, at most 1024
// workgroups.
uint32_t rows_per_group(uint32_t n) {
    uint32_t rpb = (n + 1023) / 1024;
    rpb = (rpb + kGramTile - 1) / kGramTile * kGramTile;
    return rpb < kGramTile ? kGramTile : rpb;
}

}  // namespace

#define NLHIP_EIG_P(P, CALL)                                                        \
    switch (P) {                                                                    \
        case 8: { constexpr int PP = 8; CALL; } break;                              \
        case 16: { constexpr int PP = 16; CALL; } break;                            \
        case 32: { constexpr int PP = 32; CALL; } break;                            \
        default: throw Error(NLHIP_ERR_INVALID, "block size must be 8, 16 or 32"); \
    }

EigShape eig_shape(uint32_t n, uint32_t p, int cu) {
    // G = p lanes per row, passes of 64 rows; at most 16 workgroups per CU
    const uint32_t per = 256u / p;
    const uint64_t passes = (uint64_t(n) + per - 1) / per;
    const uint64_t cap = uint64_t(cu > 0 ? cu : 1) * 16;
    EigShape e;
    e.grid = uint32_t(passes < 1 ? 1 : (passes < cap ? passes : cap));
    uint64_t rpb = (uint64_t(n) + e.grid - 1) / e.grid;
    rpb = (rpb + kSpmmPass - 1) / kSpmmPass * kSpmmPass;
    e.rows_per_block = uint32_t(rpb < kSpmmPass ? kSpmmPass : rpb);
    return e;
}

void eig_spmm(const EigShape& e, uint32_t n, uint32_t p, const uint32_t* rp, const uint32_t* ci,
              const double* val, const double* Y, double* AY, hipStream_t s) {
    if (e.one_chunk) {
        NLHIP_EIG_P(p, (k_eig_spmm<PP, true><<<e.grid, kBlock, 0, s>>>(n, e.rows_per_block, rp, ci,
                                                                    val, Y, AY)));
    } else {
        NLHIP_EIG_P(p, (k_eig_spmm<PP, false><<<e.grid, kBlock, 0, s>>>(n, e.rows_per_block, rp,
                                                                     ci, val, Y, AY)));
    }
    NLHIP_CHECK(hipGetLastError());
}

void eig_cheb(const EigShape& e, uint32_t n, uint32_t p, const uint32_t* rp, const uint32_t* ci,
              const double* val, const double* minv, const double* Y, const double* Xold,
              double* Ynew, double c, double s1, double s2, hipStream_t s) {
    if (e.one_chunk) {
        NLHIP_EIG_P(p, (k_eig_cheb<PP, true><<<e.grid, kBlock, 0, s>>>(
                           n, e.rows_per_block, rp, ci, val, minv, Y, Xold, Ynew, c, s1, s2)));
    } else {
        NLHIP_EIG_P(p, (k_eig_cheb<PP, false><<<e.grid, kBlock, 0, s>>>(
                           n, e.rows_per_block, rp, ci, val, minv, Y, Xold, Ynew, c, s1, s2)));
    }
    NLHIP_CHECK(hipGetLastError());
}

void eig_gram(uint32_t n, uint32_t p, const double* Y, const double* Z, const double* mass,
              double* G, DevBuf<double>& part, hipStream_t s) {
    const uint32_t rpb = rows_per_group(n);
    const uint32_t nblk = (n + rpb - 1) / rpb;
    part.ensure(size_t(nblk) * p * p);
    NLHIP_EIG_P(p, (k_eig_gram<PP><<<nblk, kBlock, 0, s>>>(n, Y, Z, mass, rpb, part.get())));
    k_eig_reduce<<<(p * p + 15) / 16, kBlock, 0, s>>>(nblk, p * p, part.get(), G);
    NLHIP_CHECK(hipGetLastError());
}

void eig_combine(uint32_t n, uint32_t p, const double* Y, const double* C, double* X,
                 hipStream_t s) {
    const uint32_t tiles = (n + 63) / 64;  // k_eig_combine's tiles of 64 rows
    const uint32_t g = tiles < 2048 ? (tiles ? tiles : 1) : 2048;
    NLHIP_EIG_P(p, (k_eig_combine<PP><<<g, kBlock, 0, s>>>(n, Y, C, X)));
    NLHIP_CHECK(hipGetLastError());
}

void eig_residual(uint32_t n, uint32_t p, const double* AX, const double* X, const double* mass,
                  const double* theta, double* out, DevBuf<double>& part, hipStream_t s) {
    const uint32_t rpb = rows_per_group(n);
    const uint32_t nblk = (n + rpb - 1) / rpb;
    part.ensure(size_t(nblk) * 3 * p);
    NLHIP_EIG_P(p, (k_eig_residual<PP><<<nblk, kBlock, 0, s>>>(n, AX, X, mass, theta, rpb,
                                                                part.get())));
    k_eig_reduce<<<(3 * p + 15) / 16, kBlock, 0, s>>>(nblk, 3 * p, part.get(), out);
    NLHIP_CHECK(hipGetLastError());
}

double eig_bound(uint32_t n, const uint32_t* rp, const double* val, const double* minv,
                 hipStream_t s) {
    DevBuf<unsigned long long> out(1);
    NLHIP_CHECK(hipMemsetAsync(out.get(), 0, 8, s));
    k_eig_bound<<<blocks_for(n) < 4096 ? blocks_for(n) : 4096, kBlock, 0, s>>>(n, rp, val, minv,
                                                                              out.get());
    NLHIP_CHECK(hipGetLastError());
    unsigned long long h = 0;
    NLHIP_CHECK(hipMemcpyAsync(&h, out.get(), 8, hipMemcpyDeviceToHost, s));
    NLHIP_CHECK(hipStreamSynchronize(s));
    double b;
    std::memcpy(&b, &h, 8);
    return b;
}

void eig_inverse(uint32_t n, const double* m, double* minv, hipStream_t s) {
    k_eig_inverse<<<blocks_for(n) < 4096 ? blocks_for(n) : 4096, kBlock, 0, s>>>(n, m, minv);
    NLHIP_CHECK(hipGetLastError());
}

}  // namespace nlhip
