// Least-squares normal equations on the device: what a least-squares OpenNL
// context builds row by row (nlBegin/nlEnd(NL_ROW) with NL_LEAST_SQUARES,
// used by LSCM at src/lib/OGF/mesh/commands/mesh_grob_surface_commands.cpp:
// 724-779 and by the transport scripts through voxel_grob_transport_commands
// .cpp:236-279), with the same sums in the same order.
//
// Per row r (entries in order, variable v_e, coefficient a_e):
//   S_r = rhs_r - sum over locked entries of a_e x_{v_e}   (entry order)
//   for every ordered pair of free entries (e1, e2):  M[f(e1)][f(e2)] += a_e1 a_e2
//   for every free entry e:                           b[f(e)] += a_e S_r
// Rows are taken in order.  A dynamic row (OpenNL's NLRowColumn) stores the
// first contribution to an entry and adds the later ones; compression sorts
// each row by column.
//
// Device plan: every row emits its pair products keyed by (f1, f2) in
// (row, e1, e2) order; a stable radix sort by key groups the contributions to
// each entry in that order, and one thread per entry adds them left to right.
// The right-hand side sorts the free entries by f (stable, so entry order)
// and one thread per free variable adds a_e S_r in that order.  Separately
// rounded products and sums, as the host code.
#include <hipcub/hipcub.hpp>

#include "dev_common.h"
#include "nlhip_internal.h"

namespace nlhip {

namespace {

constexpr uint32_t kNone = 0xffffffffu;

__device__ __forceinline__ double sub_rn(double a, double b) { return __dsub_rn(a, b); }

// Input checks: row_ptr monotone and within the entry count, variables in
// range, coefficients, right-hand sides and values of locked variables
// finite.  Bits of *bad: 1 row_ptr, 2 variable, 4 non-finite value.
__global__ void k_ls_check(uint32_t nrows, uint32_t nvars, uint64_t nent, const uint32_t* rp,
                           const uint32_t* var, const double* coef, uint32_t nsys,
                           const double* row_rhs, const uint32_t* free_index,
                           const double* values, uint32_t* bad) {
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t flags = 0;
    if (values)
        for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nvars; v += stride)
            if (free_index[v] == kNone)
                for (uint32_t s = 0; s < nsys; ++s)
                    if (!isfinite(values[size_t(s) * nvars + v])) flags |= 4u;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += stride) {
        const uint32_t b = rp[r], e = rp[r + 1];
        if (b > e || e > nent) {
            flags |= 1u;
            continue;
        }
        for (uint32_t k = b; k < e; ++k) {
            if (var[k] >= nvars) flags |= 2u;
            if (!isfinite(coef[k])) flags |= 4u;
        }
        if (row_rhs)
            for (uint32_t s = 0; s < nsys; ++s)
                if (!isfinite(row_rhs[size_t(s) * nrows + r])) flags |= 4u;
    }
    if (flags) atomicOr(bad, flags);
}

// Per row: S (nsys values), the number of pair products it emits, and the
// row of each of its entries.
__global__ void k_ls_rows(uint32_t nrows, const uint32_t* rp, const uint32_t* var,
                          const double* coef, const uint32_t* free_index, const double* values,
                          uint32_t nvars, uint32_t nsys, const double* row_rhs, double* S,
                          uint64_t* npairs, uint32_t* row_of) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += stride) {
        const uint32_t b = rp[r], e = rp[r + 1];
        uint64_t nf = 0;
        for (uint32_t k = b; k < e; ++k) {
            row_of[k] = r;
            nf += free_index[var[k]] != kNone ? 1u : 0u;
        }
        npairs[r] = nf * nf;
        for (uint32_t s = 0; s < nsys; ++s) {
            double acc = row_rhs ? row_rhs[size_t(s) * nrows + r] : 0.0;
            for (uint32_t k = b; k < e; ++k) {
                const uint32_t v = var[k];
                if (free_index[v] == kNone)
                    acc = sub_rn(acc, mul_rn(coef[k], values ? values[size_t(s) * nvars + v] : 0.0));
            }
            S[size_t(s) * nrows + r] = acc;
        }
    }
}

// Pair products of row r at pair_off[r]: key f1 << bits | f2, value a1 a2,
// in (e1, e2) order.
__global__ void k_ls_pairs(uint32_t nrows, const uint32_t* rp, const uint32_t* var,
                           const double* coef, const uint32_t* free_index,
                           const uint64_t* pair_off, uint32_t bits, uint64_t* keys,
                           double* vals) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += stride) {
        const uint32_t b = rp[r], e = rp[r + 1];
        uint64_t o = pair_off[r];
        for (uint32_t k1 = b; k1 < e; ++k1) {
            const uint32_t f1 = free_index[var[k1]];
            if (f1 == kNone) continue;
            const double a1 = coef[k1];
            for (uint32_t k2 = b; k2 < e; ++k2) {
                const uint32_t f2 = free_index[var[k2]];
                if (f2 == kNone) continue;
                keys[o] = (uint64_t(f1) << bits) | f2;
                vals[o] = mul_rn(a1, coef[k2]);
                ++o;
            }
        }
    }
}

// Free entries keyed by their free index (locked ones by nfree, sorted last),
// valued by the entry index; and the count of entries per free variable.
__global__ void k_ls_entries(uint64_t nent, const uint32_t* var, const uint32_t* free_index,
                             uint32_t nfree, uint32_t* keys, uint32_t* vals, uint32_t* count) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nent; k += stride) {
        const uint32_t f = free_index[var[k]];
        keys[k] = f == kNone ? nfree : f;
        vals[k] = uint32_t(k);
        if (f != kNone) atomicAdd(&count[f], 1u);
    }
}

// 1 where a run of equal keys starts.
__global__ void k_ls_heads(uint64_t n, const uint64_t* keys, uint32_t* head) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_ls_starts(uint64_t n, const uint32_t* head, const uint32_t* seg,
                            uint64_t* start) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        if (head[i]) start[seg[i]] = i;
}

// One matrix entry per run: the first contribution, then the others added in
// order; its column, and a count for its row.
__global__ void k_ls_entry_sums(uint32_t nseg, uint64_t n, const uint64_t* start,
                                const uint64_t* keys, const double* vals, uint32_t bits,
                                uint32_t* colind, double* val, uint32_t* row_count) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint64_t mask = (uint64_t(1) << bits) - 1u;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += stride) {
        const uint64_t b = start[s], e = (s + 1 < nseg) ? start[s + 1] : n;
        double v = vals[b];
        for (uint64_t i = b + 1; i < e; ++i) v = add_rn(v, vals[i]);
        const uint64_t key = keys[b];
        colind[s] = uint32_t(key & mask);
        val[s] = v;
        atomicAdd(&row_count[uint32_t(key >> bits)], 1u);
    }
}

// b[s][f] = sum over the free entries of f (entry order) of a_e S[s][row(e)],
// starting from 0.
__global__ void k_ls_rhs(uint32_t nfree, const uint32_t* occ_off, const uint32_t* occ,
                         const double* coef, const uint32_t* row_of, const double* S,
                         uint32_t nrows, uint32_t nsys, double* rhs) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < nfree; f += stride) {
        const uint32_t b = occ_off[f], e = occ_off[f + 1];
        for (uint32_t s = 0; s < nsys; ++s) {
            double acc = 0.0;
            for (uint32_t i = b; i < e; ++i) {
                const uint32_t k = occ[i];
                acc = add_rn(acc, mul_rn(coef[k], S[size_t(s) * nrows + row_of[k]]));
            }
            rhs[size_t(s) * nfree + f] = acc;
        }
    }
}

// ---- LSCM rows on the device

// Triangle t's 12 LSCM coefficients (Levy, Petitjean, Ray, Maillot 2002):
// the conformality of the map in the triangle's orthonormal frame (X along
// p1 - p0, Y = N x X), W0 = z2 - z1, W1 = z0 - z2, W2 = z1 - z0, scaled by
// 1/sqrt(|e1 x e2|): real row (Wr, -Wi) then imaginary row (Wi, Wr) over
// (u0, v0, u1, v1, u2, v2).  ok[t] = 0 for a degenerate triangle.  Every
// operation is the host builder's (workloads.cpp, lscm_coeffs), in its
// order, correctly rounded and unfused, so the coefficients are its bits.
__global__ void k_lscm_coeffs(uint32_t nt, const double* P, const uint32_t* T, double* c,
                              uint32_t* ok) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride) {
        const uint32_t* q = T + 3 * size_t(t);
        const double* p0 = P + 3 * size_t(q[0]);
        const double* p1 = P + 3 * size_t(q[1]);
        const double* p2 = P + 3 * size_t(q[2]);
        const double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        const double nr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                              e1[0] * e2[1] - e1[1] * e2[0]};
        const double dT = sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
        const double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        double* out = c + 12 * size_t(t);
        if (!(dT > 0.0) || !(l1 > 0.0)) {
            ok[t] = 0;
            for (int j = 0; j < 12; ++j) out[j] = 0.0;
            continue;
        }
        const double X[3] = {e1[0] / l1, e1[1] / l1, e1[2] / l1};
        const double Nn[3] = {nr[0] / dT, nr[1] / dT, nr[2] / dT};
        const double Y[3] = {Nn[1] * X[2] - Nn[2] * X[1], Nn[2] * X[0] - Nn[0] * X[2],
                             Nn[0] * X[1] - Nn[1] * X[0]};
        const double z1x = l1, z1y = 0.0;
        const double z2x = e2[0] * X[0] + e2[1] * X[1] + e2[2] * X[2];
        const double z2y = e2[0] * Y[0] + e2[1] * Y[1] + e2[2] * Y[2];
        const double Wr[3] = {z2x - z1x, -z2x, z1x};
        const double Wi[3] = {z2y - z1y, -z2y, z1y};
        const double sq = sqrt(dT);
        for (int j = 0; j < 3; ++j) {
            const double cr = Wr[j] / sq, ci = Wi[j] / sq;
            out[2 * j] = cr;
            out[2 * j + 1] = -ci;
            out[6 + 2 * j] = ci;
            out[6 + 2 * j + 1] = cr;
        }
        ok[t] = 1;
    }
}

// Rows of the valid triangles in triangle order: triangle t at slot s (an
// exclusive count of the valid ones before it) gives rows 2s (real) and
// 2s + 1 (imaginary), 6 entries each.
__global__ void k_lscm_rows(uint32_t nt, const uint32_t* T, const double* c, const uint32_t* ok,
                            const uint32_t* slot, uint32_t* var, double* coef) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride) {
        if (!ok[t]) continue;
        const uint32_t* q = T + 3 * size_t(t);
        const size_t base = 12 * size_t(slot[t]);
        for (int e = 0; e < 12; ++e) {
            const int k = e % 6;
            var[base + e] = 2u * q[k / 2] + uint32_t(k & 1);
            coef[base + e] = c[12 * size_t(t) + e];
        }
    }
}

__global__ void k_lscm_check(uint32_t nv, uint32_t nt, const double* P, const uint32_t* T,
                             uint32_t* bad) {
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t flags = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride)
        for (int k = 0; k < 3; ++k)
            if (T[3 * size_t(t) + k] >= nv) flags |= 2u;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += stride)
        for (int k = 0; k < 3; ++k)
            if (!isfinite(P[3 * size_t(v) + k])) flags |= 4u;
    if (flags) atomicOr(bad, flags);
}

// *bad = 1 when a triangle's vertices lie in different charts (chart c owns
// vertices [voff[c], voff[c+1])).
__global__ void k_tri_chart_check(uint32_t nt, const uint32_t* T, const uint32_t* voff,
                                  uint32_t nc, uint32_t* bad) {
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t flag = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride) {
        uint32_t chart[3];
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = T[3 * size_t(t) + k];
            uint32_t lo = 0, hi = nc;  // last c with voff[c] <= v
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (voff[mid] <= v) lo = mid;
                else hi = mid;
            }
            chart[k] = lo;
        }
        flag |= (chart[0] != chart[1] || chart[0] != chart[2]) ? 1u : 0u;
    }
    if (flag) atomicOr(bad, 1u);
}

__global__ void k_row_ptr6(uint32_t nrows, uint32_t* rp) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r <= nrows; r += stride)
        rp[r] = 6u * r;
}

uint32_t grid_for(uint64_t n) {
    const uint64_t g = (n + kBlock - 1) / kBlock;
    return uint32_t(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

// A hipCUB call with its temporary storage (kept until the call has run).
template <typename F>
void with_temp(F&& call, hipStream_t s) {
    size_t bytes = 0;
    NLHIP_CHECK(call(nullptr, bytes));
    DevBuf<uint8_t> tmp(bytes);
    NLHIP_CHECK(call(tmp.get(), bytes));
    NLHIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

void lscm_rows_device(uint32_t nv, const double* P, uint32_t nt, const uint32_t* T,
                      LSRows& rows, hipStream_t s) {
    {
        DevBuf<uint32_t> bad(1);
        NLHIP_CHECK(hipMemsetAsync(bad.get(), 0, 4, s));
        k_lscm_check<<<grid_for(nt > nv ? nt : nv), kBlock, 0, s>>>(nv, nt, P, T, bad.get());
        NLHIP_CHECK(hipGetLastError());
        uint32_t h = 0;
        NLHIP_CHECK(hipMemcpyAsync(&h, bad.get(), 4, hipMemcpyDeviceToHost, s));
        NLHIP_CHECK(hipStreamSynchronize(s));
        if (h & 2u) throw Error(NLHIP_ERR_INVALID, "triangle vertex index out of range");
        if (h & 4u) throw Error(NLHIP_ERR_NONFINITE, "vertex coordinate is NaN or Inf");
    }
    DevBuf<double> c(12 * size_t(nt ? nt : 1));
    DevBuf<uint32_t> ok(size_t(nt) + 1), slot(size_t(nt) + 1);
    NLHIP_CHECK(hipMemsetAsync(ok.get(), 0, (size_t(nt) + 1) * 4, s));
    if (nt) k_lscm_coeffs<<<grid_for(nt), kBlock, 0, s>>>(nt, P, T, c.get(), ok.get());
    NLHIP_CHECK(hipGetLastError());
    with_temp([&](void* t, size_t& b) {
        return hipcub::DeviceScan::ExclusiveSum(t, b, ok.get(), slot.get(), size_t(nt) + 1, s);
    }, s);
    uint32_t nvalid = 0;
    NLHIP_CHECK(hipMemcpyAsync(&nvalid, slot.get() + nt, 4, hipMemcpyDeviceToHost, s));
    NLHIP_CHECK(hipStreamSynchronize(s));
    if (uint64_t(nvalid) * 12 > 0xffffffffull)
        throw Error(NLHIP_ERR_UNSUPPORTED, "more than 2^32 - 1 row coefficients");
    rows.nrows = 2 * nvalid;
    rows.rp.alloc(size_t(rows.nrows) + 1);
    rows.var.alloc(12 * size_t(nvalid ? nvalid : 1));
    rows.coef.alloc(12 * size_t(nvalid ? nvalid : 1));
    k_row_ptr6<<<grid_for(size_t(rows.nrows) + 1), kBlock, 0, s>>>(rows.nrows, rows.rp.get());
    if (nt)
        k_lscm_rows<<<grid_for(nt), kBlock, 0, s>>>(nt, T, c.get(), ok.get(), slot.get(),
                                                    rows.var.get(), rows.coef.get());
    NLHIP_CHECK(hipGetLastError());
    NLHIP_CHECK(hipStreamSynchronize(s));
}

void check_triangle_charts(uint32_t nt, const uint32_t* T, const uint32_t* voff, uint32_t nc,
                           hipStream_t s) {
    DevBuf<uint32_t> bad(1);
    NLHIP_CHECK(hipMemsetAsync(bad.get(), 0, 4, s));
    if (nt) k_tri_chart_check<<<grid_for(nt), kBlock, 0, s>>>(nt, T, voff, nc, bad.get());
    NLHIP_CHECK(hipGetLastError());
    uint32_t h = 0;
    NLHIP_CHECK(hipMemcpyAsync(&h, bad.get(), 4, hipMemcpyDeviceToHost, s));
    NLHIP_CHECK(hipStreamSynchronize(s));
    if (h) throw Error(NLHIP_ERR_INVALID, "a triangle spans two charts");
}

void ls_assemble_device(const LSInput& in, LSOutput& out, hipStream_t s) {
    const uint32_t nrows = in.nrows, nvars = in.nvars, nsys = in.nsys, nfree = in.nfree;
    const uint64_t nent = in.nent;
    // checks
    {
        DevBuf<uint32_t> bad(1);
        NLHIP_CHECK(hipMemsetAsync(bad.get(), 0, 4, s));
        k_ls_check<<<grid_for(nrows > nvars ? nrows : nvars), kBlock, 0, s>>>(
            nrows, nvars, nent, in.rp, in.var, in.coef, nsys, in.row_rhs, in.free_index,
            in.values, bad.get());
        NLHIP_CHECK(hipGetLastError());
        uint32_t h = 0;
        NLHIP_CHECK(hipMemcpyAsync(&h, bad.get(), 4, hipMemcpyDeviceToHost, s));
        NLHIP_CHECK(hipStreamSynchronize(s));
        if (h & 1u) throw Error(NLHIP_ERR_INVALID, "row_ptr is not monotone or exceeds the entries");
        if (h & 2u) throw Error(NLHIP_ERR_INVALID, "variable index out of range");
        if (h & 4u) throw Error(NLHIP_ERR_NONFINITE, "coefficient or right-hand side is NaN or Inf");
    }
    // rows: S, pair counts, row of each entry
    DevBuf<double> S(size_t(nsys) * (nrows ? nrows : 1));
    DevBuf<uint64_t> npairs(size_t(nrows) + 1);
    DevBuf<uint32_t> row_of(nent ? nent : 1);
    NLHIP_CHECK(hipMemsetAsync(npairs.get(), 0, (size_t(nrows) + 1) * 8, s));
    if (nrows)
        k_ls_rows<<<grid_for(nrows), kBlock, 0, s>>>(nrows, in.rp, in.var, in.coef,
                                                     in.free_index, in.values, nvars, nsys,
                                                     in.row_rhs, S.get(), npairs.get(),
                                                     row_of.get());
    NLHIP_CHECK(hipGetLastError());
    DevBuf<uint64_t> pair_off(size_t(nrows) + 1);
    with_temp([&](void* t, size_t& b) {
        return hipcub::DeviceScan::ExclusiveSum(t, b, npairs.get(), pair_off.get(),
                                                size_t(nrows) + 1, s);
    }, s);
    uint64_t npair = 0;
    NLHIP_CHECK(hipMemcpyAsync(&npair, pair_off.get() + nrows, 8, hipMemcpyDeviceToHost, s));
    NLHIP_CHECK(hipStreamSynchronize(s));
    npairs.release();
    if (npair > 0xffffffffull)
        throw Error(NLHIP_ERR_UNSUPPORTED,
                    "more than 2^32 - 1 coefficient products in one least-squares assembly");

    uint32_t bits = 1;
    while (bits < 32 && (uint64_t(1) << bits) < nfree) ++bits;

    // matrix: pair products, stable sort by (f1, f2), one sum per entry
    uint64_t nnz = 0;
    DevBuf<uint32_t> row_count(size_t(nfree) + 1);
    NLHIP_CHECK(hipMemsetAsync(row_count.get(), 0, (size_t(nfree) + 1) * 4, s));
    if (npair) {
        DevBuf<uint64_t> k0(npair), k1(npair);
        DevBuf<double> v0(npair), v1(npair);
        k_ls_pairs<<<grid_for(nrows), kBlock, 0, s>>>(nrows, in.rp, in.var, in.coef,
                                                      in.free_index, pair_off.get(), bits,
                                                      k0.get(), v0.get());
        NLHIP_CHECK(hipGetLastError());
        hipcub::DoubleBuffer<uint64_t> kb(k0.get(), k1.get());
        hipcub::DoubleBuffer<double> vb(v0.get(), v1.get());
        with_temp([&](void* t, size_t& b) {
            return hipcub::DeviceRadixSort::SortPairs(t, b, kb, vb, npair, 0, int(2 * bits), s);
        }, s);
        const uint64_t* keys = kb.Current();
        const double* vals = vb.Current();
        DevBuf<uint32_t> head(npair), seg(npair);
        k_ls_heads<<<grid_for(npair), kBlock, 0, s>>>(npair, keys, head.get());
        NLHIP_CHECK(hipGetLastError());
        with_temp([&](void* t, size_t& b) {
            return hipcub::DeviceScan::ExclusiveSum(t, b, head.get(), seg.get(), npair, s);
        }, s);
        uint32_t last_seg = 0, last_head = 0;
        NLHIP_CHECK(hipMemcpyAsync(&last_seg, seg.get() + npair - 1, 4, hipMemcpyDeviceToHost, s));
        NLHIP_CHECK(hipMemcpyAsync(&last_head, head.get() + npair - 1, 4, hipMemcpyDeviceToHost, s));
        NLHIP_CHECK(hipStreamSynchronize(s));
        nnz = uint64_t(last_seg) + last_head;
        if (nnz > 0xffffffffull)
            throw Error(NLHIP_ERR_UNSUPPORTED, "normal equations with 2^32 or more nonzeros");
        DevBuf<uint64_t> start(nnz);
        k_ls_starts<<<grid_for(npair), kBlock, 0, s>>>(npair, head.get(), seg.get(), start.get());
        NLHIP_CHECK(hipGetLastError());
        NLHIP_CHECK(hipStreamSynchronize(s));
        head.release();
        seg.release();
        out.colind.alloc(nnz);
        out.val.alloc(nnz);
        k_ls_entry_sums<<<grid_for(nnz), kBlock, 0, s>>>(uint32_t(nnz), npair, start.get(), keys,
                                                         vals, bits, out.colind.get(),
                                                         out.val.get(), row_count.get());
        NLHIP_CHECK(hipGetLastError());
        NLHIP_CHECK(hipStreamSynchronize(s));  // before the sort buffers go
    } else {
        out.colind.alloc(1);
        out.val.alloc(1);
    }
    out.rowptr.alloc(size_t(nfree) + 1);
    with_temp([&](void* t, size_t& b) {
        return hipcub::DeviceScan::ExclusiveSum(t, b, row_count.get(), out.rowptr.get(),
                                                size_t(nfree) + 1, s);
    }, s);
    out.nnz = nnz;

    // right-hand sides: free entries sorted by free index (stable)
    DevBuf<uint32_t> occ_count(size_t(nfree) + 1), occ_off(size_t(nfree) + 1);
    NLHIP_CHECK(hipMemsetAsync(occ_count.get(), 0, (size_t(nfree) + 1) * 4, s));
    DevBuf<uint32_t> ek0(nent ? nent : 1), ek1(nent ? nent : 1), ev0(nent ? nent : 1),
        ev1(nent ? nent : 1);
    const uint32_t* occ = ev0.get();
    if (nent) {
        k_ls_entries<<<grid_for(nent), kBlock, 0, s>>>(nent, in.var, in.free_index, nfree,
                                                       ek0.get(), ev0.get(), occ_count.get());
        NLHIP_CHECK(hipGetLastError());
        uint32_t ebits = 1;
        while (ebits < 32 && (uint64_t(1) << ebits) <= nfree) ++ebits;
        hipcub::DoubleBuffer<uint32_t> kb(ek0.get(), ek1.get()), vb(ev0.get(), ev1.get());
        with_temp([&](void* t, size_t& b) {
            return hipcub::DeviceRadixSort::SortPairs(t, b, kb, vb, nent, 0, int(ebits), s);
        }, s);
        occ = vb.Current();
    }
    with_temp([&](void* t, size_t& b) {
        return hipcub::DeviceScan::ExclusiveSum(t, b, occ_count.get(), occ_off.get(),
                                                size_t(nfree) + 1, s);
    }, s);
    out.rhs.alloc(size_t(nsys) * nfree);
    k_ls_rhs<<<grid_for(nfree), kBlock, 0, s>>>(nfree, occ_off.get(), occ, in.coef,
                                                row_of.get(), S.get(), nrows, nsys,
                                                out.rhs.get());
    NLHIP_CHECK(hipGetLastError());
    NLHIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace nlhip
