// SSOR preconditioner objects (nlNewSSORPreconditioner): level schedules of
// the two triangular sweeps, built once on the host from the matrix
// pattern, and the sweeps of one application captured as a graph.  The
// arithmetic is in k_ssor.hip.
#include <algorithm>
#include <cmath>
#include <memory>

#include "nlhip_internal.h"

namespace nlhip {

namespace {

// Levels of at most this many rows run back to back in one workgroup.
constexpr uint32_t kChainRows = 256;  // one row per thread of a 256-thread workgroup

// level(i) = 1 + the deepest level among the rows i reads (j < i for the
// forward sweep, j > i for the backward one); rows counting-sorted by level
// (increasing row index within a level).
void level_schedule(uint32_t n, const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci,
                    bool upper, std::vector<uint32_t>& rows, std::vector<uint32_t>& off) {
    std::vector<uint32_t> lev(n, 0);
    uint32_t top = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t i = upper ? n - 1 - k : k;
        uint32_t l = 0;
        for (uint32_t e = rp[i]; e < rp[i + 1]; ++e) {
            const uint32_t j = ci[e];
            if (upper ? j > i : j < i) l = std::max(l, lev[j] + 1);
        }
        lev[i] = l;
        top = std::max(top, l);
    }
    off.assign(size_t(top) + 2, 0);
    for (uint32_t i = 0; i < n; ++i) ++off[lev[i] + 1];
    for (size_t l = 1; l < off.size(); ++l) off[l] += off[l - 1];
    rows.resize(n);
    std::vector<uint32_t> next(off.begin(), off.end() - 1);
    for (uint32_t i = 0; i < n; ++i) rows[next[lev[i]]++] = i;
}

}  // namespace

// One sweep on the device: rows by level and their packed entries.
struct SweepBufs {
    DevBuf<uint32_t> rows, off, eoff, col;
    DevBuf<double> val;
    SsorSweepArgs args() const {
        return SsorSweepArgs{rows.get(), off.get(), eoff.get(), col.get(), val.get()};
    }
};

class Ssor {
  public:
    struct Step {
        uint32_t l0, l1, rows;
    };
    uint32_t n = 0;
    double omega = 1.5;
    SweepBufs lo, up;
    DevBuf<double> diag, dscale, x, w, y;
    std::vector<Step> lo_steps, up_steps;
    Stream stream;
    Graph graph;
    std::mutex mu;

    SsorArgs args() const {
        SsorArgs a{};
        a.lo = lo.args();
        a.up = up.args();
        a.diag = diag.get();
        a.dscale = dscale.get();
        a.x = x.get();
        a.w = w.get();
        a.y = y.get();
        a.omega = omega;
        a.post = 2.0 - omega;
        a.n = n;
        return a;
    }
};

namespace {

// One launch per level with more than kChainRows rows, one per run of
// consecutive smaller levels.
std::vector<Ssor::Step> plan(const std::vector<uint32_t>& off) {
    std::vector<Ssor::Step> steps;
    const uint32_t nl = uint32_t(off.size() - 1);
    uint32_t l = 0;
    while (l < nl) {
        const uint32_t sz = off[l + 1] - off[l];
        if (sz > kChainRows) {
            steps.push_back({l, l + 1, sz});
            ++l;
            continue;
        }
        uint32_t e = l;
        while (e < nl && off[e + 1] - off[e] <= kChainRows) ++e;
        steps.push_back({l, e, off[e] - off[l]});
        l = e;
    }
    return steps;
}

}  // namespace

namespace {

// Level schedule of one sweep, its positions' packed entries, and the
// launch plan; uploaded on stream s.
std::vector<Ssor::Step> build_sweep(uint32_t n, const std::vector<uint32_t>& rp,
                                    const std::vector<uint32_t>& ci, const std::vector<double>& va,
                                    bool upper, SweepBufs& B, hipStream_t s) {
    std::vector<uint32_t> rows, off;
    level_schedule(n, rp, ci, upper, rows, off);
    std::vector<uint32_t> eoff(size_t(n) + 1, 0), col;
    std::vector<double> val;
    col.reserve(ci.size() / 2 + 1);
    val.reserve(ci.size() / 2 + 1);
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t i = rows[k];
        for (uint32_t e = rp[i]; e < rp[i + 1]; ++e)
            if (upper ? ci[e] > i : ci[e] < i) {
                col.push_back(ci[e]);
                val.push_back(va[e]);
            }
        eoff[k + 1] = uint32_t(col.size());
    }
    B.rows.alloc(std::max<size_t>(rows.size(), 1));
    B.off.alloc(off.size());
    B.eoff.alloc(eoff.size());
    B.col.alloc(std::max<size_t>(col.size(), 1));
    B.val.alloc(std::max<size_t>(val.size(), 1));
    copy_in(B.rows.get(), rows.data(), rows.size() * 4, s);
    copy_in(B.off.get(), off.data(), off.size() * 4, s);
    copy_in(B.eoff.get(), eoff.data(), eoff.size() * 4, s);
    copy_in(B.col.get(), col.data(), col.size() * 4, s);
    copy_in(B.val.get(), val.data(), val.size() * 8, s);
    NLHIP_CHECK(hipStreamSynchronize(s));  // the host arrays go out of scope
    return plan(off);
}

}  // namespace

Ssor* ssor_create(System& S, double omega) {
    if (!std::isfinite(omega) || omega == 0.0)
        throw Error(NLHIP_ERR_INVALID, "SSOR omega must be finite and nonzero");
    std::lock_guard<std::mutex> lock(S.mu);
    auto P = std::make_unique<Ssor>();
    const uint32_t n = S.n;
    const uint64_t nnz = S.nnz;
    P->n = n;
    P->omega = omega;
    hipStream_t s = P->stream.get();
    // the matrix on the host (a copy: the preconditioner outlives changes to M)
    std::vector<uint32_t> rp(size_t(n) + 1), ci(nnz);
    std::vector<double> va(nnz);
    NLHIP_CHECK(hipMemcpyAsync(rp.data(), S.rowptr.get(), rp.size() * 4, hipMemcpyDeviceToHost, s));
    if (nnz) {
        NLHIP_CHECK(hipMemcpyAsync(ci.data(), S.colind.get(), nnz * 4, hipMemcpyDeviceToHost, s));
        NLHIP_CHECK(hipMemcpyAsync(va.data(), S.val.get(), nnz * 8, hipMemcpyDeviceToHost, s));
    }
    P->diag.alloc(n);
    P->dscale.alloc(n);
    {
        DevBuf<uint32_t> bad(1);
        NLHIP_CHECK(hipMemsetAsync(bad.get(), 0, 4, s));
        NLHIP_CHECK(ssor_diag(n, S.rowptr.get(), S.colind.get(), S.val.get(), omega,
                              P->diag.get(), P->dscale.get(), bad.get(), s));
        uint32_t h = 0;
        NLHIP_CHECK(hipMemcpyAsync(&h, bad.get(), 4, hipMemcpyDeviceToHost, s));
        P->stream.sync();
        if (h) throw Error(NLHIP_ERR_INVALID, "SSOR needs a nonzero, finite diagonal entry in every row");
    }
    P->lo_steps = build_sweep(n, rp, ci, va, false, P->lo, s);
    P->up_steps = build_sweep(n, rp, ci, va, true, P->up, s);
    P->x.alloc(n);
    P->w.alloc(n);
    P->y.alloc(n);
    const SsorArgs a = P->args();
    P->graph.capture(s, [&] {
        for (const Ssor::Step& st : P->lo_steps)
            NLHIP_CHECK(ssor_sweep(a, false, st.l0, st.l1, st.rows, s));
        for (const Ssor::Step& st : P->up_steps)
            NLHIP_CHECK(ssor_sweep(a, true, st.l0, st.l1, st.rows, s));
        NLHIP_CHECK(ssor_scale(a, s));
    });
    return P.release();
}

void ssor_destroy(Ssor* P) { delete P; }

uint32_t ssor_n(const Ssor& P) { return P.n; }

void ssor_apply(Ssor& P, const double* x, double* y, hipStream_t s) {
    std::lock_guard<std::mutex> lock(P.mu);
    const size_t bytes = size_t(P.n) * 8;
    NLHIP_CHECK(hipMemcpyAsync(P.x.get(), x, bytes, hipMemcpyDeviceToDevice, s));
    P.graph.launch(s);
    NLHIP_CHECK(hipMemcpyAsync(y, P.y.get(), bytes, hipMemcpyDeviceToDevice, s));
    // the next application (maybe on another stream) reuses x, w, y
    NLHIP_CHECK(hipStreamSynchronize(s));
}

void precond_apply(const Precond& P, uint32_t n, const double* in, double* out, hipStream_t s) {
    if (P.ssor) {
        if (ssor_n(*P.ssor) != n) throw Error(NLHIP_ERR_INVALID, "preconditioner size mismatch");
        ssor_apply(*P.ssor, in, out, s);
    } else if (P.dinv) {
        NLHIP_CHECK(blas_dmul(n, P.dinv, in, out, s));
    } else {
        NLHIP_CHECK(hipMemcpyAsync(out, in, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace nlhip
