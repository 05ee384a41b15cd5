// Resident systems and the solve drivers: FP64 Jacobi-PCG, FP32-inner /
// FP64-outer iterative refinement, device BiCGSTAB, SpMV, and the C ABI of
// the handle-based API (include/nl_hip.h, part 2).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>

#include "nlhip_internal.h"

namespace nlhip {

// ----------------------------------------------------------------- launch

// Stream engine (code 64: the 2x2 block engine on block-form matrices of at
// least kBlockMinRows rows, the scalar one otherwise); rows per workgroup
// chosen by size (measured on LSCM systems of 300 to 10 M rows: small,
// latency-bound systems run fastest with more, emptier workgroups, large
// ones with fewer, fuller ones), at most 4 workgroups per CU (which also
// bounds the number of per-workgroup partial sums every CG launch reduces).
void default_launch(uint32_t n, uint64_t nnz, int& rpw, uint32_t& grid) {
    (void)nnz;
    rpw = 64;
    const uint64_t per = n < 32768u ? 128 : (n < kBlockMinRows ? 256 : 512);
    const uint64_t g = (uint64_t(n) + per - 1) / per;
    uint32_t cap = 1024;
    if (device_usable()) cap = uint32_t(std::max(4, device().cu_count * 4));
    grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(g, cap)));
}

// ----------------------------------------------------------------- system

static void validate_crs(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* val) {
    if (rp[0] != 0) throw Error(NLHIP_ERR_INVALID, "rowptr[0] != 0");
    for (uint32_t i = 0; i < n; ++i)
        if (rp[i + 1] < rp[i])
            throw Error(NLHIP_ERR_INVALID, "rowptr is not monotone at row " + std::to_string(i));
    const uint64_t nnz = rp[n];
    int bad_col = 0, bad_val = 0;
#pragma omp parallel for reduction(| : bad_col, bad_val) schedule(static)
    for (int64_t j = 0; j < int64_t(nnz); ++j) {
        bad_col |= (ci[j] >= n);
        bad_val |= !std::isfinite(val[j]);
    }
    if (bad_col) throw Error(NLHIP_ERR_INVALID, "colind out of range");
    if (bad_val) throw Error(NLHIP_ERR_NONFINITE, "matrix value is NaN or Inf");
}

// 16-bit column offsets for the stream engine: kept when every 1024-entry
// segment of the nonzero stream spans at most 65535 columns (banded matrices:
// grids in scan order, LSCM and Laplacian systems of meshes with local vertex
// order); the 32-bit indices stay for the other engines.  Returns false (and
// leaves ci16, segb empty) when they do not fit.
bool build_col_offsets(uint64_t nnz, const uint32_t* colind, hipStream_t s,
                       DevBuf<uint16_t>& ci16_out, DevBuf<uint32_t>& segb_out) {
    if (nnz == 0) return false;
    const uint64_t nseg = (nnz + kSegEntries - 1) >> kSegShift;
    DevBuf<uint32_t> segb(nseg + 1), span(1);
    NLHIP_CHECK(fill<int32_t>(1, 0, reinterpret_cast<int32_t*>(span.get()), s));
    NLHIP_CHECK(seg_bases(nnz, colind, segb.get(), span.get(), s));
    uint32_t h_span = 0;
    NLHIP_CHECK(hipMemcpyAsync(&h_span, span.get(), 4, hipMemcpyDeviceToHost, s));
    NLHIP_CHECK(hipStreamSynchronize(s));
    if (h_span > 0xffffu) return false;
    NLHIP_CHECK(hipMemcpyAsync(segb.get() + nseg, segb.get() + nseg - 1, 4,
                               hipMemcpyDeviceToDevice, s));
    DevBuf<uint16_t> ci16(nnz);
    NLHIP_CHECK(seg_encode(nnz, colind, segb.get(), ci16.get(), s));
    NLHIP_CHECK(hipStreamSynchronize(s));
    ci16_out = std::move(ci16);
    segb_out = std::move(segb);
    return true;
}

bool build_block_form(uint32_t n, uint64_t nnz, const uint32_t* rowptr, const uint32_t* colind,
                      const double* val, hipStream_t s, DevBuf<uint32_t>& brp_out,
                      DevBuf<uint16_t>& bci16_out, DevBuf<uint32_t>& bsegb_out,
                      DevBuf<uint32_t>& bcol_out, DevBuf<double>& bval_out) {
    if (n < 2 || (n & 1u) || nnz == 0 || (nnz & 3u)) return false;
    const uint32_t nb = n / 2;
    const uint64_t nblk = nnz / 4;
    DevBuf<uint32_t> bad(1);
    NLHIP_CHECK(fill<int32_t>(1, 0, reinterpret_cast<int32_t*>(bad.get()), s));
    NLHIP_CHECK(block_check(nb, rowptr, colind, bad.get(), s));
    uint32_t h_bad = 1;
    NLHIP_CHECK(hipMemcpyAsync(&h_bad, bad.get(), 4, hipMemcpyDeviceToHost, s));
    NLHIP_CHECK(hipStreamSynchronize(s));
    if (h_bad) return false;
    DevBuf<uint32_t> brp(size_t(nb) + 1), bcol(nblk);
    DevBuf<double> bval(block_vals_size(nblk));
    NLHIP_CHECK(block_pack<double>(nb, rowptr, colind, val, brp.get(), bcol.get(), bval.get(), s));
    DevBuf<uint16_t> bci16;
    DevBuf<uint32_t> bsegb;
    // block columns as 16-bit offsets where every segment's span allows it,
    // else the 32-bit block columns (a matrix whose vertex numbering is not
    // banded, e.g. a mesh in file order)
    if (build_col_offsets(nblk, bcol.get(), s, bci16, bsegb)) bcol.release();
    brp_out = std::move(brp);
    bval_out = std::move(bval);
    bci16_out = std::move(bci16);
    bsegb_out = std::move(bsegb);
    bcol_out = std::move(bcol);
    return true;
}

bool build_blocks(System& S) {
    return build_block_form(S.n, S.nnz, S.rowptr.get(), S.colind.get(), S.val.get(),
                            S.stream.get(), S.brp, S.bci16, S.bsegb, S.bcol, S.bval);
}

namespace {

// Jacobi diagonal, column offsets, block form, launch shape and FP32 copy of
// a system whose CRS is resident.
void finish_system(System& S, uint32_t flags) {
    hipStream_t s = S.stream.get();
    S.dinv.alloc(S.n);
    NLHIP_CHECK(jacobi_dinv<double>(S.n, S.rowptr.get(), S.colind.get(), S.val.get(),
                                    S.dinv.get(), s));
    if (!(flags & NLHIP_SYSTEM_NO_OFFSETS)) {
        build_col_offsets(S.nnz, S.colind.get(), s, S.ci16, S.segb);
        if (!(flags & NLHIP_SYSTEM_NO_BLOCKS)) build_blocks(S);
    }
    default_launch(S.n, S.nnz, S.rpw, S.grid);
    if (flags & NLHIP_SYSTEM_FP32) S.ensure_fp32();
    S.stream.sync();
}

}  // namespace

System* system_create(uint32_t n, const uint32_t* rowptr, const uint32_t* colind,
                      const double* val, uint32_t flags) {
    if (n == 0 || !rowptr) throw Error(NLHIP_ERR_INVALID, "empty system");
    device();
    // Validation runs on host copies (a device caller's arrays are fetched).
    std::vector<uint32_t> h_rp, h_ci;
    std::vector<double> h_val;
    const uint32_t* rp = rowptr;
    const uint32_t* ci = colind;
    const double* va = val;
    auto fetch = [](void* dst, const void* src, size_t bytes) {
        Stream st;
        NLHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, st.get()));
        st.sync();
    };
    if (is_device_pointer(rowptr)) {
        h_rp.resize(size_t(n) + 1);
        fetch(h_rp.data(), rowptr, (size_t(n) + 1) * 4);
        rp = h_rp.data();
    }
    const uint64_t nnz = rp[n];
    if (nnz > 0 && (!colind || !val)) throw Error(NLHIP_ERR_INVALID, "null colind/val");
    if (nnz && is_device_pointer(colind)) {
        h_ci.resize(nnz);
        fetch(h_ci.data(), colind, nnz * 4);
        ci = h_ci.data();
    }
    if (nnz && is_device_pointer(val)) {
        h_val.resize(nnz);
        fetch(h_val.data(), val, nnz * 8);
        va = h_val.data();
    }
    validate_crs(n, rp, ci, va);
    auto S = std::make_unique<System>();
    S->n = n;
    S->nnz = nnz;
    S->rowptr.alloc(size_t(n) + 1);
    S->colind.alloc(nnz);
    S->val.alloc(nnz);
    hipStream_t s = S->stream.get();
    copy_in(S->rowptr.get(), rp, (size_t(n) + 1) * 4, s);
    copy_in(S->colind.get(), ci, nnz * 4, s);
    copy_in(S->val.get(), va, nnz * 8, s);
    finish_system(*S, flags);
    return S.release();
}

System* system_adopt(uint32_t n, uint64_t nnz, DevBuf<uint32_t>&& rowptr,
                     DevBuf<uint32_t>&& colind, DevBuf<double>&& val, uint32_t flags) {
    if (n == 0) throw Error(NLHIP_ERR_INVALID, "empty system");
    auto S = std::make_unique<System>();
    S->n = n;
    S->nnz = nnz;
    S->rowptr = std::move(rowptr);
    S->colind = std::move(colind);
    S->val = std::move(val);
    finish_system(*S, flags);
    return S.release();
}

void System::ensure_fp32() {
    if (val32.get() && val32.size() >= std::max<uint64_t>(nnz, 1)) return;
    val32.alloc(nnz);
    dinv32.alloc(n);
    hipStream_t s = stream.get();
    NLHIP_CHECK(cast_f64_f32(nnz, val.get(), val32.get(), s));
    NLHIP_CHECK(jacobi_dinv<float>(n, rowptr.get(), colind.get(), val.get(), dinv32.get(), s));
    if (blocks()) {
        bval32.alloc(block_vals_size(nnz / 4));
        NLHIP_CHECK(block_pack<float>(n / 2, rowptr.get(), colind.get(), val.get(), brp.get(),
                                      nullptr, bval32.get(), s));
    }
}

uint64_t System::device_bytes() const {
    return rowptr.bytes() + colind.bytes() + ci16.bytes() + segb.bytes() + val.bytes() +
           brp.bytes() + bsegb.bytes() + bci16.bytes() + bcol.bytes() + bval.bytes() +
           bval32.bytes() +
           dinv.bytes() + val32.bytes() +
           dinv32.bytes() + ws64.x.bytes() * 7 + ws32.x.bytes() * 7 + scratch.bytes();
}

template <typename T>
void CGWorkspace<T>::ensure(uint32_t n_, uint32_t P_) {
    if (n == n_ && P == P_ && x.get()) return;
    graph.reset();
    n = n_;
    P = P_;
    x.alloc(n);
    b.alloc(n);
    r.alloc(n);
    h.alloc(n);
    d.alloc(n);
    q.alloc(n);
    part.alloc(size_t(kSpmvGridMul + 3) * P);  // d.q (up to kSpmvGridMul P), r.h, r.r, b.b
    st.alloc(1);
    st_host.ensure(2);
}
template struct CGWorkspace<double>;
template struct CGWorkspace<float>;

template <typename T>
void fill_cg_args(const System& S, const CGBuffers<T>& w, const T* dinv, CGArgs<T>& a,
                  int& code) {
    const SpmvOps<T> o = S.ops<T>();
    code = o.code;
    a.m = S.n;
    a.P = S.grid;
    a.chunk = uint32_t((uint64_t(S.n) + S.grid - 1) / S.grid);
    a.sp = spmv_work_per_block(S.n, S.grid, o.code);
    // the streamed SpMV step runs with up to kSpmvGridMul times the
    // workgroups (at least four wavefront tiles each)
    const uint64_t tile_rows = o.code >= 73 ? 128 : 64;
    const uint64_t quads = (uint64_t(S.n) + 4 * tile_rows - 1) / (4 * tile_rows);
    a.Ps = o.code >= 64 ? uint32_t(std::max<uint64_t>(
                              S.grid, std::min<uint64_t>(uint64_t(kSpmvGridMul) * S.grid, quads)))
                        : S.grid;
    a.sps = spmv_work_per_block(S.n, a.Ps, o.code);
    a.rp = o.rp;
    a.cols = o.cols;
    a.val = o.val;
    a.dinv = dinv;
    a.x = w.x;
    a.b = w.b;
    a.r = w.r;
    a.h = w.h;
    a.d = w.d;
    a.q = w.q;
    a.part_dq = w.part;  // Ps <= kSpmvGridMul P entries
    a.part_rh = w.part + size_t(kSpmvGridMul) * S.grid;
    a.part_rr = w.part + size_t(kSpmvGridMul + 1) * S.grid;
    a.part_bb = w.part + size_t(kSpmvGridMul + 2) * S.grid;
    a.st = w.st;
}
template void fill_cg_args<double>(const System&, const CGBuffers<double>&, const double*,
                                   CGArgs<double>&, int&);

template <typename T>
static void setup_args(System& S, CGWorkspace<T>& ws, const T* dinv) {
    const CGBuffers<T> w{ws.x.get(), ws.b.get(), ws.r.get(), ws.h.get(),
                         ws.d.get(), ws.q.get(), ws.part.get(), ws.st.get()};
    fill_cg_args<T>(S, w, dinv, ws.args, ws.code);
}

uint32_t auto_check_every(const System& S) {
    // about 2 ms of device work per captured graph, within [8, 128]
    const double bytes_per_iter = 12.0 * double(S.nnz) + 100.0 * double(S.n);
    const double iter_s = std::max(bytes_per_iter / 4.0e12, 6.0e-6);
    return uint32_t(std::clamp(2.0e-3 / iter_s, 8.0, 128.0));
}

// Runs the captured iterations of a started CG (arguments set, status open)
// until the device-side loop has stopped; `budget` bounds the iterations it
// can still run.  Returns the final status.
template <typename T>
static CGStatus<T> run_ladder(System& S, CGWorkspace<T>& ws, bool precond, uint32_t K,
                              uint64_t budget) {
    hipStream_t s = S.stream.get();
    if (ws.graph_rpw != ws.code || ws.graph_precond != precond ||
        std::memcmp(&ws.graph_args, &ws.args, sizeof(CGArgs<T>)) != 0)
        ws.graph.reset();
    ws.graph.ensure(K, s, [&](uint32_t k) {
        for (uint32_t i = 0; i < k; ++i) NLHIP_CHECK(cg_step<T>(ws.args, ws.code, precond, s));
    });
    ws.graph_rpw = ws.code;
    ws.graph_precond = precond;
    ws.graph_args = ws.args;
    Event ev[2];
    CGStatus<T>* host = ws.st_host.get();
    host[0].done = host[1].done = 0;
    // the device stops by itself at its cap; past that the status must say so
    const uint64_t limit = budget + 3ull * ws.graph.cap() + 8;
    uint64_t launched = 0, covered = 0;
    for (;;) {
        ws.graph.launch(launched, s);
        covered += ws.graph.iters(launched);
        const int slot = int(launched & 1);
        NLHIP_CHECK(hipMemcpyAsync(&host[slot], ws.st.get(), sizeof(CGStatus<T>),
                                   hipMemcpyDeviceToHost, s));
        ev[slot].record(s);
        ++launched;
        if (launched >= 2) {
            const int prev = int((launched - 2) & 1);
            ev[prev].sync();
            if (host[prev].done) break;
        }
        if (covered > limit) {
            S.stream.sync();
            if (host[slot].done) break;
            throw Error(NLHIP_ERR_HIP, "CG did not reach its stopping test");
        }
    }
    CGStatus<T> fin;
    NLHIP_CHECK(hipMemcpyAsync(&fin, ws.st.get(), sizeof(fin), hipMemcpyDeviceToHost, s));
    S.stream.sync();
    return fin;
}

// Re-opens a stopped CG loop: stop and done cleared, a new threshold and
// cap, and alpha = 0, because the direction launch that stopped has already
// applied the pending x -= alpha d.  The next direction launch recomputes
// rho and err from the same partials and, the test failing now, goes on
// exactly as it would have with this threshold and cap from the start
// (x -= 0 d leaves x as it is, up to the sign of a zero).
template <typename T>
static void cg_reopen(System& S, CGWorkspace<T>& ws, CGStatus<T> st, T thresh, int max_iter) {
    st.stop = 0;
    st.done = 0;
    st.alpha = T(0);
    st.thresh = thresh;
    st.max_iter = max_iter;
    NLHIP_CHECK(hipMemcpyAsync(ws.st.get(), &st, sizeof(st), hipMemcpyHostToDevice,
                               S.stream.get()));
    S.stream.sync();  // st lives on this stack
}

template <typename T>
static CGStatus<T> read_status(System& S, CGWorkspace<T>& ws) {
    CGStatus<T> st;
    NLHIP_CHECK(hipMemcpyAsync(&st, ws.st.get(), sizeof(st), hipMemcpyDeviceToHost,
                               S.stream.get()));
    S.stream.sync();
    return st;
}

// Reads |b - A x|^2 and |b|^2 (FP64) into rr, bb.
static void residual_norms(System& S, const double* b, const double* x, double* r_out,
                           double& rr, double& bb) {
    hipStream_t s = S.stream.get();
    const uint32_t P = S.grid;
    S.scratch.ensure(size_t(2) * P + 4);
    double* part = S.scratch.get();
    double* out = part + 2 * size_t(P);
    const SpmvOps<double> o = S.ops<double>();
    NLHIP_CHECK(residual64(S.n, o.rp, o.cols, o.val, x, b, r_out, part, part + P, o.code, P, s));
    NLHIP_CHECK(reduce_final(part, P, out, s));
    NLHIP_CHECK(reduce_final(part + P, P, out + 2, s));
    S.host_scalars.ensure(4);
    NLHIP_CHECK(hipMemcpyAsync(S.host_scalars.get(), out, 4 * sizeof(double),
                               hipMemcpyDeviceToHost, s));
    S.stream.sync();
    rr = S.host_scalars.get()[0];
    bb = S.host_scalars.get()[2];
}

double true_rel_residual(System& S, const double* b, const double* x) {
    double rr = 0, bb = 0;
    residual_norms(S, b, x, nullptr, rr, bb);
    return bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

// The ways of running an FP64 PCG on the system's ws64: captured graphs
// replayed with status polls into pinned memory.
struct GraphPcgOps final : PcgOps {
    System& S;
    explicit GraphPcgOps(System& s) : S(s) {}
    CGStatus<double> start(double eps) override {
        CGWorkspace<double>& ws = S.ws64;
        NLHIP_CHECK(cg_start<double>(ws.args, ws.code, S.pcg.precond, eps, 0, S.stream.get()));
        return read_status(S, ws);
    }
    void reopen(const CGStatus<double>& cur, double thresh, int cap) override {
        cg_reopen<double>(S, S.ws64, cur, thresh, cap);
    }
    CGStatus<double> run(uint32_t K, uint64_t budget) override {
        return run_ladder<double>(S, S.ws64, S.pcg.precond, K, budget);
    }
    void norms(double& rr, double& bb) override {
        residual_norms(S, S.ws64.b.get(), S.ws64.x.get(), nullptr, rr, bb);
    }
};

// Most confirmation rounds of one solve.
constexpr int kMaxConfirmRounds = 8;

void pcg_open(PcgState& P, PcgOps& ops, const nlhip_solve_params& p, bool precond, uint32_t K) {
    P = PcgState{};
    P.p = p;
    P.precond = precond;
    P.K = K;
    P.cur = ops.start(p.eps);
    P.thresh = P.cur.thresh;
    P.active = true;
}

// Starts a new recurrence from the current iterate (residual replacement):
// r = A x - b recomputed in FP64, d = h = P r, the iteration count carried
// over in `base`.
static void pcg_restart(PcgState& P, PcgOps& ops) {
    P.base = P.total();
    P.cur = ops.start(P.p.eps);
    P.thresh = P.cur.thresh;
    ++P.restarts;
}

// Runs the open solve until it stops for good or has done `until`
// iterations in all (the end of a slice).  Each stop of the device loop is
// the recursive test (err <= thresh), the slice's cap or max_iter.  After the
// test, with p.confirm, the true residual |b - A x| decides: met -> finished;
// still above eps |b| -> one more round, either
//  * continued: the same recurrence goes on with the threshold scaled down
//    by the ratio measured (the iterates stay those of a solve with a smaller
//    threshold), when the true residual is within 2x of the recursive one, or
//  * restarted: a new recurrence from the current iterate (residual
//    replacement), when rounding has made the recursive residual drift below
//    the true one by more than that, or a continued round did not lower the
//    true residual by 1 %;
// and the solve gives up (not converged) at max_iter, after kMaxConfirmRounds
// rounds, or when a restarted round did not lower the true residual by 1 %.
void pcg_drive(PcgState& P, PcgOps& ops, uint64_t until) {
    const nlhip_solve_params& p = P.p;
    while (!P.finished) {
        const uint32_t its = P.total();
        if (P.ran && its >= until) break;  // end of the slice
        const uint32_t cap = uint32_t(std::min<uint64_t>(until, p.max_iter));
        ops.reopen(P.cur, P.thresh, int(cap - P.base));
        P.cur = ops.run(P.K, uint64_t(cap - std::min(cap, its)) + 1);
        P.ran = true;
        if (!std::isfinite(P.cur.bb) || !std::isfinite(P.cur.err))
            throw Error(NLHIP_ERR_NONFINITE, "right-hand side or iterate is NaN or Inf");
        if (P.cur.err > P.thresh) {
            if (P.total() >= p.max_iter) P.finished = true;
            continue;
        }
        if (P.first_stop == ~0u) P.first_stop = P.total();
        if (!p.confirm) {
            P.finished = true;
            break;
        }
        ops.norms(P.rr, P.bb);
        const double target = (p.eps * p.eps) * P.bb;
        const bool improved = P.last_rr < 0.0 || P.rr <= 0.99 * P.last_rr;
        if (!(P.rr > target) || P.total() >= p.max_iter || P.rounds >= kMaxConfirmRounds ||
            (!improved && P.last_kind == 2)) {
            P.finished = true;
            break;
        }
        P.last_rr = P.rr;
        ++P.rounds;
        if (!(P.cur.err > 0.0) || P.rr > 4.0 * P.cur.err || !improved) {
            P.last_kind = 2;
            pcg_restart(P, ops);
        } else {
            P.last_kind = 1;
            P.thresh = P.cur.err * (target / P.rr) * 0.5;
        }
    }
}

void pcg_stats(const PcgState& P, nlhip_solve_stats& st) {
    st.iterations = P.total();
    st.outer_iterations = P.restarts;
    const double bb = P.cur.bb;
    st.cg_rel_residual = bb > 0 ? std::sqrt(P.cur.err / bb) : std::sqrt(P.cur.err);
    if (P.p.confirm && P.rr >= 0.0 && P.first_stop != ~0u)
        st.converged = !(P.rr > (P.p.eps * P.p.eps) * P.bb) ? 1 : 0;
    else
        st.converged = !(P.cur.err > P.cur.thresh) && P.first_stop != ~0u ? 1 : 0;
    st.extra_iterations = P.first_stop == ~0u ? 0u : P.total() - P.first_stop;
    st.device_seconds = P.device_seconds;
}

// Opens an FP64 PCG solve on ws64 (b and x already there): initial residual
// and status on the device, the status read back.
static void pcg_begin(System& S, const nlhip_solve_params& p, const double* dinv) {
    const bool precond = p.precond == NL_PRECOND_JACOBI;
    setup_args<double>(S, S.ws64, precond ? (dinv ? dinv : S.dinv.get()) : nullptr);
    GraphPcgOps ops(S);
    pcg_open(S.pcg, ops, p, precond, p.check_every ? p.check_every : auto_check_every(S));
}

// pcg_drive on the graph path, timed with HIP events.
static void pcg_advance(System& S, uint64_t until) {
    GraphPcgOps ops(S);
    Event e0, e1;
    e0.record(S.stream.get());
    pcg_drive(S.pcg, ops, until);
    e1.record(S.stream.get());
    e1.sync();
    S.pcg.device_seconds += 1e-3 * elapsed_ms(e0, e1);
}

static void solve_fp64(System& S, const nlhip_solve_params& p, nlhip_solve_stats& st,
                       const double* dinv) {
    pcg_begin(S, p, dinv);
    pcg_advance(S, p.max_iter);
    pcg_stats(S.pcg, st);
    S.pcg.active = false;
}

// FP32 inner PCG inside FP64 refinement.  fallback (NLHIP_PREC_MIXED_
// FALLBACK): an inner solve that uses its whole cap (inner_max_iter) ends
// the refinement and the solve finishes with FP64 CG from the current
// iterate (LAPACK DSGESV's policy).  The refinement keeps the iterate with
// the lowest FP64 residual so far: when the corrections stop improving on it
// (kRefineNoGain steps in a row, or one squared residual above
// kRefineDiverged times the lowest, as when an FP32 inner solve on an
// ill-conditioned system has lost the agreement between its recursive and
// true residuals), it goes back to that iterate and ends (with the
// fallback: continues in FP64 from there);
// every other exit also returns the best iterate.  Steps that raise the
// residual for a while are kept: CG lowers the A-norm of the error, not the
// 2-norm of the residual.
static void solve_mixed(System& S, const nlhip_solve_params& p, nlhip_solve_stats& st,
                        bool fallback) {
    S.ensure_fp32();
    CGWorkspace<double>& w64 = S.ws64;
    CGWorkspace<float>& w32 = S.ws32;
    w32.ensure(S.n, S.grid);
    const bool precond = p.precond == NL_PRECOND_JACOBI;
    setup_args<float>(S, w32, precond ? S.dinv32.get() : nullptr);
    hipStream_t s = S.stream.get();
    const uint32_t K = p.check_every ? p.check_every : auto_check_every(S);
    const uint32_t inner_cap = p.inner_max_iter ? p.inner_max_iter : p.max_iter;
    const size_t xbytes = size_t(S.n) * 8;
    double* const x_best = w64.d.get();  // free while refining
    uint32_t total = 0, outer = 0, no_gain = 0;
    double rr = 0, bb = 0, rr_best = 0;
    bool stalled = false, to_fp64 = false;
    for (;;) {
        residual_norms(S, w64.b.get(), w64.x.get(), w64.r.get(), rr, bb);
        if (outer == 0 || rr < rr_best) {
            NLHIP_CHECK(hipMemcpyAsync(x_best, w64.x.get(), xbytes, hipMemcpyDeviceToDevice, s));
            rr_best = rr;
            no_gain = 0;
        } else if (!std::isfinite(rr) || rr > kRefineDiverged * rr_best ||
                   ++no_gain >= kRefineNoGain) {
            stalled = true;
        }
        if (outer == 0 && (!std::isfinite(rr) || !std::isfinite(bb)))
            throw Error(NLHIP_ERR_NONFINITE, "right-hand side or iterate is NaN or Inf");
        if (!(rr > (p.eps * p.eps) * bb)) break;
        if (stalled) {
            to_fp64 = fallback && total < p.max_iter;
            break;
        }
        if (outer >= p.max_outer || total >= p.max_iter) break;
        const double rn = std::sqrt(rr);
        NLHIP_CHECK(ir_scale_cast(S.n, w64.r.get(), 1.0 / rn, w32.b.get(), s));
        NLHIP_CHECK(fill<float>(S.n, 0.0f, w32.x.get(), s));
        const uint32_t cap = std::min(inner_cap, p.max_iter - total);
        NLHIP_CHECK(cg_start<float>(w32.args, w32.code, precond, float(p.inner_eps), int(cap), s));
        const CGStatus<float> fin = run_ladder<float>(S, w32, precond, K, cap);
        total += uint32_t(fin.iters);
        NLHIP_CHECK(ir_update(S.n, w64.x.get(), w32.x.get(), rn, s));
        ++outer;
        if (fin.iters == 0) break;  // a zero correction: x is unchanged
        if (fallback && uint32_t(fin.iters) >= cap && total < p.max_iter) {
            residual_norms(S, w64.b.get(), w64.x.get(), w64.r.get(), rr, bb);
            if (!(rr > (p.eps * p.eps) * bb)) break;
            to_fp64 = true;
            break;
        }
    }
    // the best iterate
    if (!(rr <= rr_best)) {
        NLHIP_CHECK(hipMemcpyAsync(w64.x.get(), x_best, xbytes, hipMemcpyDeviceToDevice, s));
        rr = rr_best;
    }
    if (to_fp64) {
        nlhip_solve_params p64 = p;
        p64.max_iter = p.max_iter - total;
        nlhip_solve_stats s64{};
        solve_fp64(S, p64, s64, nullptr);
        st.iterations = total + s64.iterations;
        st.outer_iterations = outer;
        st.cg_rel_residual = s64.cg_rel_residual;
        st.converged = s64.converged;
        st.extra_iterations = s64.extra_iterations;
        return;
    }
    st.iterations = total;
    st.outer_iterations = outer;
    st.cg_rel_residual = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    st.converged = !(rr > (p.eps * p.eps) * bb) ? 1 : 0;
}

static void check_params(const nlhip_solve_params& p) {
    if (p.precision != NLHIP_PREC_FP64 && p.precision != NLHIP_PREC_MIXED &&
        p.precision != NLHIP_PREC_MIXED_FALLBACK)
        throw Error(NLHIP_ERR_INVALID, "unknown precision");
    if (p.precision == NLHIP_PREC_MIXED_FALLBACK && p.inner_max_iter == 0)
        throw Error(NLHIP_ERR_INVALID, "the fallback needs inner_max_iter > 0");
    if (p.precond != NL_PRECOND_JACOBI && p.precond != NL_PRECOND_NONE)
        throw Error(NLHIP_ERR_UNSUPPORTED, "only Jacobi or no preconditioner on the GPU");
}

void solve(System& S, const double* b, double* x, const nlhip_solve_params& p,
           nlhip_solve_stats& st, const double* dinv) {
    std::lock_guard<std::mutex> lock(S.mu);
    const auto t0 = std::chrono::steady_clock::now();
    std::memset(&st, 0, sizeof(st));
    if (take_injected_failure()) throw Error(NLHIP_ERR_INJECTED, "injected failure");
    check_params(p);
    S.pcg.active = false;
    S.ws64.ensure(S.n, S.grid);
    hipStream_t s = S.stream.get();
    copy_in(S.ws64.b.get(), b, size_t(S.n) * 8, s);
    copy_in(S.ws64.x.get(), x, size_t(S.n) * 8, s);
    Event e0, e1;
    e0.record(s);
    if (p.precision == NLHIP_PREC_FP64) solve_fp64(S, p, st, dinv);
    else solve_mixed(S, p, st, p.precision == NLHIP_PREC_MIXED_FALLBACK);
    e1.record(s);
    e1.sync();
    st.device_seconds = 1e-3 * elapsed_ms(e0, e1);
    st.rel_residual = true_rel_residual(S, S.ws64.b.get(), S.ws64.x.get());
    copy_out(x, S.ws64.x.get(), size_t(S.n) * 8, s);
    S.stream.sync();
    st.wall_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    st.status = NLHIP_OK;
}

// Sliced solves (nlhipSolveBegin / Advance / End).
static void sliced_begin(System& S, const double* b, const double* x0,
                         const nlhip_solve_params& p) {
    std::lock_guard<std::mutex> lock(S.mu);
    check_params(p);
    if (p.precision != NLHIP_PREC_FP64)
        throw Error(NLHIP_ERR_UNSUPPORTED, "sliced solves run FP64 CG only");
    S.pcg.active = false;
    S.ws64.ensure(S.n, S.grid);
    hipStream_t s = S.stream.get();
    copy_in(S.ws64.b.get(), b, size_t(S.n) * 8, s);
    if (x0) copy_in(S.ws64.x.get(), x0, size_t(S.n) * 8, s);
    else NLHIP_CHECK(fill<double>(S.n, 0.0, S.ws64.x.get(), s));
    pcg_begin(S, p, nullptr);
}

static void sliced_advance(System& S, uint32_t iters, int* finished, nlhip_solve_stats* st) {
    std::lock_guard<std::mutex> lock(S.mu);
    if (!S.pcg.active) throw Error(NLHIP_ERR_INVALID, "no sliced solve is open on this system");
    pcg_advance(S, uint64_t(uint32_t(S.pcg.cur.iters)) + iters);
    if (finished) *finished = S.pcg.finished ? 1 : 0;
    if (st) {
        std::memset(st, 0, sizeof(*st));
        pcg_stats(S.pcg, *st);
        st->rel_residual = S.pcg.rr >= 0.0 && S.pcg.bb > 0 ? std::sqrt(S.pcg.rr / S.pcg.bb)
                                                            : st->cg_rel_residual;
    }
}

static void sliced_end(System& S, double* x, nlhip_solve_stats* st) {
    std::lock_guard<std::mutex> lock(S.mu);
    if (!S.pcg.active) throw Error(NLHIP_ERR_INVALID, "no sliced solve is open on this system");
    if (st) {
        std::memset(st, 0, sizeof(*st));
        pcg_stats(S.pcg, *st);
        st->rel_residual = true_rel_residual(S, S.ws64.b.get(), S.ws64.x.get());
    }
    if (x) copy_out(x, S.ws64.x.get(), size_t(S.n) * 8, S.stream.get());
    S.stream.sync();
    S.pcg.active = false;
}

// ------------------------------------------------------------- BiCGSTAB

namespace {
struct DotHelper {
    System& S;
    DevBuf<double>& part;
    double* out;
    double* host;
    double operator()(const double* a, const double* b) {
        hipStream_t s = S.stream.get();
        NLHIP_CHECK(blas_ddot(S.n, a, b, part.get(), S.grid, out, s));
        NLHIP_CHECK(hipMemcpyAsync(host, out, sizeof(double), hipMemcpyDeviceToHost, s));
        S.stream.sync();
        return *host;
    }
};
}  // namespace

// Same operation sequence as the van der Vorst BiCGSTAB restated in the
// oracle (right preconditioning by P, r = b - A x).
uint32_t bicgstab_device(System& S, const Precond& P, const double* b, double* x, double eps,
                         uint32_t max_iter, double* rel2) {
    const uint32_t n = S.n;
    hipStream_t s = S.stream.get();
    DevBuf<double> buf(size_t(8) * n);
    double *r = buf.get(), *rt = r + n, *p = rt + n, *v = p + n, *ph = v + n, *sv = ph + n,
           *sh = sv + n, *t = sh + n;
    DevBuf<double> part(S.grid), out(2);
    S.host_scalars.ensure(4);
    DotHelper dot{S, part, out.get(), S.host_scalars.get()};
    const SpmvOps<double> A = S.ops<double>();
    auto spmv64 = [&](const double* in, double* o) {
        NLHIP_CHECK(spmv<double>(n, A.rp, A.cols, A.val, in, o, A.code, S.grid, s));
    };
    auto precond = [&](const double* in, double* o) { precond_apply(P, n, in, o, s); };
    auto copy = [&](const double* in, double* o) {
        NLHIP_CHECK(hipMemcpyAsync(o, in, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
    };
    spmv64(x, r);
    NLHIP_CHECK(blas_dscal(n, -1.0, r, s));
    NLHIP_CHECK(blas_daxpy(n, 1.0, b, r, s));
    copy(r, rt);
    NLHIP_CHECK(fill<double>(n, 0.0, p, s));
    NLHIP_CHECK(fill<double>(n, 0.0, v, s));
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    const double bb = dot(b, b);
    const double thresh = (eps * eps) * bb;
    double err = dot(r, r);
    uint32_t its = 0;
    while (err > thresh && its < max_iter) {
        const double rho_new = dot(rt, r);
        if (rho_new == 0.0) break;
        const double beta = (rho_new / rho) * (alpha / omega);
        NLHIP_CHECK(blas_daxpy(n, -omega, v, p, s));
        NLHIP_CHECK(blas_dscal(n, beta, p, s));
        NLHIP_CHECK(blas_daxpy(n, 1.0, r, p, s));
        precond(p, ph);
        spmv64(ph, v);
        const double rtv = dot(rt, v);
        if (rtv == 0.0) break;
        alpha = rho_new / rtv;
        copy(r, sv);
        NLHIP_CHECK(blas_daxpy(n, -alpha, v, sv, s));
        precond(sv, sh);
        spmv64(sh, t);
        const double tt = dot(t, t);
        NLHIP_CHECK(blas_daxpy(n, alpha, ph, x, s));
        ++its;
        if (tt == 0.0) {
            copy(sv, r);
            err = dot(r, r);
            break;
        }
        omega = dot(t, sv) / tt;
        NLHIP_CHECK(blas_daxpy(n, omega, sh, x, s));
        copy(sv, r);
        NLHIP_CHECK(blas_daxpy(n, -omega, t, r, s));
        rho = rho_new;
        err = dot(r, r);
        if (omega == 0.0) break;
    }
    S.stream.sync();
    if (rel2) *rel2 = bb > 0.0 ? err / bb : err;
    return its;
}

// PCG with any preconditioner, in the operation order of the oracle's
// pcg_core (r = A x - b, x -= alpha d; OpenNL's CG behind
// nlSolveSystemIterative, scene_graph/NL/matrix.cpp:390-393), each inner
// product reduced on the device and read back.  The fused CG (solve) stays
// the path for Jacobi and no preconditioner; this one serves SSOR.
uint32_t pcg_device(System& S, const Precond& P, const double* b, double* x, double eps,
                    uint32_t max_iter, double* rel2) {
    const uint32_t n = S.n;
    hipStream_t s = S.stream.get();
    DevBuf<double> buf(size_t(4) * n);
    double *r = buf.get(), *d = r + n, *h = d + n, *Ad = h + n;
    DevBuf<double> part(S.grid), out(2);
    S.host_scalars.ensure(4);
    DotHelper dot{S, part, out.get(), S.host_scalars.get()};
    const SpmvOps<double> A = S.ops<double>();
    auto spmv64 = [&](const double* in, double* o) {
        NLHIP_CHECK(spmv<double>(n, A.rp, A.cols, A.val, in, o, A.code, S.grid, s));
    };
    spmv64(x, r);
    NLHIP_CHECK(blas_daxpy(n, -1.0, b, r, s));
    precond_apply(P, n, r, d, s);
    NLHIP_CHECK(hipMemcpyAsync(h, d, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
    double rh = dot(r, h);
    double err = dot(r, r);
    const double bb = dot(b, b);
    const double thresh = (eps * eps) * bb;
    uint32_t its = 0;
    while (err > thresh && its < max_iter) {
        spmv64(d, Ad);
        const double alpha = rh / dot(d, Ad);
        NLHIP_CHECK(blas_daxpy(n, -alpha, d, x, s));
        NLHIP_CHECK(blas_daxpy(n, -alpha, Ad, r, s));
        precond_apply(P, n, r, h, s);
        const double rh_new = dot(r, h);
        const double beta = rh_new / rh;
        rh = rh_new;
        NLHIP_CHECK(blas_dscal(n, beta, d, s));
        NLHIP_CHECK(blas_daxpy(n, 1.0, h, d, s));
        ++its;
        err = dot(r, r);
    }
    S.stream.sync();
    if (rel2) *rel2 = bb > 0.0 ? err / bb : err;
    return its;
}

// Restarted GMRES(m) (Saad & Schultz 1986), modified Gram-Schmidt Arnoldi,
// Givens rotations, right preconditioning with P (Jacobi, SSOR or none).  Same
// method and stopping rule as the oracle's or_gmres; each inner product is
// reduced on the device and read back (a host-driven loop, like BiCGSTAB).
uint32_t gmres_device(System& S, const Precond& P, const double* b, double* x, double eps,
                      uint32_t max_iter, uint32_t m, double* rel2) {
    if (m < 1) m = 1;
    const uint32_t n = S.n;
    hipStream_t s = S.stream.get();
    DevBuf<double> Vb(size_t(m + 1) * n), buf(size_t(3) * n);
    double* V = Vb.get();
    double *w = buf.get(), *z = w + n, *u = z + n;
    DevBuf<double> part(S.grid), out(2);
    S.host_scalars.ensure(4);
    DotHelper dot{S, part, out.get(), S.host_scalars.get()};
    const SpmvOps<double> A = S.ops<double>();
    auto spmv64 = [&](const double* in, double* o) {
        NLHIP_CHECK(spmv<double>(n, A.rp, A.cols, A.val, in, o, A.code, S.grid, s));
    };
    auto precond = [&](const double* in, double* o) { precond_apply(P, n, in, o, s); };
    auto residual = [&] {  // V_0 = b - A x
        spmv64(x, V);
        NLHIP_CHECK(blas_dscal(n, -1.0, V, s));
        NLHIP_CHECK(blas_daxpy(n, 1.0, b, V, s));
    };
    std::vector<double> H(size_t(m + 1) * m, 0.0), cs(m), sn(m), g(size_t(m) + 1), y(m);
    const double bb = dot(b, b);
    const double thresh = (eps * eps) * bb;
    residual();
    double err = dot(V, V);
    uint32_t its = 0;
    while (err > thresh && its < max_iter) {
        const double beta = std::sqrt(err);
        NLHIP_CHECK(blas_dscal(n, 1.0 / beta, V, s));
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        uint32_t k = 0;
        bool breakdown = false;
        for (uint32_t j = 0; j < m && its < max_iter; ++j) {
            double* vj = V + size_t(j) * n;
            precond(vj, z);
            spmv64(z, w);
            for (uint32_t i = 0; i <= j; ++i) {
                const double h = dot(w, V + size_t(i) * n);
                H[size_t(i) * m + j] = h;
                NLHIP_CHECK(blas_daxpy(n, -h, V + size_t(i) * n, w, s));
            }
            const double hn = std::sqrt(dot(w, w));
            H[size_t(j + 1) * m + j] = hn;
            if (hn != 0.0) {
                double* vn = V + size_t(j + 1) * n;
                NLHIP_CHECK(hipMemcpyAsync(vn, w, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
                NLHIP_CHECK(blas_dscal(n, 1.0 / hn, vn, s));
            } else {
                breakdown = true;
            }
            for (uint32_t i = 0; i < j; ++i) {
                const double h0 = H[size_t(i) * m + j], h1 = H[size_t(i + 1) * m + j];
                H[size_t(i) * m + j] = cs[i] * h0 + sn[i] * h1;
                H[size_t(i + 1) * m + j] = -sn[i] * h0 + cs[i] * h1;
            }
            const double h0 = H[size_t(j) * m + j], h1 = H[size_t(j + 1) * m + j];
            const double den = std::hypot(h0, h1);
            cs[j] = den != 0.0 ? h0 / den : 1.0;
            sn[j] = den != 0.0 ? h1 / den : 0.0;
            H[size_t(j) * m + j] = den;
            H[size_t(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++its;
            k = j + 1;
            if (g[j + 1] * g[j + 1] <= thresh || breakdown) break;
        }
        for (int i = int(k) - 1; i >= 0; --i) {
            double t = g[size_t(i)];
            for (uint32_t l = uint32_t(i) + 1; l < k; ++l) t -= H[size_t(i) * m + l] * y[l];
            y[size_t(i)] = H[size_t(i) * m + size_t(i)] != 0.0 ? t / H[size_t(i) * m + size_t(i)]
                                                                : 0.0;
        }
        NLHIP_CHECK(fill<double>(n, 0.0, u, s));
        for (uint32_t i = 0; i < k; ++i) NLHIP_CHECK(blas_daxpy(n, y[i], V + size_t(i) * n, u, s));
        precond(u, z);
        NLHIP_CHECK(blas_daxpy(n, 1.0, z, x, s));
        residual();
        err = dot(V, V);
        if (breakdown) break;
    }
    S.stream.sync();
    if (rel2) *rel2 = bb > 0.0 ? err / bb : err;
    return its;
}

}  // namespace nlhip

using namespace nlhip;

namespace nlhip {

// Mean seconds of the three CG launches (direction, SpMV, update) over
// `iters` iterations of the system's launch shape (HIP events on its stream;
// S.mu held).
void profile_cg(System& S, bool fp32, uint32_t iters, double t[3]) {
    S.pcg.active = false;  // runs on ws64 / ws32
    hipStream_t st = S.stream.get();
    if (iters == 0) iters = 1;
    auto profile = [&](auto& ws, auto* dinv, auto zero) {
        using T = decltype(zero);
        ws.ensure(S.n, S.grid);
        setup_args<T>(S, ws, dinv);
        NLHIP_CHECK(fill<T>(S.n, T(1), ws.b.get(), st));
        NLHIP_CHECK(fill<T>(S.n, T(0), ws.x.get(), st));
        NLHIP_CHECK(cg_start<T>(ws.args, ws.code, true, T(0), int(iters) + 8, st));
        std::vector<Event> ev(4 * size_t(iters));
        for (uint32_t k = 0; k < iters; ++k) {
            for (int part = 0; part < 3; ++part) {
                ev[4 * k + part].record(st);
                NLHIP_CHECK(cg_step_part<T>(ws.args, ws.code, true, part, st));
            }
            ev[4 * k + 3].record(st);
        }
        S.stream.sync();
        double acc[3] = {0, 0, 0};
        for (uint32_t k = 0; k < iters; ++k)
            for (int part = 0; part < 3; ++part)
                acc[part] += 1e-3 * elapsed_ms(ev[4 * k + part], ev[4 * k + part + 1]);
        for (int part = 0; part < 3; ++part) t[part] = acc[part] / iters;
    };
    if (fp32) {
        S.ensure_fp32();
        profile(S.ws32, S.dinv32.get(), 0.0f);
    } else {
        profile(S.ws64, S.dinv.get(), 0.0);
    }
}

// Autotuner: times `iters` CG iterations (the three launches) for each
// candidate launch shape and keeps the fastest.  Candidates: the streamed
// SpMV engines, which give the same row sums (the 2x2 block engine with 1,
// 2, 4 or 8 blocks per lane per chunk, and with 4 in XCD-interleaved runs,
// when the matrix has the block form with 16-bit block columns, with 4 or 8
// when its block columns are 32-bit;
// the scalar engine with 16-bit or 32-bit columns, contiguous or
// XCD-interleaved), times workgroup counts of 1, 2, 4 and 8 per CU (the
// dot products are summed per workgroup, so the count fixes their order:
// a tuned system gives the same bits on every solve).  A first pass times
// every shape once; the four fastest and the current shape are then timed
// again, three times each with three times the iterations, interleaved, and
// each keeps its best time.  Another shape replaces the current one only
// when 1 % faster in that second pass, so timing noise rarely moves the
// choice (and with it the order of the sums) from run to run.
TuneResult autotune(System& S, bool fp32, uint32_t iters) {
    std::vector<int> codes;
    if (S.blocks() && S.bci16.get()) codes = {73, 74, 75, 76, 77};
    if (S.blocks() && !S.bci16.get()) codes = {78, 79};
    if (S.ci16.get()) {
        codes.push_back(68);
        codes.push_back(70);
    }
    codes.push_back(71);
    codes.push_back(69);
    const uint32_t cu = uint32_t(std::max(1, device().cu_count));
    const uint64_t max_grid = std::max<uint64_t>(1, (uint64_t(S.n) + 63) / 64);
    std::vector<uint32_t> grids;
    for (uint32_t m : {1u, 2u, 4u, 8u}) {
        const uint32_t g = uint32_t(std::min<uint64_t>(uint64_t(cu) * m, max_grid));
        if (std::find(grids.begin(), grids.end(), g) == grids.end()) grids.push_back(g);
    }
    const int rpw0 = S.rpw;
    const uint32_t grid0 = S.grid;
    double t[3];
    std::vector<std::pair<int, uint32_t>> shapes{{rpw0, grid0}};
    for (int code : codes)
        for (uint32_t g : grids)
            if (!(code == rpw0 && g == grid0)) shapes.emplace_back(code, g);
    auto time_shape = [&](const std::pair<int, uint32_t>& shape, uint32_t its) {
        S.rpw = shape.first;
        S.grid = shape.second;
        profile_cg(S, fp32, 2, t);  // warm-up (argument set-up, caches)
        profile_cg(S, fp32, its, t);
        return t[0] + t[1] + t[2];
    };
    std::vector<std::pair<double, size_t>> first;
    for (size_t i = 0; i < shapes.size(); ++i) first.emplace_back(time_shape(shapes[i], iters), i);
    std::sort(first.begin() + 1, first.end());
    std::vector<size_t> finalists{0};
    for (size_t i = 1; i < first.size() && finalists.size() < 5; ++i)
        finalists.push_back(first[i].second);
    std::vector<double> best_t(finalists.size(), 1e300);
    for (int rep = 0; rep < 3; ++rep)
        for (size_t f = 0; f < finalists.size(); ++f)
            best_t[f] = std::min(best_t[f], time_shape(shapes[finalists[f]], 3 * iters));
    TuneResult best{rpw0, grid0, best_t[0]};
    for (size_t f = 1; f < finalists.size(); ++f)
        if (best_t[f] < 0.99 * best_t[0] && best_t[f] < best.seconds)
            best = TuneResult{shapes[finalists[f]].first, shapes[finalists[f]].second, best_t[f]};
    S.rpw = best.rpw;
    S.grid = best.grid;
    S.ws64.graph.reset();
    S.ws32.graph.reset();
    return best;
}

}  // namespace nlhip

static System* sys(nlhip_system h) {
    if (!h) throw Error(NLHIP_ERR_INVALID, "null system handle");
    return reinterpret_cast<System*>(h);
}

extern "C" {

NLAPI void NLAPIENTRY nlhipSolveParamsDefault(nlhip_solve_params* p) {
    if (!p) return;
    p->precision = NLHIP_PREC_FP64;
    p->precond = NL_PRECOND_JACOBI;
    p->eps = 1e-10;
    p->max_iter = 5000;
    p->check_every = 0;
    p->inner_eps = 1e-4;
    p->max_outer = 30;
    p->inner_max_iter = 0;
    p->confirm = 1;
}

NLAPI int NLAPIENTRY nlhipSystemCreate(uint32_t n, const uint32_t* rowptr,
                                       const uint32_t* colind, const double* val,
                                       uint32_t flags, nlhip_system* out) {
    return guarded("nlhipSystemCreate", [&] {
        if (!out) throw Error(NLHIP_ERR_INVALID, "null output handle");
        *out = nullptr;
        System* S = system_create(n, rowptr, colind, val, flags);
        *out = reinterpret_cast<nlhip_system>(S);
    });
}

NLAPI int NLAPIENTRY nlhipSystemDestroy(nlhip_system s) {
    return guarded("nlhipSystemDestroy", [&] {
        if (!s) return;
        System* S = sys(s);
        { std::lock_guard<std::mutex> lock(S->mu); S->stream.sync(); }
        delete S;
    });
}

NLAPI int NLAPIENTRY nlhipSystemInfo(nlhip_system s, uint32_t* n, uint64_t* nnz,
                                     uint64_t* device_bytes) {
    return guarded("nlhipSystemInfo", [&] {
        System* S = sys(s);
        if (n) *n = S->n;
        if (nnz) *nnz = S->nnz;
        if (device_bytes) *device_bytes = S->device_bytes();
    });
}

NLAPI int NLAPIENTRY nlhipSystemSetLaunch(nlhip_system s, int rows_per_wave,
                                          uint32_t workgroups) {
    return guarded("nlhipSystemSetLaunch", [&] {
        System* S = sys(s);
        if (!valid_rpw(rows_per_wave))
            throw Error(NLHIP_ERR_INVALID, "rows_per_wave must be 2, 4, 8, 16 or 64..79");
        if (workgroups == 0 || workgroups > 65536)
            throw Error(NLHIP_ERR_INVALID, "workgroups must be in [1, 65536]");
        std::lock_guard<std::mutex> lock(S->mu);
        S->rpw = rows_per_wave;
        S->grid = workgroups;
        S->ws64.graph.reset();
        S->ws32.graph.reset();
    });
}

NLAPI int NLAPIENTRY nlhipSystemGetLaunch(nlhip_system s, int* rows_per_wave,
                                          uint32_t* workgroups) {
    return guarded("nlhipSystemGetLaunch", [&] {
        System* S = sys(s);
        if (rows_per_wave) *rows_per_wave = S->rpw;
        if (workgroups) *workgroups = S->grid;
    });
}

NLAPI int NLAPIENTRY nlhipSolve(nlhip_system s, const double* b, double* x,
                                const nlhip_solve_params* params, nlhip_solve_stats* stats) {
    nlhip_solve_stats local;
    nlhip_solve_stats& st = stats ? *stats : local;
    int rc = guarded("nlhipSolve", [&] {
        System* S = sys(s);
        if (!b || !x) throw Error(NLHIP_ERR_INVALID, "null b or x");
        nlhip_solve_params p;
        nlhipSolveParamsDefault(&p);
        if (params) p = *params;
        solve(*S, b, x, p, st);
    });
    st.status = rc;
    return rc;
}

NLAPI int NLAPIENTRY nlhipSolveBegin(nlhip_system s, const double* b, const double* x0,
                                     const nlhip_solve_params* params) {
    return guarded("nlhipSolveBegin", [&] {
        System* S = sys(s);
        if (!b) throw Error(NLHIP_ERR_INVALID, "null b");
        nlhip_solve_params p;
        nlhipSolveParamsDefault(&p);
        if (params) p = *params;
        sliced_begin(*S, b, x0, p);
    });
}

NLAPI int NLAPIENTRY nlhipSolveAdvance(nlhip_system s, uint32_t iters, int* finished,
                                       nlhip_solve_stats* stats) {
    return guarded("nlhipSolveAdvance", [&] { sliced_advance(*sys(s), iters, finished, stats); });
}

NLAPI int NLAPIENTRY nlhipSolveEnd(nlhip_system s, double* x, nlhip_solve_stats* stats) {
    return guarded("nlhipSolveEnd", [&] { sliced_end(*sys(s), x, stats); });
}

NLAPI int NLAPIENTRY nlhipSpMV(nlhip_system s, const double* x, double* y) {
    return guarded("nlhipSpMV", [&] {
        System* S = sys(s);
        std::lock_guard<std::mutex> lock(S->mu);
        S->pcg.active = false;  // stages through ws64
        hipStream_t st = S->stream.get();
        S->ws64.ensure(S->n, S->grid);
        copy_in(S->ws64.x.get(), x, size_t(S->n) * 8, st);
        const SpmvOps<double> o = S->ops<double>();
        NLHIP_CHECK(spmv<double>(S->n, o.rp, o.cols, o.val, S->ws64.x.get(), S->ws64.q.get(),
                                 o.code, S->grid, st));
        copy_out(y, S->ws64.q.get(), size_t(S->n) * 8, st);
        S->stream.sync();
    });
}

NLAPI int NLAPIENTRY nlhipBenchSpMV(nlhip_system s, int fp32, uint32_t reps,
                                    double* seconds_per_spmv) {
    return guarded("nlhipBenchSpMV", [&] {
        System* S = sys(s);
        std::lock_guard<std::mutex> lock(S->mu);
        S->pcg.active = false;  // runs on ws64 / ws32
        hipStream_t st = S->stream.get();
        if (reps == 0) reps = 1;
        Event e0, e1;
        if (fp32) {
            S->ensure_fp32();
            S->ws32.ensure(S->n, S->grid);
            NLHIP_CHECK(fill<float>(S->n, 1.0f, S->ws32.x.get(), st));
            const SpmvOps<float> o = S->ops<float>();
            auto run = [&] {
                NLHIP_CHECK(spmv<float>(S->n, o.rp, o.cols, o.val, S->ws32.x.get(),
                                        S->ws32.q.get(), o.code, S->grid, st));
            };
            run();
            e0.record(st);
            for (uint32_t k = 0; k < reps; ++k) run();
            e1.record(st);
        } else {
            S->ws64.ensure(S->n, S->grid);
            NLHIP_CHECK(fill<double>(S->n, 1.0, S->ws64.x.get(), st));
            const SpmvOps<double> o = S->ops<double>();
            auto run = [&] {
                NLHIP_CHECK(spmv<double>(S->n, o.rp, o.cols, o.val, S->ws64.x.get(),
                                         S->ws64.q.get(), o.code, S->grid, st));
            };
            run();
            e0.record(st);
            for (uint32_t k = 0; k < reps; ++k) run();
            e1.record(st);
        }
        e1.sync();
        if (seconds_per_spmv) *seconds_per_spmv = 1e-3 * elapsed_ms(e0, e1) / reps;
    });
}

NLAPI int NLAPIENTRY nlhipBenchCopy(uint64_t bytes, uint32_t reps, double* seconds_per_copy) {
    return guarded("nlhipBenchCopy", [&] {
        if (bytes == 0 || bytes % 16 != 0)
            throw Error(NLHIP_ERR_INVALID, "copy size must be a positive multiple of 16 bytes");
        const Device& d = device();
        if (reps == 0) reps = 1;
        DevBuf<uint8_t> src(bytes), dst(bytes);
        Stream stream;
        hipStream_t st = stream.get();
        NLHIP_CHECK(fill<int32_t>(bytes / 4, 1, reinterpret_cast<int32_t*>(src.get()), st));
        // best of a few grid sizes: 4, 16 and 64 workgroups per CU, and one
        // four-word trip per thread
        const uint64_t n16 = bytes / 16;
        const uint64_t one_trip = std::max<uint64_t>(1, (n16 + 4 * 256 - 1) / (4 * 256));
        double best = 0;
        for (uint64_t g : {uint64_t(d.cu_count) * 4, uint64_t(d.cu_count) * 16,
                           uint64_t(d.cu_count) * 64, one_trip}) {
            const uint32_t grid = uint32_t(std::min<uint64_t>(g, 1u << 30));
            NLHIP_CHECK(copy16(n16, src.get(), dst.get(), grid, st));
            Event e0, e1;
            e0.record(st);
            for (uint32_t k = 0; k < reps; ++k)
                NLHIP_CHECK(copy16(n16, src.get(), dst.get(), grid, st));
            e1.record(st);
            e1.sync();
            const double t = 1e-3 * elapsed_ms(e0, e1) / reps;
            if (best == 0 || t < best) best = t;
        }
        if (seconds_per_copy) *seconds_per_copy = best;
    });
}

NLAPI int NLAPIENTRY nlhipBenchRead(uint64_t bytes, uint32_t reps, double* seconds_per_pass) {
    return guarded("nlhipBenchRead", [&] {
        if (bytes == 0 || bytes % 16 != 0)
            throw Error(NLHIP_ERR_INVALID, "read size must be a positive multiple of 16 bytes");
        const Device& d = device();
        if (reps == 0) reps = 1;
        const uint64_t n16 = bytes / 16;
        const uint64_t one_trip = std::max<uint64_t>(1, (n16 + 8 * 256 - 1) / (8 * 256));
        const uint64_t max_grid = std::max<uint64_t>(uint64_t(d.cu_count) * 64, one_trip);
        DevBuf<uint8_t> src(bytes), out(max_grid * 256 * 16);
        Stream stream;
        hipStream_t st = stream.get();
        NLHIP_CHECK(fill<int32_t>(bytes / 4, 1, reinterpret_cast<int32_t*>(src.get()), st));
        double best = 0;
        auto time_it = [&](auto launch) {
            NLHIP_CHECK(launch());
            Event e0, e1;
            e0.record(st);
            for (uint32_t k = 0; k < reps; ++k) NLHIP_CHECK(launch());
            e1.record(st);
            e1.sync();
            const double t = 1e-3 * elapsed_ms(e0, e1) / reps;
            if (best == 0 || t < best) best = t;
        };
        // grid-stride words (plain and non-temporal loads) and one
        // contiguous run per wavefront, each over a few grid sizes
        for (uint64_t g : {uint64_t(d.cu_count) * 4, uint64_t(d.cu_count) * 16,
                           uint64_t(d.cu_count) * 64, one_trip}) {
            const uint32_t grid = uint32_t(std::min<uint64_t>(g, max_grid));
            for (bool nt : {false, true})
                time_it([&] { return read16(n16, src.get(), out.get(), grid, nt, st); });
        }
        for (uint64_t m : {2u, 4u, 8u, 16u}) {
            const uint32_t grid = uint32_t(std::min<uint64_t>(uint64_t(d.cu_count) * m, max_grid));
            time_it([&] { return read16_runs(n16, src.get(), out.get(), grid, st); });
        }
        if (seconds_per_pass) *seconds_per_pass = best;
    });
}

NLAPI int NLAPIENTRY nlhipBenchStreamMix(uint64_t blocks, uint32_t blocks_per_row, int writes,
                                         uint32_t reps, double* seconds_per_pass) {
    return guarded("nlhipBenchStreamMix", [&] {
        if (blocks_per_row == 0 || blocks < blocks_per_row)
            throw Error(NLHIP_ERR_INVALID, "need blocks >= blocks_per_row > 0");
        const Device& d = device();
        if (reps == 0) reps = 1;
        const uint64_t rows = blocks / blocks_per_row;
        DevBuf<double> vals(block_vals_size(blocks)), xg(2 * rows), rin(2 * rows),
            rout(2 * rows), out(2);
        DevBuf<uint16_t> off(blocks);
        Stream stream;
        hipStream_t st = stream.get();
        NLHIP_CHECK(fill<double>(block_vals_size(blocks), 1.0, vals.get(), st));
        NLHIP_CHECK(fill<double>(2 * rows, 1.0, xg.get(), st));
        NLHIP_CHECK(fill<double>(2 * rows, 1.0, rin.get(), st));
        NLHIP_CHECK(hipMemsetAsync(off.get(), 0, blocks * 2, st));
        double best = 0;
        for (uint64_t m : {2u, 4u, 8u, 16u}) {
            const uint32_t grid = uint32_t(uint64_t(d.cu_count) * m);
            auto launch = [&] {
                return stream_mix(blocks, blocks_per_row, vals.get(), off.get(), xg.get(),
                                  rin.get(), rout.get(), out.get(), writes != 0, grid, st);
            };
            NLHIP_CHECK(launch());
            Event e0, e1;
            e0.record(st);
            for (uint32_t k = 0; k < reps; ++k) NLHIP_CHECK(launch());
            e1.record(st);
            e1.sync();
            const double t = 1e-3 * elapsed_ms(e0, e1) / reps;
            if (best == 0 || t < best) best = t;
        }
        if (seconds_per_pass) *seconds_per_pass = best;
    });
}

// Runs `iters` PCG iterations (FP64, or FP32 with fp32 = 1) on the system
// with b = 1, launching each kernel directly and timing it with HIP events on
// the system's stream; seconds[0..2] = mean duration of the direction, SpMV
// and update launches.
NLAPI int NLAPIENTRY nlhipProfileCG(nlhip_system s, int fp32, uint32_t iters, double* seconds) {
    return guarded("nlhipProfileCG", [&] {
        System* S = sys(s);
        std::lock_guard<std::mutex> lock(S->mu);
        double t[3];
        profile_cg(*S, fp32 != 0, iters, t);
        if (seconds)
            for (int part = 0; part < 3; ++part) seconds[part] = t[part];
    });
}

NLAPI int NLAPIENTRY nlhipSystemAutotune(nlhip_system s, int fp32, uint32_t iters,
                                         int* rows_per_wave, uint32_t* workgroups,
                                         double* seconds_per_iteration) {
    return guarded("nlhipSystemAutotune", [&] {
        System* S = sys(s);
        std::lock_guard<std::mutex> lock(S->mu);
        const TuneResult r = autotune(*S, fp32 != 0, iters ? iters : 20);
        if (rows_per_wave) *rows_per_wave = r.rpw;
        if (workgroups) *workgroups = r.grid;
        if (seconds_per_iteration) *seconds_per_iteration = r.seconds;
    });
}

}  // extern "C"
