// Multiple right-hand sides of one resident system (OpenNL's NL_NB_SYSTEMS,
// reached through mesh_smooth at src/lib/OGF/mesh/commands/
// mesh_grob_surface_commands.cpp:703): host side of k_spmm.hip.
//
// The k recurrences run together and share every pass over the matrix; each
// column keeps its own scalars and stopping test, so column j follows the
// same iteration as a single-right-hand-side solve of that column.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "nlhip_internal.h"

namespace nlhip {

namespace {

uint32_t pow2_ceil(uint32_t k) {
    uint32_t g = 1;
    while (g < k) g <<= 1;
    return g;
}

// Workgroups of the multi right-hand-side launches: enough for one pass of
// G-lane row groups over the matrix, at most 16 per CU (the kernels use few
// registers, so the extra waves hide the gather latency).
uint32_t multi_grid(const System& S, uint32_t k) {
    const uint32_t per = 256u / pow2_ceil(k);
    const uint64_t passes = (uint64_t(S.n) + per - 1) / per;
    const uint64_t cap = uint64_t(std::max(1, device().cu_count)) * 16;
    return uint32_t(std::max<uint64_t>(1, std::min(passes, cap)));
}

template <typename T>
void multi_alloc(System& S, MultiWork<T>& w, uint32_t k) {
    const uint32_t P = multi_grid(S, k);
    if (w.n == S.n && w.k == k && w.P == P && w.x.get()) return;
    w.graph.reset();
    w.n = S.n;
    w.k = k;
    w.P = P;
    const size_t nk = size_t(S.n) * k;
    w.x.alloc(nk);
    w.b.alloc(nk);
    w.r.alloc(nk);
    w.h.alloc(nk);
    w.d.alloc(nk);
    w.q.alloc(nk);
    w.part.alloc(4 * size_t(P) * k);
    w.col.alloc(6 * size_t(k));
    w.iters.alloc(k);
    w.done.alloc(k);
    w.skip.alloc(k);
    w.again.alloc(k);
    w.chk.alloc(2 * size_t(k));
    w.host.ensure(2 * 32);
    w.iters_host.ensure(32);
    // 0 until a transposed X has a nonzero; the FP32 inner solves, whose X is
    // cleared before every start, never set it
    w.xnz.alloc(1);
    NLHIP_CHECK(hipMemsetAsync(w.xnz.get(), 0, sizeof(int32_t), S.stream.get()));
    MultiArgs<T>& a = w.args;
    a = MultiArgs<T>{};
    a.m = S.n;
    a.k = k;
    a.G = pow2_ceil(k);
    a.P = P;
    const uint32_t rows_per_pass = 256u / a.G;
    uint64_t rpb = (uint64_t(S.n) + P - 1) / P;
    rpb = (rpb + rows_per_pass - 1) / rows_per_pass * rows_per_pass;
    a.rows_per_block = uint32_t(std::max<uint64_t>(rpb, rows_per_pass));
    a.one_chunk = spmm_one_chunk(S.n, S.rowptr.get(), a.G, S.stream.get()) ? 1u : 0u;
    a.rp = S.rowptr.get();
    a.ci = S.colind.get();
    a.cols = S.cols();
    a.Pdq = P;
    if (k <= kMultiStreamMaxK) {
        // stream SpMM: the 64-row tiles dealt to as many workgroups as the
        // device holds at once (no second, partly filled round), a multiple
        // of 4 tiles (one quarter per wavefront) to each, at most P workgroups
        const uint64_t ntiles = (uint64_t(S.n) + 63) / 64;
        const uint64_t resident = uint64_t(std::max(1, device().cu_count)) *
                                  multi_stream_per_cu<T>(k, a.cols.ci16 != nullptr);
        const uint64_t want = std::min<uint64_t>({P, resident, (ntiles + 3) / 4});
        uint64_t sp = (ntiles + want - 1) / want;
        sp = (sp + 3) / 4 * 4;
        a.sp = uint32_t(sp);
        a.Pdq = uint32_t((ntiles + sp - 1) / sp);
    }
    a.x = w.x.get();
    a.b = w.b.get();
    a.r = w.r.get();
    a.h = w.h.get();
    a.d = w.d.get();
    a.q = w.q.get();
    a.part_dq = w.part.get();
    a.part_rh = w.part.get() + size_t(P) * k;
    a.part_rr = w.part.get() + 2 * size_t(P) * k;
    a.part_bb = w.part.get() + 3 * size_t(P) * k;
    a.rho = w.col.get();
    a.alpha = w.col.get() + k;
    a.beta = w.col.get() + 2 * size_t(k);
    a.err = w.col.get() + 3 * size_t(k);
    a.thresh = w.col.get() + 4 * size_t(k);
    a.bb = w.col.get() + 5 * size_t(k);
    a.iters = w.iters.get();
    a.done = w.done.get();
    a.skip = nullptr;
    a.x_nonzero = w.xnz.get();
}

// Runs the started (or re-opened) recurrences until every column has
// stopped; `budget` bounds the iterations a column can still run.
template <typename T>
void multi_loop(System& S, MultiWork<T>& w, bool precond, int max_iter, uint32_t K,
                uint64_t budget) {
    hipStream_t s = S.stream.get();
    if (w.graph_precond != precond || w.graph_max_iter != max_iter ||
        std::memcmp(&w.graph_args, &w.args, sizeof(MultiArgs<T>)) != 0)
        w.graph.reset();
    w.graph.ensure(K, s, [&](uint32_t k) {
        for (uint32_t i = 0; i < k; ++i)
            NLHIP_CHECK(multi_iteration<T>(w.args, precond, max_iter, s));
    });
    w.graph_precond = precond;
    w.graph_max_iter = max_iter;
    w.graph_args = w.args;
    Event ev[2];
    int32_t* host = w.host.get();
    const uint32_t k = w.k;
    // every column stopped, by the done flags copied to polling slot i
    auto stopped = [&](int i) {
        for (uint32_t j = 0; j < k; ++j)
            if (!host[32 * i + j]) return false;
        return true;
    };
    const uint64_t limit = budget + 3ull * w.graph.cap() + 8;
    uint64_t launched = 0, covered = 0;
    for (;;) {
        w.graph.launch(launched, s);
        covered += w.graph.iters(launched);
        const int slot = int(launched & 1);
        NLHIP_CHECK(hipMemcpyAsync(&host[32 * slot], w.done.get(), k * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, s));
        ev[slot].record(s);
        ++launched;
        if (launched >= 2) {
            const int prev = int((launched - 2) & 1);
            ev[prev].sync();
            if (stopped(prev)) break;
        }
        if (covered > limit) {
            S.stream.sync();
            if (stopped(slot)) break;
            throw Error(NLHIP_ERR_HIP, "multi right-hand-side CG did not reach its stopping test");
        }
    }
    S.stream.sync();
}

template <typename T>
void run_multi(System& S, MultiWork<T>& w, bool precond, T eps, int max_iter, uint32_t K) {
    NLHIP_CHECK(multi_start<T>(w.args, precond, eps, max_iter, S.stream.get()));
    multi_loop<T>(S, w, precond, max_iter, K, uint64_t(std::max(max_iter, 0)));
}

// Iterations per captured graph: about 0.25 ms of device work, 3 to 16.
// The loop checks a graph's stop flag after queuing the next graph, so a
// solve that stops inside graph i still runs graph i + 1 as launches that
// exit at once (about 10 us per iteration), while each graph costs about
// 9 us more than its launches; multi-RHS solves often stop after a few
// iterations (mesh smoothing: 5 to 60), where short graphs waste least.
uint32_t multi_check_every(const System& S, uint32_t k) {
    const double bytes = 12.0 * double(S.nnz) + 108.0 * double(S.n) * k;
    const double iter_s = std::max(bytes / 4.0e12, 2.0e-5);
    return uint32_t(std::clamp(2.5e-4 / iter_s, 3.0, 16.0));
}

// FP64 |b_j - A x_j|^2 and |b_j|^2 for every column into rr, bb.
void multi_norms(System& S, uint32_t k, std::vector<double>& rr, std::vector<double>& bb) {
    MultiArgs<double> a = S.m64.args;
    a.val = S.val.get();
    NLHIP_CHECK(multi_residual64(a, S.stream.get()));
    rr.resize(k);
    bb.resize(k);
    NLHIP_CHECK(hipMemcpyAsync(rr.data(), a.err, k * 8, hipMemcpyDeviceToHost, S.stream.get()));
    NLHIP_CHECK(hipMemcpyAsync(bb.data(), a.bb, k * 8, hipMemcpyDeviceToHost, S.stream.get()));
    S.stream.sync();
}

// |b_j - A x_j|^2 and |b_j|^2 without touching what the stopped FP64
// recurrences need to continue (R, the r.h / r.r partials, the column
// scalars): the residual goes to Q and the d.q partials, the norms to chk.
void confirm_norms(System& S, uint32_t k, std::vector<double>& rr, std::vector<double>& bb) {
    MultiWork<double>& w = S.m64;
    MultiArgs<double> a = w.args;
    a.val = S.val.get();
    a.r = w.q.get();
    a.part_rr = a.part_dq;
    a.err = w.chk.get();
    a.bb = w.chk.get() + k;
    NLHIP_CHECK(multi_residual64(a, S.stream.get()));
    rr.resize(k);
    bb.resize(k);
    NLHIP_CHECK(hipMemcpyAsync(rr.data(), a.err, k * 8, hipMemcpyDeviceToHost, S.stream.get()));
    NLHIP_CHECK(hipMemcpyAsync(bb.data(), a.bb, k * 8, hipMemcpyDeviceToHost, S.stream.get()));
    S.stream.sync();
}

constexpr int kMaxConfirmRounds = 8;

}  // namespace

void solve_multi(System& S, uint32_t k, const double* B, double* X, const nlhip_solve_params& p,
                 nlhip_solve_stats* per, nlhip_solve_stats* total) {
    if (k == 0 || k > 32) throw Error(NLHIP_ERR_UNSUPPORTED, "1 to 32 right-hand sides");
    if (p.precond != NL_PRECOND_JACOBI && p.precond != NL_PRECOND_NONE)
        throw Error(NLHIP_ERR_UNSUPPORTED, "only Jacobi or no preconditioner on the GPU");
    std::lock_guard<std::mutex> lock(S.mu);
    if (take_injected_failure()) throw Error(NLHIP_ERR_INJECTED, "injected failure");
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = S.stream.get();
    multi_alloc(S, S.m64, k);
    MultiWork<double>& w = S.m64;
    const size_t nk = size_t(S.n) * k;
    // stage column-major B and X through q, then interleave
    copy_in(w.q.get(), B, nk * 8, s);
    NLHIP_CHECK(multi_layout<double>(S.n, k, w.q.get(), w.b.get(), true, s));
    copy_in(w.q.get(), X, nk * 8, s);
    NLHIP_CHECK(hipMemsetAsync(w.xnz.get(), 0, sizeof(int32_t), s));
    NLHIP_CHECK(multi_layout<double>(S.n, k, w.q.get(), w.x.get(), true, s, w.xnz.get()));
    const bool precond = p.precond == NL_PRECOND_JACOBI;
    const uint32_t K = p.check_every ? p.check_every : multi_check_every(S, k);
    Event e0, e1;
    e0.record(s);
    std::vector<int32_t> its(k, 0);
    std::vector<uint32_t> outer(k, 0);
    std::vector<double> cgrel(k, 0.0);
    std::vector<uint8_t> conv(k, 0);
    std::vector<uint32_t> first(k, ~0u), extra(k, 0);
    if (p.precision == NLHIP_PREC_FP64) {
        w.args.val = S.val.get();
        w.args.dinv = precond ? S.dinv.get() : nullptr;
        w.args.skip = nullptr;
        run_multi<double>(S, w, precond, p.eps, int(p.max_iter), K);
        std::vector<double> err(k), thr(k), bb(k);
        auto read = [&] {
            NLHIP_CHECK(hipMemcpyAsync(its.data(), w.iters.get(), k * 4, hipMemcpyDeviceToHost, s));
            NLHIP_CHECK(hipMemcpyAsync(err.data(), w.args.err, k * 8, hipMemcpyDeviceToHost, s));
            NLHIP_CHECK(hipMemcpyAsync(thr.data(), w.args.thresh, k * 8, hipMemcpyDeviceToHost, s));
            NLHIP_CHECK(hipMemcpyAsync(bb.data(), w.args.bb, k * 8, hipMemcpyDeviceToHost, s));
            S.stream.sync();
        };
        read();
        for (uint32_t j = 0; j < k; ++j) {
            if (!std::isfinite(err[j]) || !std::isfinite(bb[j]))
                throw Error(NLHIP_ERR_NONFINITE, "right-hand side or iterate is NaN or Inf");
            if (!(err[j] > thr[j])) first[j] = uint32_t(its[j]);
        }
        // confirmation of the true residual, column by column (see
        // nlhip_solve_params::confirm): a column whose recursive residual met
        // eps but whose |b - A x| did not continues its recurrence with a
        // lower threshold; the others stay stopped
        std::vector<double> rr, tb;
        if (p.confirm) {
            std::vector<double> last(k, -1.0);
            std::vector<int32_t> again(k);
            for (int round = 0;; ++round) {
                confirm_norms(S, k, rr, tb);
                bool any = false;
                for (uint32_t j = 0; j < k; ++j) {
                    again[j] = 0;
                    if (err[j] > thr[j]) continue;
                    if (first[j] == ~0u) first[j] = uint32_t(its[j]);
                    const double target = (p.eps * p.eps) * tb[j];
                    if (rr[j] > target && uint32_t(its[j]) < p.max_iter && err[j] > 0.0 &&
                        (last[j] < 0.0 || rr[j] <= 0.99 * last[j]) && round < kMaxConfirmRounds) {
                        again[j] = 1;
                        thr[j] = err[j] * (target / rr[j]) * 0.5;
                        last[j] = rr[j];
                        any = true;
                    }
                }
                if (!any) break;
                NLHIP_CHECK(hipMemcpyAsync(w.args.thresh, thr.data(), k * 8, hipMemcpyHostToDevice, s));
                NLHIP_CHECK(hipMemcpyAsync(w.again.get(), again.data(), k * 4,
                                           hipMemcpyHostToDevice, s));
                NLHIP_CHECK(multi_reopen(w.args, w.again.get(), s));
                multi_loop<double>(S, w, precond, int(p.max_iter), K, p.max_iter);
                read();
            }
        }
        for (uint32_t j = 0; j < k; ++j) {
            cgrel[j] = bb[j] > 0 ? std::sqrt(err[j] / bb[j]) : std::sqrt(err[j]);
            conv[j] = p.confirm ? (first[j] != ~0u && !(rr[j] > (p.eps * p.eps) * tb[j]) ? 1 : 0)
                                : (!(err[j] > thr[j]) ? 1 : 0);
            extra[j] = first[j] == ~0u ? 0u : uint32_t(its[j]) - first[j];
        }
    } else if (p.precision == NLHIP_PREC_MIXED) {
        S.ensure_fp32();
        multi_alloc(S, S.m32, k);
        MultiWork<float>& w32 = S.m32;
        w32.args.val = S.val32.get();
        w32.args.dinv = precond ? S.dinv32.get() : nullptr;
        w32.args.skip = w32.skip.get();
        S.mscale.ensure(2 * size_t(k));
        const uint32_t inner_cap = p.inner_max_iter ? p.inner_max_iter : p.max_iter;
        std::vector<int32_t> skip(k), refined(k, 0), save(k), back(k), stalled(k, 0);
        std::vector<uint32_t> no_gain(k, 0);
        std::vector<double> sc(2 * size_t(k)), rr_best(k, 0.0);
        // each column's iterate with the lowest FP64 residual so far (w.d is
        // free while refining); the refinement of a column ends back at it
        // (kRefineNoGain / kRefineDiverged, as for one system)
        double* const x_best = w.d.get();
        int32_t* const inner_host = w32.iters_host.get();
        bool first = true, finish = false, pending = false;
        for (;;) {
            std::vector<double> rr, bb;
            multi_norms(S, k, rr, bb);  // r64 = B - A X in w.r (waits for the stream)
            if (pending) {
                // the previous step's inner iterations, copied behind it
                bool progress = false;
                for (uint32_t j = 0; j < k; ++j) {
                    if (skip[j]) continue;
                    its[j] += inner_host[j];
                    outer[j] += 1;
                    progress |= inner_host[j] > 0;
                }
                finish = !progress;  // this pass settles every column and ends
                pending = false;
            }
            if (first)
                for (uint32_t j = 0; j < k; ++j)
                    if (!std::isfinite(rr[j]) || !std::isfinite(bb[j]))
                        throw Error(NLHIP_ERR_NONFINITE,
                                    "right-hand side or iterate is NaN or Inf");
            uint32_t active = 0;
            bool any_save = false, any_back = false;
            for (uint32_t j = 0; j < k; ++j) {
                save[j] = back[j] = 0;
                if (first || (refined[j] && rr[j] < rr_best[j])) {
                    save[j] = 1;
                    rr_best[j] = rr[j];
                    no_gain[j] = 0;
                } else if (refined[j] && (!std::isfinite(rr[j]) ||
                                          rr[j] > kRefineDiverged * rr_best[j] ||
                                          ++no_gain[j] >= kRefineNoGain)) {
                    stalled[j] = 1;
                }
                const bool budget = uint32_t(its[j]) >= p.max_iter || outer[j] >= p.max_outer;
                const bool done = !(rr[j] > (p.eps * p.eps) * bb[j]) || budget || stalled[j] ||
                                  finish;
                if (done && !(rr[j] <= rr_best[j])) {  // back to the best iterate
                    back[j] = 1;
                    rr[j] = rr_best[j];
                }
                any_save |= save[j] != 0;
                any_back |= back[j] != 0;
                conv[j] = !(rr[j] > (p.eps * p.eps) * bb[j]) ? 1 : 0;
                cgrel[j] = bb[j] > 0 ? std::sqrt(rr[j] / bb[j]) : std::sqrt(rr[j]);
                skip[j] = done ? 1 : 0;
                refined[j] = done ? 0 : 1;
                const double rn = std::sqrt(rr[j]);
                sc[j] = skip[j] ? 0.0 : 1.0 / rn;
                sc[k + j] = skip[j] ? 0.0 : rn;
                active += skip[j] ? 0 : 1;
            }
            first = false;
            if (any_save) {
                NLHIP_CHECK(hipMemcpyAsync(w.again.get(), save.data(), k * 4,
                                           hipMemcpyHostToDevice, s));
                NLHIP_CHECK(multi_copy_cols(S.n, k, x_best, w.x.get(), w.again.get(), s));
            }
            if (any_back) {
                NLHIP_CHECK(hipMemcpyAsync(w.again.get(), back.data(), k * 4,
                                           hipMemcpyHostToDevice, s));
                NLHIP_CHECK(multi_copy_cols(S.n, k, w.x.get(), x_best, w.again.get(), s));
            }
            if (active == 0) break;
            NLHIP_CHECK(hipMemcpyAsync(w32.skip.get(), skip.data(), k * 4, hipMemcpyHostToDevice, s));
            NLHIP_CHECK(hipMemcpyAsync(S.mscale.get(), sc.data(), sc.size() * 8,
                                       hipMemcpyHostToDevice, s));
            NLHIP_CHECK(multi_scale_cast(w.args, w32.args, S.mscale.get(), s));
            run_multi<float>(S, w32, precond, float(p.inner_eps), int(inner_cap), K);
            NLHIP_CHECK(hipMemcpyAsync(inner_host, w32.iters.get(), k * 4, hipMemcpyDeviceToHost, s));
            NLHIP_CHECK(multi_ir_update(w.args, w32.args, S.mscale.get() + k, s));
            pending = true;  // counted after the next residual, behind which it lands
        }
    } else {
        throw Error(NLHIP_ERR_INVALID, "unknown precision");
    }
    e1.record(s);
    e1.sync();
    const double dev_s = 1e-3 * elapsed_ms(e0, e1);
    std::vector<double> rr, bb;
    multi_norms(S, k, rr, bb);
    NLHIP_CHECK(multi_layout<double>(S.n, k, w.x.get(), w.q.get(), false, s));
    copy_out(X, w.q.get(), nk * 8, s);
    S.stream.sync();
    const double wall_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    uint32_t it_max = 0, nconv = 0;
    double worst = 0.0;
    for (uint32_t j = 0; j < k; ++j) {
        const double rel = bb[j] > 0 ? std::sqrt(rr[j] / bb[j]) : std::sqrt(rr[j]);
        it_max = std::max(it_max, uint32_t(its[j]));
        nconv += conv[j];
        worst = std::max(worst, rel);
        if (per) {
            nlhip_solve_stats& st = per[j];
            std::memset(&st, 0, sizeof(st));
            st.iterations = uint32_t(its[j]);
            st.outer_iterations = outer[j];
            st.converged = conv[j];
            st.rel_residual = rel;
            st.cg_rel_residual = cgrel[j];
            st.device_seconds = dev_s;
            st.wall_seconds = wall_s;
            st.extra_iterations = extra[j];
        }
    }
    if (total) {
        std::memset(total, 0, sizeof(*total));
        total->iterations = it_max;
        total->extra_iterations = *std::max_element(extra.begin(), extra.end());
        total->converged = nconv == k ? 1 : 0;
        total->rel_residual = worst;
        total->device_seconds = dev_s;
        total->wall_seconds = wall_s;
    }
}

}  // namespace nlhip

using namespace nlhip;

extern "C" {

NLAPI int NLAPIENTRY nlhipSolveMulti(nlhip_system s, uint32_t k, const double* B, double* X,
                                     const nlhip_solve_params* params,
                                     nlhip_solve_stats* per_rhs, nlhip_solve_stats* total) {
    return guarded("nlhipSolveMulti", [&] {
        if (!s) throw Error(NLHIP_ERR_INVALID, "null system handle");
        if (!B || !X) throw Error(NLHIP_ERR_INVALID, "null B or X");
        nlhip_solve_params p;
        nlhipSolveParamsDefault(&p);
        if (params) p = *params;
        solve_multi(*reinterpret_cast<System*>(s), k, B, X, p, per_rhs, total);
    });
}

}  // extern "C"
